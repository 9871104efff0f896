// Two-level multigrid preconditioner (the role of QUDA's multigrid behind MULTIGRIDSolverParams,
// quda_multigrid_params.h:12-113) for the even-odd operator: kernels (mg.hip) and host state (mg.cpp).
//
// The coarse space aggregates the null vectors of the unpreconditioned operator M over blocks of the
// full lattice, split by chirality (gamma_5 is diagonal in the DeGrand-Rossi basis: spins 0,1 and 2,3):
// coarse degrees of freedom per block D = 2 nvec.  The coarse operator is the Galerkin product
// A_c = P^dag M P with a nearest-neighbour (9-point) block stencil.  The V-cycle preconditions the
// even-odd Schur operator A = A_oo - D_oe A_ee^-1 D_eo / 4 through the identity A^-1 = (M^-1)_oo: MR
// smoothing on A, restriction of the odd residual, a GCR solve of the coarse system, prolongation of
// the odd part, MR smoothing.
#pragma once
#include "op.h"

namespace cmx {

constexpr int MG_MAXVEC = 16;           // null vectors
constexpr int MG_MAXD = 2 * MG_MAXVEC;  // coarse degrees of freedom per block
constexpr int MG_NST = 9;               // coarse stencil: the block, then +x, -x, +y, -y, +z, -z, +t, -t
constexpr int MG_MAXKRY = 32;           // coarse GCR directions

struct MGGeom {
  Geom g;          // fine lattice
  int B[4], Lc[4]; // block extents, coarse extents
  int Nb, nvec, D, bvol;
};

struct MGVecs {
  const float4* P[MG_MAXVEC];  // null vectors: full lattice (even then odd), FP32, 6 chunks of 2 Vh
  // FP16 copies in the library's FP16 fermion format (3 chunks of 2 Vh, one norm per site), read by the
  // restriction and prolongation of an FP16 V-cycle (half the bytes); null when absent
  const void* H[MG_MAXVEC];
  const float* Hn[MG_MAXVEC];
};

// block orthonormalisation of the null vectors per (block, chirality), FP64 sums
void launch_mg_orth(const MGGeom& m, float4* const* P, hipStream_t s);
// out[b * D + c * nvec + k] = sum over the block's sites of P_k^c(x)^dag v(x): full = true: v a full-lattice
// vector (view of the whole vector, stride 2 Vh); false: v an odd-checkerboard vector (its own stride)
void launch_mg_restrict(const MGGeom& m, const MGVecs& P, int prec, VRef v, bool full, float2* out, hipStream_t s);
// out (+)= P e (precision prec): full = true: out a full-lattice vector; false: odd checkerboard
void launch_mg_prolong(const MGGeom& m, const MGVecs& P, int prec, const float2* e, VRef out, bool full, bool add,
                       hipStream_t s);
// probe vector: full-lattice FP64, P_k^c on the blocks of colour col, zero elsewhere
void launch_mg_probe(const MGGeom& m, const MGVecs& P, const int* colour, int col, int k, int c, VRef out,
                     hipStream_t s);
// FP32 odd-checkerboard Z2 noise (+-1/sqrt 2 per real) from (seed, site)
void launch_mg_noise(VRef out, int n, unsigned long long seed, hipStream_t s);
// the coarse operator's column dof from v = M probe under a colouring of period 2 (1) per direction:
// each block's entries whose neighbour has colour col, from the block or its two faces in that direction
void launch_mg_face_restrict(const MGGeom& m, const MGVecs& P, const int* colour, const int* nbr, int col, int dof,
                             VRef v, float2* Ac, hipStream_t s);
// The coarse operator Ac holds per block b and stencil entry d a D x D matrix, column-major
// (element (r, c) at ((b * 9 + d) * D + c) * D + r), so the D lanes of a block read a column together.
// mg_scatter: element (r, dof) of entry d of block b = y[b][r], for the first stencil entry d whose
// block has colour col
void launch_mg_scatter(const MGGeom& m, const int* nbr, const int* colour, int col, int dof, const float2* y,
                       float2* Ac, hipStream_t s);
// y = A_c x; skip (nullable): a device flag, the launch does nothing while it is set
void launch_mg_coarse_apply(const MGGeom& m, const int* nbr, const float2* Ac, const float2* x, float2* y,
                            const int* skip, hipStream_t s);
// the same with the entries' FP16 copy Ah (complex half pairs, same layout), for an FP16 V-cycle
void launch_mg_coarse_apply_h(const MGGeom& m, const int* nbr, const void* Ah, const float2* x, float2* y,
                              const int* skip, hipStream_t s);
// h[i] = a[i] rounded to FP16 (n complex entries)
void launch_mg_to_half(long n, const float2* a, void* h, hipStream_t s);
// coarse vectors (n complex): partials[j][block] = sum over the block's elements of conj(v_j) w, j < m
int mg_cdot_blocks(int n);
void launch_mg_cdot(const float2* const* v, int m, const float2* w, int n, double2* partials, const int* skip,
                    hipStream_t s);
// The coarse GCR's scalars live on the device (S: doubles, F: flags), so a V-cycle needs no host
// synchronisation.  mg_gcr_fin sums the partials of m dot products and, by stage: INIT (|r|^2 of the
// right-hand side: target and flags), STEP (the Gram-Schmidt coefficients of direction j, its
// normalisation, alpha, the residual norm recursion, the stopping flag), RESTART (|r|^2 recomputed).
enum { MGS_B2 = 0, MGS_R2 = 1, MGS_STOP = 2, MGS_INV = 3, MGS_AR = 4, MGS_AI = 5, MGS_C = 8,
       MGS_N = MGS_C + 2 * MG_MAXKRY };
enum { MGF_SKIP = 0, MGF_UPD = 1, MGF_N = 4 };
enum { MG_FIN_INIT = 0, MG_FIN_STEP = 1, MG_FIN_RESTART = 2 };
void launch_mg_gcr_fin(const double2* partials, int nblk, int m, int j, int stage, double tol2, double* S, int* F,
                       hipStream_t s);
// GCR update with the device scalars: when F[UPD]: q_j = (q_j - sum_i c_i q_i) inv; p_j likewise;
// x += alpha p_j; r -= alpha q_j
struct MGUpdate {
  float2* q[MG_MAXKRY];
  float2* p[MG_MAXKRY];
  int j, n;
  const double* S;
  const int* F;
  float2 *x, *r;
};
void launch_mg_update(const MGUpdate& u, hipStream_t s);

struct MGState;
// Builds the null space, the block-orthonormal prolongator and the coarse operator on op
void mg_setup(Op& op, const cmx_mg_params& p, cmx_mg_result& res);
// the V-cycle's parameters (smoothing steps, omega, coarse tolerance, iterations and directions) changed
// on a set-up multigrid; blocking and nvec must equal the set-up's
void mg_update(Op& op, const cmx_mg_params& p);
// z = K v: one V-cycle on the odd-checkerboard system at precision prec (v, z in that precision)
void mg_precondition(Op& op, int prec, const Vec& v, Vec& z);
class BatchPlan;
// z[i] = K v[i] for solver contexts ops[i] that share one multigrid, stream and launch shape: the
// restriction, coarse solve and prolongation one source after another, the MR smoothing steps of all
// sources in one batched grid per kernel (plan); each z[i] equals mg_precondition's bit for bit
void mg_precondition_batch(const std::vector<Op*>& ops, int prec, const std::vector<const Vec*>& v,
                           const std::vector<Vec*>& z, BatchPlan& plan);
// test and inspection access: null vectors (host order, FP64 copies), the coarse operator, and the
// restriction / prolongation / coarse operator on host arrays
void mg_get_nullvecs(Op& op, double* out);
void mg_get_coarse_op(Op& op, double* out);
void mg_restrict_host(Op& op, const double* full_in, double* coarse_out);
void mg_prolong_host(Op& op, const double* coarse_in, double* full_out);
void mg_coarse_apply_host(Op& op, const double* in, double* out);
void mg_info(Op& op, int* nb, int* d);

}  // namespace cmx
