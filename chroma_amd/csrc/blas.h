// Fused BLAS-1 over fermion views, with FP64 reductions and device-resident solver scalars.
//
// Solver scalars (alpha, beta, |r|^2, ...) never leave the device inside an iteration: reductions
// write one partial per work-group, a one-work-group "finalize" kernel sums the partials in a fixed
// order (deterministic) and runs a short scalar program (e.g. a = c/d, the convergence test), and
// the vector kernels read their coefficients from the same device array.  The host only polls the
// convergence flag every few iterations, so a whole iteration can be replayed from a HIP graph.
#pragma once
#include "cmx_internal.h"

namespace cmx {

// Scalar slots (doubles).  Complex scalars occupy two consecutive slots (re, im).
enum Slot {
  S_C = 0, S_CP, S_D, S_A, S_B, S_RSDSQ, S_NEGA,
  // reliable CG (reliable_cg.cc:60-170)
  S_RSQ = 8, S_RNORM, S_R0NORM, S_MAXRX, S_MAXRR, S_DELTA, S_BETA, S_CHINORM,
  // BiCGStab (invbicgstab.cc), complex pairs
  S_RHO = 20, S_RHOP = 22, S_ALPHA = 24, S_OMEGA = 26, S_BETAC = 28, S_CTMP = 30, S_TNORM = 32,
  S_NALPHA = 34, S_NOMEGA = 36, S_TR = 38,
  // Krylov-basis solvers (FGMRES-DR): coefficient columns, norms and inverse norms
  S_KRY = 64,
  // MR preconditioner (invmr.cc:120-160): c = <Mr,r> (2), d, a (2), -a (2), |v|^2, cp, omega, RsdMR
  S_MR = 480,
  S_NSLOTS = 512
};
// Flags (ints)
enum Flag {
  F_DONE = 0,   // iteration kernels return immediately once set
  F_CONV,       // converged (or stopped); copied into F_DONE by the next finalize
  F_NCOUNT,     // iteration count at convergence
  F_K,          // iterations completed
  F_MAXIT,      // MaxIter
  F_UPDR,       // reliable update of r this iteration
  F_UPDX,       // reliable update of x (group-wise) this iteration
  F_NUPD,       // number of reliable updates
  F_ERR,        // 0 ok, CMX_ERR_NOCONV, CMX_ERR_BREAKDOWN
  F_MRDONE,     // MR preconditioner converged: its remaining iterations return immediately
  F_NFLAGS = 16
};

// Finalize programs
enum Prog {
  PG_STORE = 0,      // S[slot+r] = red[r]
  PG_CG_INIT,        // after r = chi - A psi: cp, c = cp, convergence at k = 0
  PG_CG_D,           // after |M p|^2: F_DONE = F_CONV; d; a = c/d; -a
  PG_CG_R,           // after r -= a A p, |r|^2: k++, convergence, b = cp/c, c = cp
  PG_RCG_INIT,       // reliable CG start: r_sq, rNorm = r0Norm = maxrx = maxrr
  PG_RCG_D,          // F_DONE = F_CONV; d; a = c/d
  PG_RCG_R,          // r_sq, rNorm, maxrx/maxrr, updateX/updateR decision, convergence if no update
  PG_RCG_UPD,        // after the FP64 reliable update: r_sq, rNorm, maxrr, (r0Norm, maxrx), convergence
  PG_BICG_RHO,       // rho = <r0,r>; breakdown test; beta = (rho/rho_prev)(alpha/omega)
  PG_BICG_ALPHA,     // ctmp = <r0,v>; alpha = rho/ctmp; rho_prev = rho
  PG_BICG_OMEGA,     // |t|^2, <t,r>: omega = <t,r>/|t|^2
  PG_BICG_R,         // |r|^2: k++, convergence
  // reliable BiCGStab (reliable_bicgstab.cc:14-290)
  PG_RBICG_INIT,     // after b = chi - A psi: r_sq, norms, rho = r_sq, rho_prev = alpha = omega = 1, beta = 0
  PG_RBICG_ALPHA,    // F_DONE = F_CONV; ctmp = <r0,v>; alpha = rho/ctmp; rho_prev = rho
  PG_RBICG_R,        // |r|^2, rho = <r0,r>: updateX / updateR; without an update: k++, convergence, beta
  PG_RBICG_UPD,      // after the FP64 update of r: r_sq, norms; k++, convergence, beta
  // MR (invmr.cc:100-160), run with the F_MRDONE skip flag
  PG_MR_INIT,        // cp = |r|^2; rsd_sq = RsdMR^2 |v|^2; F_MRDONE = cp <= rsd_sq
  PG_MR_A,           // d = |Mr|^2, c = <Mr, r>: a = omega c / d
  PG_MR_R,           // cp = |r|^2; F_MRDONE = cp <= rsd_sq
};

// A fermion view for BLAS kernels
struct VRef {
  void* v;
  float* nrm;
  int stride;
};

struct BlasLaunch {
  int n;            // sites
  VRef v[4];
  double* S;        // scalars
  int* F;           // flags
  int ia, ib;       // scalar slots used as coefficients
  double* partials; // per work-group reductions (nred each)
  int skip_flag;    // return when F[skip_flag] != 0 (-1: never)
  int run_flag;     // run only when F[run_flag] != 0 (-1: always)
};

// Functor ids
enum BlasOp {
  B_COPY = 0,       // v0 = v1                                     (precision conversion allowed)
  B_ZERO,           // v0 = 0
  B_NORM2,          // red0 = |v0|^2
  B_XMY_NORM,       // v0 = v1 - v2 ; red0 = |v0|^2
  B_XMY_NORM_COPY,  // v0 = v1 - v2 ; v3 = v0 ; red0 = |v0|^2
  B_AXPY_NORM,      // v0 += S[ia] v1 ; red0 = |v0|^2
  B_AXPY,           // v0 += S[ia] v1
  B_XPAY,           // v0 = v1 + S[ia] v0
  B_CG_UPDATE,      // v0 += S[ia] v1 ; v1 = v2 + S[ib] v1             (psi, p, r)
  B_AXPY2_NORM,     // v0 += S[ia] v2 ; v1 -= S[ia] v3 ; red0 = |v1|^2 (x, r, p, Ap)
  B_ADD_ZERO_COPY,  // v0 += v1 ; v1 = 0 ; v2 = v3                   (psi += x_dble; x = 0; b = r_dble)
  B_ADD,            // v0 += v1
  B_CDOT,           // red0,1 = <v0, v1>
  B_BICG_P,         // v0 = v1 + S[ib](v0 - S[ia] v2)                (p = r + beta (p - omega v))
  B_CAXPY,          // v0 += S[ia] v1 (complex)
  B_TNORM_TDOT,     // red0 = |v0|^2, red1,2 = <v0, v1>
  B_BICG_UPDATE,    // v0 += S[ia] v1 + S[ib] v2 ; v1 -= S[ia] v3 ; red0 = |v1|^2  (psi, r, p, t)
  B_CAXPY2,         // v0 += S[ia] v1 ; v2 += S[ib] v3 (complex)    (r -= alpha v ; x += alpha p)
  B_RBICG_UPDATE,   // v0 += S[ia] v1 ; v1 -= S[ia] v2 ; red0 = |v1|^2, red1,2 = <v3, v1>
                    //                                     (x += omega s ; r = s - omega t ; |r|^2, <r0|r>)
  B_SCALE,          // v0 = S[ia] v1 (real)
  B_AXPY_CNORM,     // v0 += S[ia] v1 (complex) ; red0 = |v0|^2
  B_AXPBY,          // v0 = S[ia] v1 + S[ib] v0 (real)
  B_MR_UPDATE,      // v0 += S[ia] v1 ; v1 += S[ib] v2 (complex) ; red0 = |v1|^2  (psi += a r ; r -= a M r)
  B_NOPS
};

int blas_nred(int op);
int blas_blocks(int n);

// Multi-vector kernels of the Krylov-basis solvers (krylov.hip).  A basis is up to KRY_MAXV views of
// one precision; coefficients are complex pairs in the scalar array S.
constexpr int KRY_MAXV = 48;
constexpr int KRY_BLOCKS = 1024;  // work-groups of the reductions (grid-stride), 256 threads each
struct MultiLaunch {
  int n;                 // sites
  int m;                 // basis vectors in use
  VRef w;                // the single vector
  VRef v[KRY_MAXV];      // the basis
  double* S;
  int coef;              // S slot of the first complex coefficient
  double sgn;            // coefficient multiplier (+1 / -1)
  int overwrite;         // maxpy: w = sum (instead of w += sum)
  double* partials;      // KRY_BLOCKS x nred
};
// partials[b][2i, 2i+1] = <v_i, w> over work-group b's sites
void launch_mdot(const MultiLaunch& L, int pv, int pw, hipStream_t s);
// w (+)= sgn * sum_i S[coef+2i] v_i ; optionally partials[b][0] = |w|^2
void launch_maxpy(const MultiLaunch& L, int pv, int pw, bool norm, hipStream_t s);
// out[r] = sum_b partials[b][r] for r < nred (one work-group per column, fixed order);
// mode 1 additionally writes out[nred] = 1 / sqrt(out[0])
void launch_colsum(const double* partials, int nb, int nred, double* out, int mode, hipStream_t s);
// bg != nullptr: the batched grid of bg (records of the same op and precisions)
void launch_blas(int op, const BlasLaunch& b, const int prec[4], hipStream_t s, const BatchGrid* bg = nullptr);

// Finalize: one work-group sums nb x nred partials in a fixed order into red[], stores them at
// S[slot..] (slot >= 0) and runs the scalar program prog
struct FinLaunch {
  const double* partials;
  int nb, nred;
  double* S;
  int* F;
  int prog, slot;
};
void launch_finalize(const FinLaunch& L, hipStream_t s);
// one finalize per problem: tab is a device array of nproblems records
void launch_finalize_batch(const FinLaunch* tab, int nproblems, hipStream_t s);

}  // namespace cmx
