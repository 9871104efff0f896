// Two-level multigrid: set-up (null space, prolongator, Galerkin coarse operator) and the V-cycle used
// as the right preconditioner of FGMRES-DR.  The parameters are the two-level form of Chroma's
// MULTIGRIDSolverParams (quda_multigrid_params.h:12-113); the reference hands them to QUDA, whose
// multigrid is not in the reference tree: the method here is the standard aggregation multigrid for
// Wilson-type operators (chirality-split block aggregates of relaxed near-null vectors, Galerkin coarse
// operator, MR smoothing, GCR on the coarse level).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>

#include "batch.h"
#include "fgmres.h"
#include "mg.h"
#include "setup.h"

namespace cmx {

struct MGState {
  cmx_mg_params par{};
  MGGeom m{};
  int period[4] = {1, 1, 1, 1};
  int ncol = 1;
  std::vector<std::unique_ptr<Vec>> P;  // null vectors, full lattice FP32
  std::vector<std::unique_ptr<Vec>> P16;  // their FP16 copies (restriction and prolongation at FP16)
  MGVecs pv{};
  DevBuf Ac, nbr, colour;
  DevBuf Ah;  // FP16 copy of Ac (the coarse solve of an FP16 V-cycle)
  // coarse GCR work space
  DevBuf xc, rc, bc, part, S, F;
  std::vector<DevBuf> pk, qk;
};

namespace {

MGState& state(Op& op) {
  CMX_REQUIRE(op.mg, CMX_ERR_STATE, "multigrid: cmx_mg_setup has not been called on this operator");
  return *static_cast<MGState*>(op.mg.get());
}

// the unpreconditioned operator M on full-lattice FP64 vectors (eoprec_linop.h:147-165):
// out_cb = A_cb in_cb - 1/2 D in_{1-cb}
void apply_full64(Op& op, const Vec& in, Vec& out) {
  const int Vh = (int)op.Vh;
  HopLaunch h0 = op.hop_base(P64);
  h0.isign = +1;
  for (int cb = 0; cb < 2; ++cb) {
    HopLaunch h = h0;
    h.cb = cb;
    const VRef vin = in.view((1 - cb) * Vh), vout = out.view(cb * Vh), vx = in.view(cb * Vh);
    h.in = vin.v, h.in_nrm = nullptr, h.in_stride = vin.stride;
    h.out = vout.v, h.out_nrm = nullptr, h.out_stride = vout.stride;
    h.x = vx.v, h.x_nrm = nullptr, h.x_stride = vx.stride;
    if (op.clover()) {
      h.epi = EPI_CLOV_XPAY;
      h.clov = cb == 0 ? op.clov64_even.p : op.clov_odd[P64].p;
    } else {
      h.epi = EPI_XPAY;
      h.clov = nullptr;
    }
    h.a = op.fact;
    h.b = -0.5;
    launch_hop(h, op.stream);
  }
  CMX_HIP(hipGetLastError());
}

// even part of a full FP64 vector from its odd part so that (M x)_e = 0: x_e = 1/2 A_ee^-1 D_eo x_o
void reconstruct_even64(Op& op, Vec& F) {
  const int Vh = (int)op.Vh;
  HopLaunch h = op.hop_base(P64);
  h.isign = +1;
  h.cb = 0;
  const VRef vin = F.view(Vh), vout = F.view(0);
  h.in = vin.v, h.in_nrm = nullptr, h.in_stride = vin.stride;
  h.out = vout.v, h.out_nrm = nullptr, h.out_stride = vout.stride;
  if (op.clover()) {
    h.epi = EPI_CLOV;
    h.clov = op.clovinv_even[P64].p;
    h.b = 0.5;
  } else {
    h.epi = EPI_NONE;
    h.clov = nullptr;
    h.b = 0.5 / op.fact;
  }
  h.a = 0.0;
  launch_hop(h, op.stream);
  CMX_HIP(hipGetLastError());
}

// copy n sites between views of any precision through the FP64 staging buffer
void convert(Op& op, int n, const VRef& src, int psrc, const VRef& dst, int pdst) {
  spinor_unpack(psrc, n, src.v, src.nrm, src.stride, op.staging.as<double>(), op.stream);
  spinor_pack(pdst, n, op.staging.as<double>(), dst.v, dst.nrm, dst.stride, op.stream);
  CMX_HIP(hipGetLastError());
}

double norm2_host(Op& op, const Vec& v) {
  const int nb = blas_blocks(v.n);
  op.blas(B_NORM2, {&v});
  op.finalize(nb, 1, PG_STORE, S_TR);
  CMX_HIP(hipMemcpyAsync(op.hS + S_TR, static_cast<double*>(op.state_S.p) + S_TR, sizeof(double),
                         hipMemcpyDeviceToHost, op.stream));
  op.sync();
  return op.hS[S_TR];
}

// x = A_c^-1 b by GCR(nk) with restarts, to |r| <= tol |b| or maxit steps (x, b: coarse vectors).  Every
// scalar stays on the device: the kernels after convergence return at once, and the host never waits.
void coarse_gcr(MGState& st, Op& op, const float2* b, float2* x, bool half) {
  const MGGeom& m = st.m;
  const int n = m.Nb * m.D;
  const int nk = st.par.coarse_nkrylov;
  const int nblk = mg_cdot_blocks(n);
  const double tol2 = st.par.coarse_tol * st.par.coarse_tol;
  float2* r = st.rc.as<float2>();
  double* S = st.S.as<double>();
  int* F = st.F.as<int>();
  const int* skip = F + MGF_SKIP;
  double2* part = st.part.as<double2>();
  CMX_HIP(hipMemsetAsync(x, 0, sizeof(float2) * n, op.stream));
  CMX_HIP(hipMemcpyAsync(r, b, sizeof(float2) * n, hipMemcpyDeviceToDevice, op.stream));
  const float2* rv[1] = {r};
  launch_mg_cdot(rv, 1, r, n, part, nullptr, op.stream);
  launch_mg_gcr_fin(part, nblk, 1, 0, MG_FIN_INIT, tol2, S, F, op.stream);
  int j = 0;
  for (int it = 0; it < st.par.coarse_maxit; ++it) {
    float2* pj = st.pk[j].as<float2>();
    float2* qj = st.qk[j].as<float2>();
    CMX_HIP(hipMemcpyAsync(pj, r, sizeof(float2) * n, hipMemcpyDeviceToDevice, op.stream));
    if (half)
      launch_mg_coarse_apply_h(m, st.nbr.as<int>(), st.Ah.p, pj, qj, skip, op.stream);
    else
      launch_mg_coarse_apply(m, st.nbr.as<int>(), st.Ac.as<float2>(), pj, qj, skip, op.stream);
    // <q_i, q_j> (i < j), <q_j, q_j>, conj(<q_j, r>) in one pass
    const float2* vs[MG_MAXKRY + 2];
    for (int i = 0; i <= j; ++i) vs[i] = st.qk[i].as<float2>();
    vs[j + 1] = r;
    launch_mg_cdot(vs, j + 2, qj, n, part, skip, op.stream);
    launch_mg_gcr_fin(part, nblk, j + 2, j, MG_FIN_STEP, tol2, S, F, op.stream);
    MGUpdate u;
    std::memset(&u, 0, sizeof(u));
    for (int i = 0; i <= j; ++i) u.q[i] = st.qk[i].as<float2>(), u.p[i] = st.pk[i].as<float2>();
    u.j = j;
    u.n = n;
    u.S = S;
    u.F = F;
    u.x = x;
    u.r = r;
    launch_mg_update(u, op.stream);
    if (++j == nk && it + 1 < st.par.coarse_maxit) {  // restart: |r|^2 recomputed from its vector
      j = 0;
      launch_mg_cdot(rv, 1, r, n, part, skip, op.stream);
      launch_mg_gcr_fin(part, nblk, 1, 0, MG_FIN_RESTART, tol2, S, F, op.stream);
    }
  }
  CMX_HIP(hipGetLastError());
}

}  // namespace

void mg_setup(Op& op, const cmx_mg_params& p, cmx_mg_result& res) {
  const auto t0 = std::chrono::steady_clock::now();
  CMX_REQUIRE(op.eo(), CMX_ERR_UNSUPPORTED, "multigrid: even-odd Wilson or clover operators");
  CMX_REQUIRE(p.nvec >= 1 && p.nvec <= MG_MAXVEC, CMX_ERR_ARG, "multigrid: NullVectors must be 1..16");
  CMX_REQUIRE(p.nu_pre >= 0 && p.nu_post >= 0 && p.coarse_maxit >= 0 && p.setup_maxit >= 0, CMX_ERR_ARG,
              "multigrid: negative iteration count");
  CMX_REQUIRE(p.coarse_nkrylov >= 1 && p.coarse_nkrylov <= MG_MAXKRY, CMX_ERR_ARG,
              "multigrid: PrecondGCRNKrylov must be 1..32");
  CMX_REQUIRE(p.prec == P64 || p.prec == P32 || p.prec == P16, CMX_ERR_ARG,
              "multigrid: Precision must be DOUBLE, SINGLE or HALF");
  CMX_REQUIRE(p.smoother == 0 || p.smoother == 1, CMX_ERR_ARG, "multigrid: smoother must be 0 (MR) or 1 (network)");
  std::unique_ptr<MGState> st(new MGState());
  st->par = p;
  MGGeom& m = st->m;
  m.g = op.g;
  m.Nb = 1;
  m.bvol = 1;
  for (int mu = 0; mu < 4; ++mu) {
    const int B = p.blocking[mu];
    CMX_REQUIRE(B >= 2 && B % 2 == 0 && op.g.L[mu] % B == 0, CMX_ERR_ARG,
                "multigrid: Blocking extents must be even and divide the lattice");
    m.B[mu] = B;
    m.Lc[mu] = op.g.L[mu] / B;
    m.Nb *= m.Lc[mu];
    m.bvol *= B;
  }
  m.nvec = p.nvec;
  m.D = 2 * p.nvec;
  const int Vh = (int)op.Vh;
  std::memset(&res, 0, sizeof(res));
  op.ensure_prec(P32);
  const int rp = p.prec;  // storage precision of the relaxation
  op.ensure_prec(rp);

  // null space: relaxed random vectors on the odd checkerboard, extended to the full lattice
  Vec x(rp, Vh), zero(rp, Vh), r(rp, Vh), mr(rp, Vh), F(P64, 2 * Vh);
  Vec x32(P32, rp == P32 ? 0 : Vh);
  op.blas(B_ZERO, {&zero});
  for (int k = 0; k < m.nvec; ++k) {
    // the noise start in FP32, stored at the relaxation's precision
    Vec& start = rp == P32 ? x : x32;
    launch_mg_noise(start.view(), Vh, p.seed * 0x100000001b3ull + 7919ull * (unsigned long long)(k + 1), op.stream);
    if (rp != P32) convert(op, Vh, x32.view(), P32, x.view(), rp);
    MRPrecond relax(op, rp, Vh, p.setup_maxit, 0.0, 1.0, r, mr);
    relax.apply(x, zero);
    op.apply(rp, mr.view(), x.view(), +1, nullptr);
    const double ax = norm2_host(op, mr), xx = norm2_host(op, x);
    CMX_REQUIRE(xx > 0, CMX_ERR_BREAKDOWN, "multigrid: a null vector relaxed to zero");
    res.nullvec_rel_residual_max = std::max(res.nullvec_rel_residual_max, std::sqrt(ax / xx));
    convert(op, Vh, x.view(), rp, F.view(Vh), P64);
    reconstruct_even64(op, F);
    st->P.emplace_back(new Vec(P32, 2 * Vh));
    convert(op, 2 * Vh, F.view(0), P64, st->P.back()->view(0), P32);
  }
  float4* pp[MG_MAXVEC] = {};
  for (int k = 0; k < m.nvec; ++k) {
    pp[k] = st->P[k]->buf.as<float4>();
    st->pv.P[k] = pp[k];
  }
  launch_mg_orth(m, pp, op.stream);
  CMX_HIP(hipGetLastError());
  op.sync();
  const auto t1 = std::chrono::steady_clock::now();

  // block stencil and colouring.  Even (or unit) coarse extents: period 2 (1) per direction, and each
  // probe is restricted per block and face (launch_mg_face_restrict): 2^4 colours.  Otherwise: period
  // p_mu >= 3 dividing the extent, so the 9 stencil entries of a block fall on distinct colours, and the
  // whole-block restriction is scattered to the entry of that colour.
  bool face = true;
  for (int mu = 0; mu < 4; ++mu) face = face && (m.Lc[mu] == 1 || m.Lc[mu] % 2 == 0);
  std::vector<int> nbr((size_t)m.Nb * MG_NST), col(m.Nb);
  st->ncol = 1;
  for (int mu = 0; mu < 4; ++mu) {
    int per = m.Lc[mu];
    if (face) {
      per = m.Lc[mu] == 1 ? 1 : 2;
    } else if (per > 2) {
      for (int q = 3; q <= m.Lc[mu]; ++q)
        if (m.Lc[mu] % q == 0) {
          per = q;
          break;
        }
    }
    st->period[mu] = per;
    st->ncol *= per;
  }
  for (int b = 0; b < m.Nb; ++b) {
    int bc[4], rr = b;
    for (int mu = 0; mu < 4; ++mu) bc[mu] = rr % m.Lc[mu], rr /= m.Lc[mu];
    int c = 0, s = 1;
    for (int mu = 0; mu < 4; ++mu) c += (bc[mu] % st->period[mu]) * s, s *= st->period[mu];
    col[b] = c;
    auto idx = [&](const int (&y)[4]) { return y[0] + m.Lc[0] * (y[1] + m.Lc[1] * (y[2] + m.Lc[2] * y[3])); };
    nbr[(size_t)b * MG_NST] = b;
    for (int mu = 0; mu < 4; ++mu)
      for (int sgn = 0; sgn < 2; ++sgn) {
        int y[4] = {bc[0], bc[1], bc[2], bc[3]};
        y[mu] = (y[mu] + (sgn == 0 ? 1 : m.Lc[mu] - 1)) % m.Lc[mu];
        nbr[(size_t)b * MG_NST + 1 + 2 * mu + sgn] = idx(y);
      }
  }
  st->nbr.alloc(sizeof(int) * nbr.size());
  st->colour.alloc(sizeof(int) * col.size());
  CMX_HIP(hipMemcpy(st->nbr.p, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice));
  CMX_HIP(hipMemcpy(st->colour.p, col.data(), sizeof(int) * col.size(), hipMemcpyHostToDevice));

  // the coarse operator by probing: for each colour and coarse degree of freedom, M applied to the
  // prolonged unit vectors of all blocks of that colour, restricted, and scattered to the stencil entry
  // of each block whose neighbour has that colour
  const size_t nac = (size_t)m.Nb * MG_NST * m.D * m.D;
  st->Ac.alloc(sizeof(float2) * nac);
  CMX_HIP(hipMemsetAsync(st->Ac.p, 0, sizeof(float2) * nac, op.stream));
  Vec G(P64, 2 * Vh);
  DevBuf y(sizeof(float2) * (size_t)m.Nb * m.D);
  for (int cl = 0; cl < st->ncol; ++cl)
    for (int c = 0; c < 2; ++c)
      for (int k = 0; k < m.nvec; ++k) {
        launch_mg_probe(m, st->pv, st->colour.as<int>(), cl, k, c, F.view(0), op.stream);
        apply_full64(op, F, G);
        if (face) {
          launch_mg_face_restrict(m, st->pv, st->colour.as<int>(), st->nbr.as<int>(), cl, c * m.nvec + k,
                                  G.view(0), st->Ac.as<float2>(), op.stream);
        } else {
          launch_mg_restrict(m, st->pv, P64, G.view(0), true, y.as<float2>(), op.stream);
          launch_mg_scatter(m, st->nbr.as<int>(), st->colour.as<int>(), cl, c * m.nvec + k, y.as<float2>(),
                            st->Ac.as<float2>(), op.stream);
        }
        ++res.n_probes;
      }
  CMX_HIP(hipGetLastError());
  op.sync();
  const auto t2 = std::chrono::steady_clock::now();

  // FP16 copies of the null vectors: an FP16 V-cycle restricts and prolongs with them (half the bytes of
  // the FP32 vectors; the coarse operator stays the Galerkin product of the FP32 ones)
  for (int k = 0; k < m.nvec; ++k) {
    st->P16.emplace_back(new Vec(P16, 2 * Vh));
    op.blas(B_COPY, {st->P16.back().get(), st->P[k].get()});
    st->pv.H[k] = st->P16.back()->buf.p;
    st->pv.Hn[k] = st->P16.back()->nrm();
  }

  // the coarse operator's FP16 copy: an FP16 V-cycle's coarse solve reads half the bytes (the solve
  // stops at a relative residual of coarse_tol, far above FP16 rounding)
  st->Ah.alloc(sizeof(float2) / 2 * nac);
  launch_mg_to_half((long)nac, st->Ac.as<float2>(), st->Ah.p, op.stream);
  CMX_HIP(hipGetLastError());

  // coarse GCR work space
  const size_t cv = sizeof(float2) * (size_t)m.Nb * m.D;
  st->xc.alloc(cv);
  st->rc.alloc(cv);
  st->bc.alloc(cv);
  st->pk.resize(p.coarse_nkrylov);
  st->qk.resize(p.coarse_nkrylov);
  for (int j = 0; j < p.coarse_nkrylov; ++j) st->pk[j].alloc(cv), st->qk[j].alloc(cv);
  const int nbk = mg_cdot_blocks(m.Nb * m.D);
  st->part.alloc(sizeof(double2) * nbk * (MG_MAXKRY + 2));
  st->S.alloc(sizeof(double) * MGS_N);
  st->F.alloc(sizeof(int) * MGF_N);

  op.mg = std::shared_ptr<void>(st.release(), [](void* q) { delete static_cast<MGState*>(q); });
  res.n_blocks = m.Nb;
  res.coarse_dof = m.D;
  res.nullspace_seconds = std::chrono::duration<double>(t1 - t0).count();
  res.coarse_op_seconds = std::chrono::duration<double>(t2 - t1).count();
  res.setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (p.verbose)
    std::printf("MG setup: %d blocks x %d dof, %d probes, null space %.2f s, coarse operator %.2f s, "
                "max |A x|/|x| %.3g\n",
                m.Nb, m.D, res.n_probes, res.nullspace_seconds, res.coarse_op_seconds, res.nullvec_rel_residual_max);
}

void mg_precondition(Op& op, int prec, const Vec& v, Vec& z) {
  MGState& st = state(op);
  const MGGeom& m = st.m;
  const int n = (int)op.Vh;
  CMX_REQUIRE(v.prec == prec && z.prec == prec && v.n == n && z.n == n, CMX_ERR_ARG, "multigrid: field mismatch");
  Vec &r = op.work(310, prec, n), &mr = op.work(311, prec, n), &t = op.work(312, prec, n);
  op.blas(B_ZERO, {&z});
  const Vec* res = &v;
  if (st.par.nu_pre > 0) {
    MRPrecond(op, prec, n, st.par.nu_pre, 0.0, st.par.omega, r, mr).apply(z, v);
    op.apply(prec, mr.view(), z.view(), +1, nullptr);
    op.blas(B_XMY_NORM, {&t, &v, &mr});  // t = v - A z
    res = &t;
  }
  launch_mg_restrict(m, st.pv, prec, res->view(), false, st.bc.as<float2>(), op.stream);
  coarse_gcr(st, op, st.bc.as<float2>(), st.xc.as<float2>(), prec == P16);
  launch_mg_prolong(m, st.pv, prec, st.xc.as<float2>(), z.view(), false, true, op.stream);
  CMX_HIP(hipGetLastError());
  if (st.par.smoother == 1) {  // the learned network on the residual of the coarse correction
    op.apply(prec, mr.view(), z.view(), +1, nullptr);
    op.blas(B_XMY_NORM, {&t, &v, &mr});  // t = v - A z
    learned_precond_apply(op, prec, t, r);
    op.blas(B_ADD, {&z, &r});            // z += N t
  }
  if (st.par.nu_post > 0) MRPrecond(op, prec, n, st.par.nu_post, 0.0, st.par.omega, r, mr).apply(z, v);
}

void mg_update(Op& op, const cmx_mg_params& p) {
  MGState& st = state(op);
  for (int mu = 0; mu < 4; ++mu)
    CMX_REQUIRE(p.blocking[mu] == st.par.blocking[mu], CMX_ERR_ARG,
                "multigrid update: Blocking differs from the set-up (a new null space needs cmx_mg_setup)");
  CMX_REQUIRE(p.nvec == st.par.nvec, CMX_ERR_ARG,
              "multigrid update: NullVectors differs from the set-up (a new null space needs cmx_mg_setup)");
  CMX_REQUIRE(p.nu_pre >= 0 && p.nu_post >= 0 && p.coarse_maxit >= 0, CMX_ERR_ARG,
              "multigrid: negative iteration count");
  CMX_REQUIRE(p.coarse_nkrylov >= 1 && p.coarse_nkrylov <= MG_MAXKRY, CMX_ERR_ARG,
              "multigrid: PrecondGCRNKrylov must be 1..32");
  CMX_REQUIRE(p.smoother == 0 || p.smoother == 1, CMX_ERR_ARG, "multigrid: smoother must be 0 (MR) or 1 (network)");
  op.sync();  // a V-cycle in flight may still read the GCR directions
  const size_t cv = sizeof(float2) * (size_t)st.m.Nb * st.m.D;
  for (int j = (int)st.pk.size(); j < p.coarse_nkrylov; ++j) {
    st.pk.emplace_back(cv);
    st.qk.emplace_back(cv);
  }
  // the set-up fields (setup_maxit, seed, prec) keep the values the null space was built with
  const int setup_maxit = st.par.setup_maxit, prec = st.par.prec;
  const unsigned long long seed = st.par.seed;
  st.par = p;
  st.par.setup_maxit = setup_maxit;
  st.par.seed = seed;
  st.par.prec = prec;
}

void mg_precondition_batch(const std::vector<Op*>& ops, int prec, const std::vector<const Vec*>& v,
                           const std::vector<Vec*>& z, BatchPlan& plan) {
  MGState& st = state(*ops[0]);
  const MGGeom& m = st.m;
  const int n = (int)ops[0]->Vh;
  if (ops.size() < 2 || st.par.nu_pre > 0 || st.par.smoother != 0) {
    for (size_t i = 0; i < ops.size(); ++i) mg_precondition(*ops[i], prec, *v[i], *z[i]);
    return;
  }
  // the coarse correction, one source after another on the shared stream (the coarse work space is shared)
  for (size_t i = 0; i < ops.size(); ++i) {
    Op& op = *ops[i];
    CMX_REQUIRE(op.mg.get() == ops[0]->mg.get() && op.stream == ops[0]->stream, CMX_ERR_ARG,
                "multigrid batch: contexts must share the multigrid and the stream");
    CMX_REQUIRE(v[i]->prec == prec && z[i]->prec == prec && v[i]->n == n && z[i]->n == n, CMX_ERR_ARG,
                "multigrid: field mismatch");
    op.blas(B_ZERO, {z[i]});
    launch_mg_restrict(m, st.pv, prec, v[i]->view(), false, st.bc.as<float2>(), op.stream);
    coarse_gcr(st, op, st.bc.as<float2>(), st.xc.as<float2>(), prec == P16);
    launch_mg_prolong(m, st.pv, prec, st.xc.as<float2>(), z[i]->view(), false, true, op.stream);
    CMX_HIP(hipGetLastError());
  }
  if (st.par.nu_post <= 0) return;
  // the MR smoothing of every source, recorded per source and run as one batched grid per kernel
  std::vector<Recorder> recs(ops.size());
  for (size_t i = 0; i < ops.size(); ++i) {
    Op& op = *ops[i];
    Vec &r = op.work(310, prec, n), &mr = op.work(311, prec, n);
    RecordScope rs(recs[i]);
    MRPrecond(op, prec, n, st.par.nu_post, 0.0, st.par.omega, r, mr).apply(*z[i], *v[i]);
  }
  plan.refresh(recs, ops[0]->stream);
  plan.run(ops[0]->stream);
}

void mg_info(Op& op, int* nb, int* d) {
  MGState& st = state(op);
  *nb = st.m.Nb;
  *d = st.m.D;
}

void mg_get_nullvecs(Op& op, double* out) {
  MGState& st = state(op);
  const int V = 2 * (int)op.Vh;
  for (int k = 0; k < st.m.nvec; ++k) op.unpack(out + (size_t)k * V * 24, V, *st.P[k]);
}

void mg_get_coarse_op(Op& op, double* out) {
  MGState& st = state(op);
  const size_t nac = (size_t)st.m.Nb * MG_NST * st.m.D * st.m.D;
  std::vector<float2> h(nac);
  CMX_HIP(hipMemcpy(h.data(), st.Ac.p, sizeof(float2) * nac, hipMemcpyDeviceToHost));
  // device entries are column-major; the host copy is [block][entry][row][column]
  const size_t D = (size_t)st.m.D;
  for (size_t e = 0; e < (size_t)st.m.Nb * MG_NST; ++e)
    for (size_t r = 0; r < D; ++r)
      for (size_t c = 0; c < D; ++c) {
        const float2 v = h[(e * D + c) * D + r];
        const size_t o = (e * D + r) * D + c;
        out[2 * o] = v.x, out[2 * o + 1] = v.y;
      }
}

void mg_restrict_host(Op& op, const double* full_in, double* coarse_out) {
  MGState& st = state(op);
  const int V = 2 * (int)op.Vh;
  Vec F(P64, V);
  op.pack(full_in, V, F);
  launch_mg_restrict(st.m, st.pv, P64, F.view(0), true, st.bc.as<float2>(), op.stream);
  const size_t n = (size_t)st.m.Nb * st.m.D;
  std::vector<float2> h(n);
  CMX_HIP(hipMemcpyAsync(h.data(), st.bc.p, sizeof(float2) * n, hipMemcpyDeviceToHost, op.stream));
  op.sync();
  for (size_t i = 0; i < n; ++i) coarse_out[2 * i] = h[i].x, coarse_out[2 * i + 1] = h[i].y;
}

void mg_prolong_host(Op& op, const double* coarse_in, double* full_out) {
  MGState& st = state(op);
  const int V = 2 * (int)op.Vh;
  const size_t n = (size_t)st.m.Nb * st.m.D;
  std::vector<float2> h(n);
  for (size_t i = 0; i < n; ++i) h[i] = make_float2((float)coarse_in[2 * i], (float)coarse_in[2 * i + 1]);
  CMX_HIP(hipMemcpy(st.xc.p, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  Vec F(P64, V);
  launch_mg_prolong(st.m, st.pv, P64, st.xc.as<float2>(), F.view(0), true, false, op.stream);
  op.unpack(full_out, V, F);
}

void mg_coarse_apply_host(Op& op, const double* in, double* out) {
  MGState& st = state(op);
  const size_t n = (size_t)st.m.Nb * st.m.D;
  std::vector<float2> h(n);
  for (size_t i = 0; i < n; ++i) h[i] = make_float2((float)in[2 * i], (float)in[2 * i + 1]);
  CMX_HIP(hipMemcpy(st.xc.p, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  launch_mg_coarse_apply(st.m, st.nbr.as<int>(), st.Ac.as<float2>(), st.xc.as<float2>(), st.rc.as<float2>(),
                         nullptr, op.stream);
  CMX_HIP(hipMemcpyAsync(h.data(), st.rc.p, sizeof(float2) * n, hipMemcpyDeviceToHost, op.stream));
  op.sync();
  for (size_t i = 0; i < n; ++i) out[2 * i] = h[i].x, out[2 * i + 1] = h[i].y;
}

}  // namespace cmx
