// Krylov solvers on the device.  Each iteration is a fixed sequence of kernel launches whose
// coefficients live in the device scalar array; the host launches iterations in batches and polls
// the convergence flag in between (see blas.h).  Iteration counts and stopping rules are those of
// the reference solvers:
//   CG_INVERTER        LinOpSysSolverCG (syssolver_linop_cg.h:57-96) + InvCG2_a (invcg2.cc:70-232)
//   BICGSTAB_INVERTER  InvBiCGStab_a (invbicgstab.cc:10-202)
//   RELIABLE CG        LinOpSysSolverReliableCGClover (syssolver_linop_rel_cg_clover.h:104-152) +
//                      RelInvCG_a (reliable_cg.cc:10-188): FP32/FP16 inner CG, FP64 reliable updates
//   RELIABLE BICGSTAB  LinOpSysSolverReliableBiCGStabClover (syssolver_linop_rel_bicgstab_clover.h:104-140)
//                      + RelInvBiCGStab_a (reliable_bicgstab.cc:14-290)
// The MdagM system solvers (psi = (M^dag M)^-1 chi) are the same iterations without the M^dag chi
// source preparation (CG, reliable CG), or two BiCGStab solves, M^dag Y = chi then M psi = Y
// (syssolver_mdagm_bicgstab.h:64-100, syssolver_mdagm_rel_bicgstab_clover.h:104-176).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "batch.h"
#include "fgmres.h"
#include "mg.h"
#include "op.h"

namespace cmx {


namespace {

struct Timer {
  hipEvent_t a, b;
  hipStream_t s;
  explicit Timer(hipStream_t s_) : s(s_) {
    CMX_HIP(hipEventCreate(&a));
    CMX_HIP(hipEventCreate(&b));
  }
  ~Timer() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  void start() { CMX_HIP(hipEventRecord(a, s)); }
  double stop() {
    CMX_HIP(hipEventRecord(b, s));
    CMX_HIP(hipEventSynchronize(b));
    float ms = 0;
    CMX_HIP(hipEventElapsedTime(&ms, a, b));
    return ms * 1e-3;
  }
};

void reset_state(Op& op, double rsd_sq, int maxit, double delta) {
  std::memset(op.hS, 0, sizeof(double) * S_NSLOTS);
  std::memset(op.hF, 0, sizeof(int) * F_NFLAGS);
  op.hS[S_RSDSQ] = rsd_sq;
  op.hS[S_DELTA] = delta;
  // BiCGStab: rho_0 = alpha = omega = 1 (invbicgstab.cc:64-69)
  op.hS[S_RHOP] = 1.0;
  op.hS[S_ALPHA] = 1.0;
  op.hS[S_OMEGA] = 1.0;
  op.hF[F_MAXIT] = maxit;
  CMX_HIP(hipMemcpyAsync(op.state_S.p, op.hS, sizeof(double) * S_NSLOTS, hipMemcpyHostToDevice, op.stream));
  CMX_HIP(hipMemcpyAsync(op.state_F.p, op.hF, sizeof(int) * F_NFLAGS, hipMemcpyHostToDevice, op.stream));
  op.sync();
}

void read_state(Op& op) {
  CMX_HIP(hipMemcpyAsync(op.hS, op.state_S.p, sizeof(double) * S_NSLOTS, hipMemcpyDeviceToHost, op.stream));
  CMX_HIP(hipMemcpyAsync(op.hF, op.state_F.p, sizeof(int) * F_NFLAGS, hipMemcpyDeviceToHost, op.stream));
  op.sync();
}

// |v|^2 on the host (one synchronisation)
double norm2(Op& op, const Vec& v) {
  op.blas(B_NORM2, {&v});
  op.finalize(blas_blocks(v.n), 1, PG_STORE, S_TR);
  CMX_HIP(hipMemcpyAsync(op.hS + S_TR, static_cast<double*>(op.state_S.p) + S_TR, sizeof(double),
                         hipMemcpyDeviceToHost, op.stream));
  op.sync();
  return op.hS[S_TR];
}

// Replays a fixed sequence of launches on the operator's stream, optionally from a captured HIP graph
// kept in the operator (keyed by solver tag, sequence length and the solution / source buffers; the
// solver's other vectors are the operator's persistent workspace).  The sequence must not allocate,
// synchronise or copy to the host; the solvers run it once eagerly before the first capture, so every
// lazy allocation has happened.
struct GraphCache {
  Op& op;
  bool on;
  long tag;
  const void *psi, *chi;
  GraphCache(Op& o, bool use, long tag_, const Vec& psi_, const Vec& chi_)
      : op(o), on(use), tag(tag_), psi(psi_.buf.p), chi(chi_.buf.p) {}
  template <typename Seq> void run(long key, Seq&& seq) {
    if (!on) {
      seq();
      return;
    }
    for (auto& e : op.graphs)
      if (e.tag == tag && e.batch == key && e.psi == psi && e.chi == chi) {
        if (e.gen == op.ws_gen) {
          CMX_HIP(hipGraphLaunch(e.exec, op.stream));
          return;
        }
        (void)hipGraphExecDestroy(e.exec);  // the workspace moved: capture again
        e = op.graphs.back();
        op.graphs.pop_back();
        break;
      }
    hipGraph_t gr = nullptr;
    CMX_HIP(hipStreamBeginCapture(op.stream, hipStreamCaptureModeThreadLocal));
    try {
      seq();
    } catch (...) {
      (void)hipStreamEndCapture(op.stream, &gr);
      if (gr) (void)hipGraphDestroy(gr);
      throw;
    }
    CMX_HIP(hipStreamEndCapture(op.stream, &gr));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    CMX_HIP(e);
    op.graphs.push_back(Op::GraphEntry{tag, key, op.ws_gen, psi, chi, ex});
    CMX_HIP(hipGraphLaunch(ex, op.stream));
  }
};

// Runs iterations in batches until the device reports convergence (or a stop).  With p.graph the
// batches after the first are replayed from captured graphs.
template <typename Iter>
int run_batches(Op& op, const cmx_solver_params& p, long tag, const Vec& psi, const Vec& chi, Iter&& iter) {
  int batch = p.check_every > 0 ? p.check_every : 4;
  long launched = 0;
  GraphCache gc(op, p.graph != 0, tag, psi, chi);
  bool first = true;
  while (true) {
    if (first) {
      for (int j = 0; j < batch; ++j) iter();
      first = false;
    } else {
      gc.run(batch, [&] {
        for (int j = 0; j < batch; ++j) iter();
      });
    }
    launched += batch;
    CMX_HIP(hipMemcpyAsync(op.hF, op.state_F.p, sizeof(int) * F_NFLAGS, hipMemcpyDeviceToHost, op.stream));
    op.sync();
    if (op.hF[F_CONV]) break;
    if (launched > (long)p.max_iter + 8) break;  // the device stops itself at MaxIter; safety net
    if (p.check_every <= 0 && batch < 32) batch *= 2;
  }
  return op.hF[F_NCOUNT];
}

double flops_per_apply(const Op& op) {
  const double nc_ns = 12.0;
  switch (op.par.kind) {
    case CMX_OP_EO_CLOVER: return (2 * 1320.0 + 2 * 552.0 + 4 * nc_ns) * op.Vh;  // eoprec_clover_linop_w.cc:299-306
    case CMX_OP_EO_WILSON: return (2 * 1320.0 + 6 * nc_ns) * op.Vh;             // eoprec_wilson_linop_w.cc:247-251
    default: return (1320.0 + 4 * nc_ns) * op.V;                                 // unprec_wilson_linop_w.cc:109-113
  }
}

// ---- CG_INVERTER: CGNE on M^dag M in FP64 ---------------------------------------------------------
// mdagm: chi is already the source of M^dag M psi = chi (MdagMSysSolverCG, syssolver_mdagm_cg.h:70)
// The pieces (setup, one iteration, finish) serve the single solve and the batched solves.
struct CGState {
  Op& op;
  const cmx_solver_params& p;
  Vec& psi;
  const Vec& chi;
  bool mdagm;
  int n, nb;
  Vec &chi_tmp, &r, &pv, &mp, &mmp;
  CGState(Op& o, const cmx_solver_params& p_, Vec& psi_, const Vec& chi_, bool mdagm_)
      : op(o), p(p_), psi(psi_), chi(chi_), mdagm(mdagm_), n(o.subset_sites()), nb(blas_blocks(n)),
        chi_tmp(o.work(10, P64, n)), r(o.work(11, P64, n)), pv(o.work(12, P64, n)), mp(o.work(13, P64, n)),
        mmp(o.work(14, P64, n)) {}
  // chi_tmp = M^dag chi, the tolerance, and the start r = chi_tmp - M^dag M psi ; p = r
  void setup(Timer* tm) {
    if (mdagm)  // syssolver_linop_cg.h:65-66
      op.blas(B_COPY, {&chi_tmp, &chi});
    else
      op.apply(P64, chi_tmp.view(), chi.view(), -1, nullptr);
    const double chi_sq = norm2(op, chi_tmp);
    reset_state(op, p.rsd_target * p.rsd_target * chi_sq, p.max_iter, p.delta);
    if (tm) tm->start();
    // invcg2.cc:109-131
    op.apply(P64, mp.view(), psi.view(), +1, nullptr);
    op.apply(P64, mmp.view(), mp.view(), -1, nullptr);
    op.blas(B_XMY_NORM, {&r, &chi_tmp, &mmp});
    op.finalize(nb, 1, PG_CG_INIT);
    op.blas(B_COPY, {&pv, &r});
  }
  void iter() {
    int* F = op.state_F.as<int>();
    const int np = op.apply(P64, mp.view(), pv.view(), +1, op.partials.as<double>(), F + F_DONE);  // Mp, |Mp|^2
    op.finalize(np, 1, PG_CG_D);                                                                 // a = c/d
    op.apply(P64, mmp.view(), mp.view(), -1, nullptr, F + F_DONE);                               // M^dag M p
    op.blas(B_AXPY_NORM, {&r, &mmp}, S_NEGA, 0, F_DONE);                                         // r -= a M^dag M p
    op.finalize(nb, 1, PG_CG_R);                                                                 // cp, b, test
    op.blas(B_CG_UPDATE, {&psi, &pv, &r}, S_A, S_B, F_DONE);                                     // psi += a p; p = r + b p
  }
  // results once the device has stopped (state read back)
  void finish(double seconds, cmx_solver_result& res) {
    res.n_count = op.hF[F_NCOUNT];
    res.solve_seconds = seconds;
    const int err = op.hF[F_ERR];
    res.gflops = (res.n_count * (2 * flops_per_apply(op) + 10 * 12.0 * n) + 2 * flops_per_apply(op)) /
                 std::max(seconds, 1e-12) * 1e-9;
    if (err == CMX_ERR_BREAKDOWN) throw Error(CMX_ERR_BREAKDOWN, "CG: breakdown (|M p| = 0 or not finite)");
    if (err == CMX_ERR_NOCONV) res.n_count = p.max_iter;  // InvCG2 returns MaxCG with a warning
  }
};

void solve_cg(Op& op, const cmx_solver_params& p, Vec& psi, const Vec& chi, bool mdagm, cmx_solver_result& res) {
  CGState st(op, p, psi, chi, mdagm);
  Timer tm(op.stream);
  st.setup(&tm);
  run_batches(op, p, 1 + (mdagm ? 100 : 0), psi, chi, [&] { st.iter(); });
  read_state(op);
  st.finish(tm.stop(), res);
}

// ---- BICGSTAB_INVERTER ----------------------------------------------------------------------------
void solve_bicgstab(Op& op, const cmx_solver_params& p, Vec& psi, const Vec& chi, int isign, cmx_solver_result& res) {
  const int n = op.subset_sites();
  Vec &r = op.work(20, P64, n), &r0 = op.work(21, P64, n), &pv = op.work(22, P64, n), &v = op.work(23, P64, n),
      &t = op.work(24, P64, n);
  const int nb = blas_blocks(n);
  const double chi_sq = norm2(op, chi);
  reset_state(op, p.rsd_target * p.rsd_target * chi_sq, p.max_iter, p.delta);
  Timer tm(op.stream);
  tm.start();
  int* F = op.state_F.as<int>();
  // r = r0 = chi - A psi ; p = v = 0   (invbicgstab.cc:37-62)
  op.apply(P64, r0.view(), psi.view(), isign, nullptr);
  op.blas(B_XMY_NORM, {&r, &chi, &r0});
  op.blas(B_COPY, {&r0, &r});
  op.blas(B_ZERO, {&pv});
  op.blas(B_ZERO, {&v});
  auto iter = [&]() {
    op.blas(B_CDOT, {&r0, &r}, 0, 0, F_DONE);  // rho = <r0|r>
    op.finalize(nb, 2, PG_BICG_RHO);           // beta = (rho/rho_prev)(alpha/omega)
    op.blas(B_BICG_P, {&pv, &r, &v}, S_OMEGA, S_BETAC, F_DONE);  // p = r + beta (p - omega v)
    op.apply(P64, v.view(), pv.view(), isign, nullptr, F + F_DONE);  // v = A p
    op.blas(B_CDOT, {&r0, &v}, 0, 0, F_DONE);
    op.finalize(nb, 2, PG_BICG_ALPHA);                            // alpha = rho / <r0|v>
    op.blas(B_CAXPY, {&r, &v}, S_NALPHA, 0, F_DONE);              // s = r - alpha v
    op.apply(P64, t.view(), r.view(), isign, nullptr, F + F_DONE);   // t = A s
    op.blas(B_TNORM_TDOT, {&t, &r}, 0, 0, F_DONE);
    op.finalize(nb, 3, PG_BICG_OMEGA);                            // omega = <t|s>/|t|^2
    op.blas(B_BICG_UPDATE, {&psi, &r, &pv, &t}, S_OMEGA, S_ALPHA, F_DONE);  // psi += omega s + alpha p; r = s - omega t
    op.finalize(nb, 1, PG_BICG_R);
  };
  res.n_count = run_batches(op, p, 2 + (isign < 0 ? 100 : 0), psi, chi, iter);
  read_state(op);
  res.solve_seconds = tm.stop();
  res.gflops = res.n_count * (2 * flops_per_apply(op) + 80 * 12.0 * n) / std::max(res.solve_seconds, 1e-12) * 1e-9;
  if (op.hF[F_ERR] == CMX_ERR_BREAKDOWN) throw Error(CMX_ERR_BREAKDOWN, "BiCGStab breakdown");
  if (op.hF[F_ERR] == CMX_ERR_NOCONV) res.n_count = p.max_iter;
}

// ---- reliable-update mixed-precision CG on M^dag M ----------------------------------------------------
struct RelCGState {
  Op& op;
  const cmx_solver_params& p;
  Vec& psi;
  const Vec& chi;
  bool mdagm;
  int n, ps, nb;
  Vec &chi_tmp, &b, &r_dble, &x_dble, &tmp1, &tmp2;
  Vec &r, &x, &pv, &mp, &mmp;
  static int check_prec(int ps) {
    CMX_REQUIRE(ps == P32 || ps == P16, CMX_ERR_ARG, "reliable CG: sloppy precision must be single or half");
    return ps;
  }
  RelCGState(Op& o, const cmx_solver_params& p_, Vec& psi_, const Vec& chi_, bool mdagm_)
      : op(o), p(p_), psi(psi_), chi(chi_), mdagm(mdagm_), n(o.subset_sites()), ps(check_prec(p_.sloppy_prec)),
        nb(blas_blocks(n)), chi_tmp(o.work(30, P64, n)), b(o.work(31, P64, n)), r_dble(o.work(32, P64, n)),
        x_dble(o.work(33, P64, n)), tmp1(o.work(34, P64, n)), tmp2(o.work(35, P64, n)), r(o.work(36, P32, n)),
        x(o.work(37, P32, n)), pv(o.work(38, ps, n)), mp(o.work(39, ps, n)), mmp(o.work(40, ps, n)) {}
  void setup(Timer* tm) {
    // chi_tmp = M^dag chi (syssolver_linop_rel_cg_clover.h:121-125); the MdagM solver takes chi as is
    // (syssolver_mdagm_rel_cg_clover.h:126-133)
    if (mdagm)
      op.blas(B_COPY, {&chi_tmp, &chi});
    else
      op.apply(P64, chi_tmp.view(), chi.view(), -1, nullptr);
    const double chi_sq = norm2(op, chi_tmp);
    reset_state(op, p.rsd_target * p.rsd_target * chi_sq, p.max_iter, p.delta);
    op.ensure_prec(ps);
    if (tm) tm->start();
    // b = chi - A^dag A psi ; r = b ; x = 0   (reliable_cg.cc:44-64)
    op.apply(P64, tmp1.view(), psi.view(), +1, nullptr);
    op.apply(P64, tmp2.view(), tmp1.view(), -1, nullptr);
    op.blas(B_XMY_NORM_COPY, {&b, &chi_tmp, &tmp2, &r});
    op.finalize(nb, 1, PG_RCG_INIT);
    op.blas(B_ZERO, {&x});
    op.blas(B_ZERO, {&pv});
  }
  void inner() {
    int* F = op.state_F.as<int>();
    op.blas(B_XPAY, {&pv, &r}, S_BETA, 0, F_DONE);                                          // p = r + beta p
    const int np = op.apply(ps, mp.view(), pv.view(), +1, op.partials.as<double>(), F + F_DONE);
    op.finalize(np, 1, PG_RCG_D);                                                           // c = r_sq; a = c/d
    op.apply(ps, mmp.view(), mp.view(), -1, nullptr, F + F_DONE);
    op.blas(B_AXPY2_NORM, {&x, &r, &pv, &mmp}, S_A, 0, F_DONE);                             // x += a p; r -= a mmp
    op.finalize(nb, 1, PG_RCG_R);                                                           // updateX / updateR
  }
  // r = b - A^dag A x in FP64 (reliable_cg.cc:120-136)
  void update_r(int skip, int run) {
    int* F = op.state_F.as<int>();
    op.blas(B_COPY, {&x_dble, &x}, 0, 0, skip, run);
    op.apply(P64, tmp1.view(), x_dble.view(), +1, nullptr, skip >= 0 ? F + skip : nullptr, run >= 0 ? F + run : nullptr);
    op.apply(P64, tmp2.view(), tmp1.view(), -1, nullptr, skip >= 0 ? F + skip : nullptr, run >= 0 ? F + run : nullptr);
    op.blas(B_XMY_NORM_COPY, {&r_dble, &b, &tmp2, &r}, 0, 0, skip, run);
  }
  // The update step with the decisions taken on the device: the FP64 kernels run when the flags ask
  // for them and return at once otherwise (same arithmetic as the host-decided step)
  void update_gated() {
    update_r(F_DONE, F_UPDR);
    op.finalize(nb, 1, PG_RCG_UPD);                                                         // norms, convergence
    op.blas(B_ADD_ZERO_COPY, {&psi, &x, &b, &r_dble}, 0, 0, F_DONE, F_UPDX);                // psi += x; x = 0; b = r
  }
  void finish(double seconds, cmx_solver_result& res) {
    res.n_count = op.hF[F_NCOUNT];
    res.solve_seconds = seconds;
    res.n_reliable = op.hF[F_NUPD];
    res.gflops = (res.n_count + 1) * (2 * flops_per_apply(op) + 16 * 12.0 * n) / std::max(seconds, 1e-12) * 1e-9;
    if (op.hF[F_ERR] == CMX_ERR_BREAKDOWN) throw Error(CMX_ERR_BREAKDOWN, "reliable CG: breakdown");
    if (op.hF[F_ERR] == CMX_ERR_NOCONV)
      throw Error(CMX_ERR_NOCONV, "reliable CG: no convergence in " + std::to_string(p.max_iter) + " iterations");
  }
};

void solve_relcg(Op& op, const cmx_solver_params& p, Vec& psi, const Vec& chi, bool mdagm, cmx_solver_result& res) {
  RelCGState st(op, p, psi, chi, mdagm);
  Timer tm(op.stream);
  st.setup(&tm);
  // One host synchronisation per iteration: the flags of the updateR / updateX decision are read
  // back, and the FP64 reliable-update kernels are launched only when the device asks for them
  // (an early-exit launch of a full-lattice grid costs tens of microseconds, more than the read-back).
  long launched = 0;
  GraphCache gc(op, p.graph != 0, 3 + 10 * st.ps + (mdagm ? 100 : 0), psi, chi);
  auto inner = [&]() { st.inner(); };
  while (true) {
    if (launched == 0)
      inner();  // eager: performs the lazy allocations of the operator's scratch vectors
    else
      gc.run(0, inner);
    CMX_HIP(hipMemcpyAsync(op.hF, op.state_F.p, sizeof(int) * F_NFLAGS, hipMemcpyDeviceToHost, op.stream));
    if (p.verbose)
      CMX_HIP(hipMemcpyAsync(op.hS, op.state_S.p, sizeof(double) * 16, hipMemcpyDeviceToHost, op.stream));
    op.sync();
    if (p.verbose)  // reliable_cg.cc:108-110
      std::printf("RelCG: k = %d  r_sq = %.6e  d = %.6e  a = %.6e  updateR %d updateX %d err %d\n", op.hF[F_K],
                  op.hS[S_RSQ], op.hS[S_D], op.hS[S_A], op.hF[F_UPDR], op.hF[F_UPDX], op.hF[F_ERR]);
    if (op.hF[F_DONE]) break;  // converged (or stopped) in the previous iteration
    if (op.hF[F_UPDR]) st.update_r(-1, -1);
    op.finalize(st.nb, 1, PG_RCG_UPD);                                                      // norms, convergence
    if (op.hF[F_UPDX]) op.blas(B_ADD_ZERO_COPY, {&psi, &st.x, &st.b, &st.r_dble});          // psi += x; x = 0; b = r
    if (++launched > (long)p.max_iter + 4) break;                                           // device stops at MaxIter
  }
  op.blas(B_ADD, {&psi, &st.x});                                                            // psi += x (reliable_cg.cc:158-163)
  read_state(op);
  st.finish(tm.stop(), res);
}

// ---- reliable-update mixed-precision BiCGStab on M (isign) ------------------------------------------
// Vectors r, r0, p, v, t are stored in the sloppy precision (the operator's input and output), the
// accumulated solution x in FP32 (FP64 when the sloppy precision is FP64); b, x_dble, r_dble, tmp are
// FP64.  The reference's five BLAS passes per iteration are fused into four-vector passes:
// r -= alpha v with x += alpha p, and x += omega s with r = s - omega t, |r|^2 and <r0|r>.
void solve_relbicgstab(Op& op, const cmx_solver_params& p, Vec& psi, const Vec& chi, int isign,
                       cmx_solver_result& res) {
  const int n = op.subset_sites();
  const int ps = p.sloppy_prec;
  CMX_REQUIRE(ps == P64 || ps == P32 || ps == P16, CMX_ERR_ARG, "reliable BiCGStab: bad sloppy precision");
  const int px = ps == P64 ? P64 : P32;
  const int nb = blas_blocks(n);
  Vec b(P64, n), tmp(P64, n), r_dble(P64, n), x_dble(P64, n);
  Vec r(ps, n), r0(ps, n), pv(ps, n), v(ps, n), t(ps, n), x(px, n);
  const double chi_sq = norm2(op, chi);  // rsd_sq = Rsd^2 |chi|^2 (:65)
  reset_state(op, p.rsd_target * p.rsd_target * chi_sq, p.max_iter, p.delta);
  Timer tm(op.stream);
  tm.start();
  int* F = op.state_F.as<int>();
  // b = chi - A psi ; r = b ; r0 = b ; x = p = v = 0   (:57-79)
  op.apply(P64, tmp.view(), psi.view(), isign, nullptr);
  op.blas(B_XMY_NORM_COPY, {&b, &chi, &tmp, &r});
  op.finalize(nb, 1, PG_RBICG_INIT);
  op.blas(B_COPY, {&r0, &r});
  op.blas(B_ZERO, {&x});
  op.blas(B_ZERO, {&pv});
  op.blas(B_ZERO, {&v});
  long launched = 0;
  while (true) {
    op.blas(B_BICG_P, {&pv, &r, &v}, S_OMEGA, S_BETAC);                        // p = r + beta (p - omega v)
    op.apply(ps, v.view(), pv.view(), isign, nullptr);                         // v = A p
    op.blas(B_CDOT, {&r0, &v});
    op.finalize(nb, 2, PG_RBICG_ALPHA);                                        // alpha = rho / <r0|v>
    op.blas(B_CAXPY2, {&r, &v, &x, &pv}, S_NALPHA, S_ALPHA, F_DONE);           // s = r - alpha v; x += alpha p
    op.apply(ps, t.view(), r.view(), isign, nullptr, F + F_DONE);              // t = A s
    op.blas(B_TNORM_TDOT, {&t, &r}, 0, 0, F_DONE);
    op.finalize(nb, 3, PG_BICG_OMEGA);                                         // omega = <t|s> / |t|^2
    op.blas(B_RBICG_UPDATE, {&x, &r, &t, &r0}, S_OMEGA, 0, F_DONE);            // x += omega s; r = s - omega t
    op.finalize(nb, 3, PG_RBICG_R);                                            // updateX / updateR
    CMX_HIP(hipMemcpyAsync(op.hF, op.state_F.p, sizeof(int) * F_NFLAGS, hipMemcpyDeviceToHost, op.stream));
    op.sync();
    if (op.hF[F_DONE] || op.hF[F_ERR]) break;  // breakdown inside this iteration
    if (op.hF[F_UPDR]) {                       // r = b - A x in FP64 (:216-256)
      op.blas(B_COPY, {&x_dble, &x});
      op.apply(P64, tmp.view(), x_dble.view(), isign, nullptr);
      op.blas(B_XMY_NORM_COPY, {&r_dble, &b, &tmp, &r});
      op.finalize(nb, 1, PG_RBICG_UPD);
      if (op.hF[F_UPDX]) op.blas(B_ADD_ZERO_COPY, {&psi, &x, &b, &r_dble});  // psi += x; x = 0; b = r
      CMX_HIP(hipMemcpyAsync(op.hF, op.state_F.p, sizeof(int) * F_NFLAGS, hipMemcpyDeviceToHost, op.stream));
      op.sync();
    }
    if (op.hF[F_CONV]) break;
    if (++launched > (long)p.max_iter + 4) break;  // the device stops at MaxIter
  }
  op.blas(B_ADD, {&psi, &x});  // psi += x (:259-266)
  res.n_count = op.hF[F_NCOUNT];
  read_state(op);
  res.solve_seconds = tm.stop();
  res.n_reliable = op.hF[F_NUPD];
  res.gflops = (res.n_count + 1) * (2 * flops_per_apply(op) + 80 * 12.0 * n) / std::max(res.solve_seconds, 1e-12) * 1e-9;
  if (op.hF[F_ERR] == CMX_ERR_BREAKDOWN) throw Error(CMX_ERR_BREAKDOWN, "reliable BiCGStab: breakdown");
  if (op.hF[F_ERR] == CMX_ERR_NOCONV)
    throw Error(CMX_ERR_NOCONV, "reliable BiCGStab: no convergence in " + std::to_string(p.max_iter) + " iterations");
}

// one pass of the selected solver on A psi = chi (isign = +1), or on M^dag M psi = chi (mdagm)
void solve_pass(Op& op, const cmx_solver_params& q, Vec& psi, const Vec& chi, bool mdagm, cmx_solver_result& part) {
  if (q.solver == CMX_SOLVER_FGMRESDR) {
    CMX_REQUIRE(!mdagm, CMX_ERR_UNSUPPORTED, "FGMRES-DR is a LinOp system solver (no MdagM form)");
    solve_fgmresdr(op, q, psi, chi, part);
    return;
  }
  switch (q.solver) {
    case CMX_SOLVER_CG: solve_cg(op, q, psi, chi, mdagm, part); return;
    case CMX_SOLVER_RELIABLE_CG: solve_relcg(op, q, psi, chi, mdagm, part); return;
    case CMX_SOLVER_BICGSTAB:
    case CMX_SOLVER_RELIABLE_BICGSTAB: break;
    default: throw Error(CMX_ERR_ARG, "solve: unknown solver");
  }
  auto pass = [&](Vec& x, const Vec& rhs, int isign, cmx_solver_result& r) {
    if (q.solver == CMX_SOLVER_BICGSTAB)
      solve_bicgstab(op, q, x, rhs, isign, r);
    else
      solve_relbicgstab(op, q, x, rhs, isign, r);
  };
  if (!mdagm) {
    pass(psi, chi, +1, part);
    return;
  }
  // M^dag Y = chi from the guess Y = M psi, then M psi = Y from the guess psi
  const int n = op.subset_sites();
  Vec& Y = op.work(4, P64, n);
  op.apply(P64, Y.view(), psi.view(), +1, nullptr);
  cmx_solver_result r1, r2;
  std::memset(&r1, 0, sizeof(r1));
  std::memset(&r2, 0, sizeof(r2));
  pass(Y, chi, -1, r1);
  pass(psi, Y, +1, r2);
  part.n_count = r1.n_count + r2.n_count;
  part.n_reliable = r1.n_reliable + r2.n_reliable;
  part.solve_seconds = r1.solve_seconds + r2.solve_seconds;
  part.gflops = part.solve_seconds > 0
                    ? (r1.gflops * r1.solve_seconds + r2.gflops * r2.solve_seconds) / part.solve_seconds
                    : 0.0;
}

// resid = |chi - A psi| in FP64 (syssolver_linop_cg.h:80-87), or |chi - M^dag M psi|
// (syssolver_mdagm_cg.h:77-85), and resid / |chi|; workspace slots 2, 3
void true_residual(Op& op, const Vec& psi, const Vec& chi, bool mdagm, cmx_solver_result& res) {
  const int n = op.subset_sites();
  Vec &t = op.work(2, P64, n), &rr = op.work(3, P64, n);
  op.apply(P64, t.view(), psi.view(), +1, nullptr);
  if (mdagm) {
    op.apply(P64, rr.view(), t.view(), -1, nullptr);
    op.blas(B_COPY, {&t, &rr});
  }
  op.blas(B_XMY_NORM, {&rr, &chi, &t});
  op.finalize(blas_blocks(n), 1, PG_STORE, S_TR);
  op.blas(B_NORM2, {&chi});
  op.finalize(blas_blocks(n), 1, PG_STORE, S_TR + 1);
  CMX_HIP(hipMemcpyAsync(op.hS + S_TR, static_cast<double*>(op.state_S.p) + S_TR, 2 * sizeof(double),
                         hipMemcpyDeviceToHost, op.stream));
  op.sync();
  res.resid = std::sqrt(op.hS[S_TR]);
  res.rel_resid = op.hS[S_TR + 1] > 0 ? res.resid / std::sqrt(op.hS[S_TR + 1]) : res.resid;
}

}  // namespace

// ---- batched solves (batch.h) ---------------------------------------------------------------------
// Each problem's iteration is recorded once (the reliable-update step in its flag-gated form, so that
// the sequence is fixed), the recordings are merged into batched stages, and the host replays the
// stages until every problem has stopped.  Problems that stop early return at the first instruction of
// each later kernel.
namespace {

// candidate launch shapes for a batch: the plane-grid kernel needs it admitted on every lattice
std::vector<HopConfig> batch_candidates(const std::vector<Op*>& ops, int ps) {
  bool fast_ok = ps != P64;
  for (Op* o : ops) {
    HopLaunch h = o->hop_base(ps);
    HopConfig c;
    c.fast = 1;
    hop_apply_config(h, c);
    if (!(h.fast & 1)) fast_ok = false;
  }
  std::vector<HopConfig> c;
  for (int block : {64, 128, 256}) {
    c.push_back(HopConfig{block, 0, 1, 0, 0});
    c.push_back(HopConfig{block, 0, 1, 0, HOP_FAST_STREAM});
    if (fast_ok) {
      c.push_back(HopConfig{block, 0, 1, 0, 1});  // links kept in L2 (shared by problems on one gauge field)
      c.push_back(HopConfig{block, 0, 1, 0, 1 | HOP_FAST_STREAM});
    }
  }
  return c;
}

}  // namespace

namespace {

// Device bytes of one FGMRES-DR solve's workspace (the solution, source and residual vectors in FP64,
// the bases and spares in the basis precision, the staging and MR vectors in the operator precision)
size_t fgmres_workspace_bytes(const Op& op, const cmx_solver_params& p) {
  const long n = op.subset_sites();
  const int ps = p.sloppy_prec, pb = p.basis_prec < 0 ? ps : p.basis_prec;
  const int nd = std::max(0, p.n_defl), total = p.n_krylov + nd;
  size_t b = 6 * spinor_bytes(P64, n);
  b += (size_t)(2 * total + 1 + 2 * nd + 1 + (nd > 0 && pb != P64 ? 1 : 0)) * spinor_bytes(pb, n);
  if (pb != ps) b += 2 * spinor_bytes(ps, n);
  if (p.precond == CMX_PRECOND_MR) b += 2 * spinor_bytes(ps, n);
  b += 2 * spinor_bytes(ps, n);  // the operator's even-checkerboard temporaries of a solver context
  // the batched network: a mixed field (8 paths x 3 colours x 8 B per site) and two block outputs per source
  if (p.precond == CMX_PRECOND_LEARNED) b += (size_t)192 * n + 2 * spinor_bytes(ps, n);
  return b;
}

// FGMRES-DR over n problems: every solve is a FgmresDR stepping object on its own solver context; at each
// tick every solve in its Arnoldi loop issues one step, the streams are synchronised once, and each solve
// takes its new column (the Givens update, on the host).  Cycle ends and restarts run per solve.  Each
// solve performs exactly the kernels and host arithmetic of cmx_solve, so its result is the same bit for
// bit.  As many solves run at once as their workspaces fit in free device memory; the others wait.
void solve_batch_fgmres(const std::vector<Op*>& ops, const cmx_solver_params& p, int max_concurrent,
                        double* const* psi_h, const double* const* chi_h, cmx_solver_result* res,
                        cmx_batch_result& bres) {
  const int np = (int)ops.size();
  std::vector<hipStream_t> streams;
  for (Op* o : ops)
    if (std::find(streams.begin(), streams.end(), o->stream) == streams.end()) streams.push_back(o->stream);
  auto sync_all = [&]() {
    for (hipStream_t s : streams) CMX_HIP(hipStreamSynchronize(s));
  };
  // concurrency from free memory and the per-solve workspace
  size_t free_b = 0, total_b = 0;
  CMX_HIP(hipMemGetInfo(&free_b, &total_b));
  size_t need = 0;
  for (Op* o : ops) need = std::max(need, fgmres_workspace_bytes(*o, p));
  int conc = (int)std::max<size_t>(1, std::min<size_t>((size_t)np, (size_t)(0.85 * (double)free_b) / std::max<size_t>(need, 1)));
  if (max_concurrent > 0) conc = std::min(conc, max_concurrent);

  // problem i runs on solver context ctx[i]: its own, or the context of a finished problem on the same
  // operator (no new workspace)
  std::vector<Op*> ctx(ops);
  auto same_operator = [](const Op& a, const Op& b) {
    return a.gd == b.gd && std::memcmp(&a.par, &b.par, sizeof(a.par)) == 0 && a.net == b.net && a.ptc == b.ptc &&
           a.mg == b.mg && std::memcmp(a.hcfg, b.hcfg, sizeof(a.hcfg)) == 0 &&
           std::memcmp(a.recon, b.recon, sizeof(a.recon)) == 0 && a.stream == b.stream;
  };
  std::vector<cmx_solver_params> q(np, p);
  for (int i = 0; i < np; ++i) {
    std::memset(&res[i], 0, sizeof(res[i]));
    if (p.true_residual) q[i].rsd_target = 0.5 * p.rsd_target;
  }
  std::vector<std::unique_ptr<FgmresDR>> run(np);
  std::vector<int> stepping(np, 0);  // 1: inside an Arnoldi cycle
  std::vector<int> err(np, CMX_OK);
  std::vector<std::string> msg(np);
  std::vector<int> pending, active;
  std::vector<char> packed(np, 0), unpacked(np, 0);
  for (int i = 0; i < np; ++i) pending.push_back(i);
  // the source and guess on the device and the workspace allocated
  auto prepare = [&](int i) {
    Op& op = *ctx[i];
    const int n = op.subset_sites();
    op.pack(chi_h[i], n, op.work(1, P64, n));
    op.pack(psi_h[i], n, op.work(0, P64, n));
    op.work(2, P64, n);
    op.work(3, P64, n);
    FgmresDR::reserve(op, q[i]);
    packed[i] = 1;
  };
  // the timed region: device time of the solves, without the host transfers of the solves that start
  // late or finish early (they pause it)
  struct Seg {
    hipEvent_t a = nullptr, b = nullptr;
    ~Seg() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } seg;
  CMX_HIP(hipEventCreate(&seg.a));
  CMX_HIP(hipEventCreate(&seg.b));
  double timed = 0;
  bool timing = false;
  hipStream_t s0 = ctx[0]->stream;
  auto resume = [&]() {
    if (timing) return;
    CMX_HIP(hipEventRecord(seg.a, s0));
    timing = true;
  };
  auto pause = [&]() {
    if (!timing) return;
    sync_all();
    CMX_HIP(hipEventRecord(seg.b, s0));
    CMX_HIP(hipEventSynchronize(seg.b));
    float ms = 0;
    CMX_HIP(hipEventElapsedTime(&ms, seg.a, seg.b));
    timed += ms * 1e-3;
    timing = false;
  };
  auto start = [&](int i) {
    Op& op = *ctx[i];
    const int n = op.subset_sites();
    if (!packed[i]) {
      pause();
      prepare(i);
      resume();
    }
    run[i].reset(new FgmresDR(op, q[i], op.work(0, P64, n), op.work(1, P64, n)));
    stepping[i] = 0;
  };
  // the first wave is prepared before the timed region
  for (int k = 0; k < std::min(conc, np); ++k) prepare(k);
  sync_all();
  // a solve that has stopped: its result and true residual, a true-residual restart, or its solution
  // to the host (and a solver context of the batch's own frees its workspace)
  auto stop = [&](int i) -> bool {
    Op& op = *ctx[i];
    const int n = op.subset_sites();
    cmx_solver_result part;
    std::memset(&part, 0, sizeof(part));
    try {
      run[i]->result(part);
    } catch (const Error& e) {
      err[i] = e.code;
      msg[i] = e.what();
    }
    run[i].reset();
    res[i].n_count += part.n_count;
    res[i].n_reliable += part.n_reliable;
    pause();  // the check after the solve, as cmx_solve's, is not solve time
    true_residual(op, op.work(0, P64, n), op.work(1, P64, n), false, res[i]);
    resume();
    if (err[i] == CMX_OK && p.true_residual && res[i].rel_resid > p.rsd_target && res[i].n_restarts < 8 &&
        res[i].n_count < p.max_iter) {
      q[i].rsd_target = q[i].rsd_target * std::max(0.02, 0.5 * p.rsd_target / res[i].rel_resid);
      q[i].max_iter = std::max(1, p.max_iter - res[i].n_count);
      res[i].n_restarts++;
      start(i);
      return false;
    }
    // the solution goes to the host now only when a waiting solve needs this workspace; otherwise after
    // the timed region
    if (!pending.empty()) {
      pause();
      op.unpack(psi_h[i], n, op.work(0, P64, n));
      resume();
      unpacked[i] = 1;
      int next = -1;
      for (int j : pending)
        if (!packed[j] && same_operator(op, *ops[j])) {
          next = j;
          break;
        }
      if (next >= 0) {
        ctx[next] = &op;  // the waiting solve takes this context and its workspace
      } else if (op.shared_stream) {  // a clone made for this batch
        op.ws.clear();
        ++op.ws_gen;
      }
    }
    return true;
  };

  BatchPlan apply_plan, mg_plan;
  resume();
  long ticks = 0;
  while (!pending.empty() || !active.empty()) {
    while ((int)active.size() < conc && !pending.empty()) {
      const int i = pending.front();
      pending.erase(pending.begin());
      start(i);
      active.push_back(i);
    }
    // cycle starts (and solves that stopped before any step)
    std::vector<int> still;
    for (int i : active) {
      bool live = true;
      while (live && !stepping[i]) {
        if (run[i]->finished()) {
          live = !stop(i);
        } else {
          run[i]->begin_cycle();
          stepping[i] = 1;
        }
      }
      if (live) still.push_back(i);
    }
    active = still;
    if (active.empty()) continue;
    // the learned network of all solves on one operator's network and links in one launch per kernel
    std::vector<char> pre(np, 0);
    {
      std::vector<const Vec*> vin;
      std::vector<Vec*> zout;
      std::vector<int> who;
      for (int i : active) {
        const Vec* v = nullptr;
        Vec* z = nullptr;
        if (!run[i]->learned_step_io(v, z)) continue;
        if (!who.empty()) {
          const Op &a = *ctx[who[0]], &b = *ctx[i];
          if (a.gd != b.gd || a.net != b.net || a.ptc != b.ptc || a.stream != b.stream) continue;
        }
        vin.push_back(v);
        zout.push_back(z);
        who.push_back(i);
      }
      if (who.size() >= 2) {
        learned_precond_apply_batch(*ctx[who[0]], p.sloppy_prec, vin, zout);
        for (int i : who) pre[i] = 1;
      }
    }
    // the multigrid V-cycles of the solves that share one multigrid: the smoothing of all of them batched
    {
      std::vector<const Vec*> vin;
      std::vector<Vec*> zout;
      std::vector<Op*> ops;
      std::vector<int> who;
      for (int i : active) {
        const Vec* v = nullptr;
        Vec* z = nullptr;
        if (pre[i] || !run[i]->mg_step_io(v, z)) continue;
        if (!who.empty()) {
          const Op &a = *ctx[who[0]], &b = *ctx[i];
          if (a.mg.get() != b.mg.get() || a.gd != b.gd || a.stream != b.stream ||
              std::memcmp(&a.hcfg[p.sloppy_prec][0], &b.hcfg[p.sloppy_prec][0], sizeof(HopConfig)) != 0)
            continue;
        }
        vin.push_back(v);
        zout.push_back(z);
        ops.push_back(ctx[i]);
        who.push_back(i);
      }
      if (who.size() >= 2) {
        mg_precondition_batch(ops, p.sloppy_prec, vin, zout, mg_plan);
        for (int i : who) pre[i] = 1;
      }
    }
    // the other solves' preconditioners, one solve at a time, before any operator reads z_j
    for (int i : active)
      if (!pre[i]) {
        run[i]->issue_precondition();
        pre[i] = 1;
      }
    // the operator of all solves that share one operator's fields and launch shapes in one batched grid
    // per kernel (the CG batch's plan, its records refreshed every tick)
    std::vector<char> app(np, 0);
    {
      std::vector<int> who;
      std::vector<Recorder> recs;
      for (int i : active) {
        const Vec* z = nullptr;
        Vec* w = nullptr;
        if (!run[i]->apply_step_io(z, w)) continue;
        if (!who.empty()) {
          const Op &a = *ctx[who[0]], &b = *ctx[i];
          if (a.gd != b.gd || a.stream != b.stream || a.Vh != b.Vh || a.par.kind != b.par.kind ||
              std::memcmp(&a.hcfg[p.sloppy_prec][0], &b.hcfg[p.sloppy_prec][0], sizeof(HopConfig)) != 0)
            continue;
        }
        recs.emplace_back();
        {
          RecordScope rs(recs.back());
          ctx[i]->apply(p.sloppy_prec, w->view(), z->view(), +1, nullptr);
        }
        who.push_back(i);
      }
      if (who.size() >= 2) {
        apply_plan.refresh(recs, ctx[who[0]]->stream);
        apply_plan.run(ctx[who[0]]->stream);
        for (int i : who) app[i] = 1;
      }
    }
    for (int i : active) run[i]->issue_step(pre[i] != 0, app[i] != 0);
    sync_all();
    ++ticks;
    for (int i : active)
      if (!run[i]->finish_step()) {
        run[i]->end_cycle();
        stepping[i] = 0;
      }
  }
  pause();
  bres.solve_seconds = timed;
  bres.n_iter = (int)ticks;
  for (int i = 0; i < np; ++i)
    if (!unpacked[i]) ctx[i]->unpack(psi_h[i], ctx[i]->subset_sites(), ctx[i]->work(0, P64, ctx[i]->subset_sites()));
  for (int i = 0; i < np; ++i)
    if (err[i] != CMX_OK) throw Error(err[i], "solve_batch: problem " + std::to_string(i) + ": " + msg[i]);
}

}  // namespace

void solve_batch(const std::vector<Op*>& ops, const cmx_solver_params& p, const cmx_batch_params& bp,
                 double* const* psi_h, const double* const* chi_h, bool mdagm, cmx_solver_result* res,
                 cmx_batch_result& bres) {
  const auto t0 = std::chrono::steady_clock::now();
  const int np = (int)ops.size();
  CMX_REQUIRE(np >= 1, CMX_ERR_ARG, "solve_batch: no problems");
  CMX_REQUIRE(p.max_iter > 0 && p.rsd_target > 0, CMX_ERR_ARG, "solve: MaxIter and RsdTarget must be positive");
  CMX_REQUIRE(p.solver == CMX_SOLVER_CG || p.solver == CMX_SOLVER_RELIABLE_CG || p.solver == CMX_SOLVER_FGMRESDR,
              CMX_ERR_UNSUPPORTED, "solve_batch: CMX_SOLVER_CG, CMX_SOLVER_RELIABLE_CG or CMX_SOLVER_FGMRESDR");
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < i; ++j)
      CMX_REQUIRE(ops[i] != ops[j], CMX_ERR_ARG, "solve_batch: each problem needs its own solver context");
  if (p.solver == CMX_SOLVER_FGMRESDR) {
    CMX_REQUIRE(!mdagm, CMX_ERR_UNSUPPORTED, "FGMRES-DR is a LinOp system solver (no MdagM form)");
    std::memset(&bres, 0, sizeof(bres));
    solve_batch_fgmres(ops, p, bp.max_concurrent, psi_h, chi_h, res, bres);
    bres.total_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < np; ++i) {
      res[i].solve_seconds = bres.solve_seconds;
      res[i].total_seconds = bres.total_seconds;
    }
    return;
  }
  std::memset(&bres, 0, sizeof(bres));
  const bool rel = p.solver == CMX_SOLVER_RELIABLE_CG;
  const int ps = rel ? p.sloppy_prec : P64;
  if (rel) RelCGState::check_prec(ps);
  hipStream_t s0 = ops[0]->stream;
  // the operators' launch shapes are restored on exit
  std::vector<HopConfig> saved(np);
  for (int i = 0; i < np; ++i) saved[i] = ops[i]->hcfg[ps][0];
  struct Restore {
    const std::vector<Op*>& ops;
    const std::vector<HopConfig>& saved;
    int ps;
    ~Restore() {
      for (size_t i = 0; i < ops.size(); ++i) ops[i]->hcfg[ps][0] = saved[i];
    }
  } restore{ops, saved, ps};

  std::vector<Vec*> psi(np), chi(np);
  std::vector<cmx_solver_params> q(np, p);
  for (int i = 0; i < np; ++i) {
    Op& op = *ops[i];
    const int n = op.subset_sites();
    psi[i] = &op.work(0, P64, n);
    chi[i] = &op.work(1, P64, n);
    op.work(2, P64, n);
    op.work(3, P64, n);
    op.pack(chi_h[i], n, *chi[i]);
    op.pack(psi_h[i], n, *psi[i]);
    std::memset(&res[i], 0, sizeof(res[i]));
    // true-residual mode aims the first pass a little below the target, as the single solve does
    if (p.true_residual) q[i].rsd_target = 0.5 * p.rsd_target;
  }

  // launch shape: timed candidates (one batched operator application of every problem), or the
  // operators' own
  if (bp.tune) {
    std::vector<std::unique_ptr<Vec>> va, vb, vc;
    for (Op* o : ops) {
      const int n = o->subset_sites();
      o->ensure_prec(ps);
      va.emplace_back(new Vec(ps, n));
      vb.emplace_back(new Vec(ps, n));
      vc.emplace_back(new Vec(ps, n));
      o->blas(B_ZERO, {va.back().get()});
    }
    for (Op* o : ops) o->sync();
    float best = 1e30f;
    HopConfig bestc{};
    hipEvent_t e0, e1;
    CMX_HIP(hipEventCreate(&e0));
    CMX_HIP(hipEventCreate(&e1));
    for (const HopConfig& c : batch_candidates(ops, ps)) {
      for (Op* o : ops) o->hcfg[ps][0] = c;
      std::vector<Recorder> recs(np);
      for (int i = 0; i < np; ++i) {
        RecordScope rs(recs[i]);
        Op& op = *ops[i];
        op.apply(ps, vb[i]->view(), va[i]->view(), +1, op.partials.as<double>());
        op.apply(ps, vc[i]->view(), vb[i]->view(), -1, nullptr);
      }
      BatchPlan plan;
      plan.build(recs);
      // sustained rate, as in Op::autotune: about 4 ms untimed, then about 4 ms timed
      CMX_HIP(hipEventRecord(e0, s0));
      plan.run(s0);
      CMX_HIP(hipEventRecord(e1, s0));
      CMX_HIP(hipEventSynchronize(e1));
      float t1 = 0;
      CMX_HIP(hipEventElapsedTime(&t1, e0, e1));
      const int nrep = std::max(4, std::min(200, (int)(4.0f / std::max(t1, 1e-3f))));
      for (int k = 0; k < nrep; ++k) plan.run(s0);
      CMX_HIP(hipEventRecord(e0, s0));
      for (int k = 0; k < nrep; ++k) plan.run(s0);
      CMX_HIP(hipEventRecord(e1, s0));
      CMX_HIP(hipEventSynchronize(e1));
      float ms = 0;
      CMX_HIP(hipEventElapsedTime(&ms, e0, e1));
      ms /= (float)nrep;
      if (ms < best) best = ms, bestc = c;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (Op* o : ops) o->hcfg[ps][0] = bestc;
    CMX_HIP(hipStreamSynchronize(s0));
  }
  {
    const HopLaunch h = ops[0]->hop_base(ps);
    bres.hop_block = h.block;
    bres.hop_fast = h.fast;
  }

  // passes: every problem, then (true-residual mode) the problems whose FP64 residual is above target,
  // restarted from their current solution with a tightened inner tolerance
  std::vector<int> act(np);
  for (int i = 0; i < np; ++i) act[i] = i;
  for (int restart = 0; !act.empty(); ++restart) {
    const int na = (int)act.size();
    std::vector<std::unique_ptr<CGState>> cg(na);
    std::vector<std::unique_ptr<RelCGState>> rcg(na);
    for (int k = 0; k < na; ++k) {
      const int i = act[k];
      if (rel) {
        rcg[k].reset(new RelCGState(*ops[i], q[i], *psi[i], *chi[i], mdagm));
        rcg[k]->setup(nullptr);
      } else {
        cg[k].reset(new CGState(*ops[i], q[i], *psi[i], *chi[i], mdagm));
        cg[k]->setup(nullptr);
      }
    }
    for (int i : act) ops[i]->sync();
    // one iteration per problem, recorded (the reliable update in its flag-gated form)
    std::vector<Recorder> recs(na);
    for (int k = 0; k < na; ++k) {
      RecordScope rs(recs[k]);
      if (rel) {
        rcg[k]->inner();
        rcg[k]->update_gated();
      } else {
        cg[k]->iter();
      }
    }
    // replay until every problem has stopped (the device stops each at its MaxIter); the problems
    // that have stopped leave the batched grid at the next check
    std::vector<int> live(na);
    for (int k = 0; k < na; ++k) live[k] = k;
    BatchPlan plan;
    auto rebuild = [&]() {
      std::vector<Recorder> sub;
      sub.reserve(live.size());
      for (int k : live) sub.push_back(recs[k]);
      plan.build(sub);
    };
    rebuild();
    if (restart == 0) {
      bres.n_stages = (int)plan.stages();
      bres.blocks_per_iter = plan.blocks();
    }
    int maxit = 0;
    for (int i : act) maxit = std::max(maxit, q[i].max_iter);
    Timer tm(s0);
    tm.start();
    int batch = p.check_every > 0 ? p.check_every : 4;
    long launched = 0;
    hipGraphExec_t gx = nullptr;
    int gx_len = 0;
    auto drop_graph = [&]() {
      if (gx) (void)hipGraphExecDestroy(gx);
      gx = nullptr;
    };
    try {
      while (true) {
        if (bp.graph && launched > 0) {
          if (!gx || gx_len != batch) {
            drop_graph();
            hipGraph_t gr = nullptr;
            CMX_HIP(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
            for (int j = 0; j < batch; ++j) plan.run(s0);
            CMX_HIP(hipStreamEndCapture(s0, &gr));
            const hipError_t e = hipGraphInstantiate(&gx, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            CMX_HIP(e);
            gx_len = batch;
          }
          CMX_HIP(hipGraphLaunch(gx, s0));
        } else {
          for (int j = 0; j < batch; ++j) plan.run(s0);
        }
        launched += batch;
        for (int k : live) {
          Op* o = ops[act[k]];
          CMX_HIP(hipMemcpyAsync(o->hF, o->state_F.p, sizeof(int) * F_NFLAGS, hipMemcpyDeviceToHost, s0));
        }
        CMX_HIP(hipStreamSynchronize(s0));
        std::vector<int> still;
        for (int k : live)
          if (!ops[act[k]]->hF[F_CONV]) still.push_back(k);
        if (still.empty() || launched > (long)maxit + 8) break;
        if (still.size() < live.size()) {
          live = still;
          drop_graph();
          rebuild();
        }
        if (p.check_every <= 0 && batch < 32) batch *= 2;
      }
    } catch (...) {
      drop_graph();
      throw;
    }
    drop_graph();
    const double secs = tm.stop();
    bres.solve_seconds += secs;
    bres.n_iter += (int)launched;
    // per problem: final update, results, true residual
    std::vector<int> again;
    for (int k = 0; k < na; ++k) {
      const int i = act[k];
      Op& op = *ops[i];
      if (rel) op.blas(B_ADD, {psi[i], &rcg[k]->x});  // psi += x (reliable_cg.cc:158-163)
      read_state(op);
      cmx_solver_result part;
      std::memset(&part, 0, sizeof(part));
      if (rel)
        rcg[k]->finish(secs, part);
      else
        cg[k]->finish(secs, part);
      res[i].n_count += part.n_count;
      res[i].n_reliable += part.n_reliable;
      res[i].gflops += part.gflops * secs;
      true_residual(op, *psi[i], *chi[i], mdagm, res[i]);
      if (p.true_residual && res[i].rel_resid > p.rsd_target && restart < 8 && res[i].n_count < p.max_iter) {
        q[i].rsd_target = q[i].rsd_target * std::max(0.02, 0.5 * p.rsd_target / res[i].rel_resid);
        q[i].max_iter = std::max(1, p.max_iter - res[i].n_count);
        res[i].n_restarts++;
        again.push_back(i);
      }
    }
    act = again;
  }
  for (int i = 0; i < np; ++i) {
    Op& op = *ops[i];
    op.unpack(psi_h[i], op.subset_sites(), *psi[i]);
    res[i].solve_seconds = bres.solve_seconds;
    res[i].gflops = bres.solve_seconds > 0 ? res[i].gflops / bres.solve_seconds : 0.0;
  }
  bres.total_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int i = 0; i < np; ++i) res[i].total_seconds = bres.total_seconds;
}

void solve(Op& op, const cmx_solver_params& p, double* psi_h, const double* chi_h, bool mdagm, cmx_solver_result& res) {
  CMX_REQUIRE(p.max_iter > 0 && p.rsd_target > 0, CMX_ERR_ARG, "solve: MaxIter and RsdTarget must be positive");
  const auto t0 = std::chrono::steady_clock::now();
  const int n = op.subset_sites();
  Vec &psi = op.work(0, P64, n), &chi = op.work(1, P64, n);
  op.pack(chi_h, n, chi);
  op.pack(psi_h, n, psi);
  std::memset(&res, 0, sizeof(res));
  cmx_solver_params q = p;
  // true-residual mode aims the first pass a little below the target, so that a restart is rare
  if (p.true_residual) q.rsd_target = 0.5 * p.rsd_target;
  // the true-residual workspace is allocated before the solver's graphs are captured
  op.work(2, P64, n);
  op.work(3, P64, n);
  for (int restart = 0;; ++restart) {
    cmx_solver_result part;
    std::memset(&part, 0, sizeof(part));
    solve_pass(op, q, psi, chi, mdagm, part);
    res.n_count += part.n_count;
    res.n_reliable += part.n_reliable;
    res.solve_seconds += part.solve_seconds;
    res.gflops += part.gflops * part.solve_seconds;
    true_residual(op, psi, chi, mdagm, res);
    if (!p.true_residual || res.rel_resid <= p.rsd_target || restart >= 8 || res.n_count >= p.max_iter) break;
    // restart from psi with the inner tolerance scaled by the observed true/inner ratio
    q.rsd_target = q.rsd_target * std::max(0.02, 0.5 * p.rsd_target / res.rel_resid);
    q.max_iter = std::max(1, p.max_iter - res.n_count);
    res.n_restarts++;
  }
  res.gflops = res.solve_seconds > 0 ? res.gflops / res.solve_seconds : 0.0;
  op.unpack(psi_h, n, psi);
  res.total_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (p.verbose)
    std::printf("CMX_SOLVER: %d iterations. Rsd = %g Relative Rsd = %g  (%.3f s, %.1f GFLOP/s)\n", res.n_count,
                res.resid, res.rel_resid, res.solve_seconds, res.gflops);
}

}  // namespace cmx
