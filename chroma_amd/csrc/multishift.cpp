// Multishift CG (MInvCG2_a, minvcg2.cc:72-355; MdagMMultiSysSolverCG, multi_syssolver_mdagm_cg.h:58-100):
// (M^dag M + shifts[s]) psi[s] = chi for every shift with one Krylov space (Jegerlehner,
// hep-lat/9708029), in FP64 on the device.  The vectors live on the device; the scalar recurrences of
// the shifted systems (z, bs, as) are a few flops per shift and run on the host in double, as the
// reference runs them, from the two reductions each iteration reads back (d = |M p|^2, c = |r|^2).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "blas.h"
#include "op.h"

namespace cmx {

namespace {

constexpr int MS_MAX = 64;             // shifts per solve
constexpr int MS_D = S_KRY;            // device-written: d = |M p0|^2
constexpr int MS_C = S_KRY + 1;        // device-written: c = |r|^2
constexpr int MS_B = S_KRY + 2;        // host-written: b
constexpr int MS_SIG = S_KRY + 3;      // host-written: the shift of the residual check
constexpr int MS_X = S_KRY + 8;        // host-written: a, then (z_s, as_s) for every shift
constexpr int MS_NBS = MS_X + 1 + 2 * MS_MAX;  // host-written: -bs_s
static_assert(MS_NBS + MS_MAX <= S_MR, "multishift slots overlap the MR slots");

void put(Op& op, int slot, const double* v, int n) {
  std::memcpy(op.hS + slot, v, sizeof(double) * n);
  CMX_HIP(hipMemcpyAsync(static_cast<double*>(op.state_S.p) + slot, op.hS + slot, sizeof(double) * n,
                         hipMemcpyHostToDevice, op.stream));
}
double get(Op& op, int slot) {
  CMX_HIP(hipMemcpyAsync(op.hS + slot, static_cast<double*>(op.state_S.p) + slot, sizeof(double),
                         hipMemcpyDeviceToHost, op.stream));
  op.sync();
  return op.hS[slot];
}

}  // namespace

void solve_multishift(Op& op, const cmx_solver_params& p, int ns, const double* shifts, const double* rsd,
                      double* psi_h, const double* chi_h, cmx_solver_result* res) {
  CMX_REQUIRE(ns >= 1 && ns <= MS_MAX, CMX_ERR_ARG, "solve_multishift: 1.." + std::to_string(MS_MAX) + " shifts");
  CMX_REQUIRE(p.max_iter > 0, CMX_ERR_ARG, "solve_multishift: MaxIter must be positive");
  for (int s = 0; s < ns; ++s)
    CMX_REQUIRE(rsd[s] > 0 && std::isfinite(shifts[s]), CMX_ERR_ARG, "solve_multishift: bad shift or RsdCG");
  const auto t0 = std::chrono::steady_clock::now();
  const int n = op.subset_sites();
  const int nb = blas_blocks(n);
  Vec &chi = op.work(1, P64, n), &r = op.work(11, P64, n), &p0 = op.work(12, P64, n), &mp = op.work(13, P64, n),
      &mmp = op.work(14, P64, n);
  std::vector<std::unique_ptr<Vec>> psi, pv;
  for (int s = 0; s < ns; ++s) {
    psi.emplace_back(new Vec(P64, n));
    pv.emplace_back(new Vec(P64, n));
  }
  op.pack(chi_h, n, chi);
  hipEvent_t e0, e1;
  CMX_HIP(hipEventCreate(&e0));
  CMX_HIP(hipEventCreate(&e1));
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } evg{e0, e1};
  CMX_HIP(hipEventRecord(e0, op.stream));

  // isz: the smallest shift decides the first convergence test (:99-104, :247)
  int isz = 0;
  for (int s = 1; s < ns; ++s)
    if (shifts[s] < shifts[isz]) isz = s;
  for (int s = 0; s < ns; ++s) op.blas(B_ZERO, {psi[s].get()});  // (:118-121)
  op.blas(B_NORM2, {&chi});
  op.finalize(nb, 1, PG_STORE, MS_C);
  const double chi_norm_sq = get(op, MS_C);
  int n_count = 0;
  if (std::sqrt(chi_norm_sq) >= 1e-10) {  // a zero source returns the zero solutions (:133-153)
    std::vector<double> rsd_sq(ns), bs(ns), z[2] = {std::vector<double>(ns), std::vector<double>(ns)};
    for (int s = 0; s < ns; ++s) rsd_sq[s] = chi_norm_sq * rsd[s] * rsd[s];
    double cp = chi_norm_sq;
    // r = p0 = p_s = chi (:161-175)
    op.blas(B_COPY, {&r, &chi});
    op.blas(B_COPY, {&p0, &chi});
    for (int s = 0; s < ns; ++s) op.blas(B_COPY, {pv[s].get(), &chi});
    // b = -|r|^2 / |M p0|^2 ; r += b M^dag M p0 (:179-194)
    int np = op.apply(P64, mp.view(), p0.view(), +1, op.partials.as<double>());
    op.finalize(np, 1, PG_STORE, MS_D);
    op.apply(P64, mmp.view(), mp.view(), -1, nullptr);
    double d = get(op, MS_D);
    CMX_REQUIRE(d > 0 && std::isfinite(d), CMX_ERR_BREAKDOWN, "multishift CG: |M p| = 0 or not finite");
    double b = -cp / d;
    put(op, MS_B, &b, 1);
    op.blas(B_AXPY_NORM, {&r, &mmp}, MS_B);
    op.finalize(nb, 1, PG_STORE, MS_C);
    // shifted bs and z (:197-217); psi_s = -bs_s chi (:221-225)
    int iz = 1;
    std::vector<double> nbs(ns);
    for (int s = 0; s < ns; ++s) {
      z[1 - iz][s] = 1.0;
      z[iz][s] = 1.0 / (1.0 - shifts[s] * b);
      bs[s] = b * z[iz][s];
      nbs[s] = -bs[s];
    }
    put(op, MS_NBS, nbs.data(), ns);
    for (int s = 0; s < ns; ++s) op.blas(B_SCALE, {psi[s].get(), &chi}, MS_NBS + s);
    double c = get(op, MS_C);
    std::vector<char> conv(ns, 0);
    bool convP = c < rsd_sq[isz];
    std::vector<double> x(1 + 2 * ns);
    for (int k = 1; k <= p.max_iter && !convP; ++k) {
      // p0 = r + a p0 ; p_s = z_s r + as_s p_s (:253-273)
      const double a = c / cp;
      x[0] = a;
      for (int s = 0; s < ns; ++s) {
        x[1 + 2 * s] = z[iz][s];
        x[2 + 2 * s] = conv[s] ? 0.0 : a * z[iz][s] * bs[s] / (z[1 - iz][s] * b);
      }
      put(op, MS_X, x.data(), 1 + 2 * ns);
      op.blas(B_XPAY, {&p0, &r}, MS_X);
      for (int s = 0; s < ns; ++s)
        if (!conv[s]) op.blas(B_AXPBY, {pv[s].get(), &r}, MS_X + 1 + 2 * s, MS_X + 2 + 2 * s);
      cp = c;
      // b = -cp / |M p0|^2 ; r += b M^dag M p0 ; c = |r|^2 (:275-294)
      np = op.apply(P64, mp.view(), p0.view(), +1, op.partials.as<double>());
      op.finalize(np, 1, PG_STORE, MS_D);
      op.apply(P64, mmp.view(), mp.view(), -1, nullptr);
      d = get(op, MS_D);
      CMX_REQUIRE(d > 0 && std::isfinite(d), CMX_ERR_BREAKDOWN, "multishift CG: |M p| = 0 or not finite");
      const double bp = b;
      b = -cp / d;
      put(op, MS_B, &b, 1);
      op.blas(B_AXPY_NORM, {&r, &mmp}, MS_B);
      op.finalize(nb, 1, PG_STORE, MS_C);
      c = get(op, MS_C);
      // shifted z and bs (:297-308) ; psi_s -= bs_s p_s (:312-322)
      iz = 1 - iz;
      for (int s = 0; s < ns; ++s)
        if (!conv[s]) {
          const double z0 = z[1 - iz][s], z1 = z[iz][s];
          z[iz][s] = z0 * z1 * bp / (b * a * (z1 - z0) + z1 * bp * (1.0 - shifts[s] * b));
          bs[s] = b * z[iz][s] / z0;
          nbs[s] = -bs[s];
        }
      put(op, MS_NBS, nbs.data(), ns);
      for (int s = 0; s < ns; ++s)
        if (!conv[s]) op.blas(B_AXPY, {psi[s].get(), pv[s].get()}, MS_NBS + s);
      // per-shift test on the shifted residual estimate c z_s^2 (:326-339)
      convP = true;
      for (int s = 0; s < ns; ++s) {
        if (!conv[s]) conv[s] = c * z[iz][s] * z[iz][s] < rsd_sq[s];
        convP = convP && conv[s];
      }
      n_count = k;
      if (p.verbose) std::printf("MInvCG2: k = %d |r| = %g\n", k, std::sqrt(c));
    }
  }
  CMX_HIP(hipEventRecord(e1, op.stream));
  CMX_HIP(hipEventSynchronize(e1));
  float ms = 0;
  CMX_HIP(hipEventElapsedTime(&ms, e0, e1));
  // FP64 residual of every shifted system: |chi - (M^dag M + shift_s) psi_s|
  for (int s = 0; s < ns; ++s) {
    op.apply(P64, mp.view(), psi[s]->view(), +1, nullptr);
    op.apply(P64, mmp.view(), mp.view(), -1, nullptr);
    put(op, MS_SIG, &shifts[s], 1);
    op.blas(B_AXPY, {&mmp, psi[s].get()}, MS_SIG);
    op.blas(B_XMY_NORM, {&r, &chi, &mmp});
    op.finalize(nb, 1, PG_STORE, MS_C);
    const double rr = get(op, MS_C);
    cmx_solver_result& q = res[s];
    std::memset(&q, 0, sizeof(q));
    q.n_count = n_count;
    q.resid = std::sqrt(rr);
    q.rel_resid = chi_norm_sq > 0 ? q.resid / std::sqrt(chi_norm_sq) : q.resid;
    q.solve_seconds = ms * 1e-3;
    op.unpack(psi_h + (size_t)s * 24 * n, n, *psi[s]);
  }
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int s = 0; s < ns; ++s) res[s].total_seconds = wall;
  if (n_count >= p.max_iter)  // QDP_error_exit("too many CG iterations") (:357-359)
    throw Error(CMX_ERR_NOCONV, "multishift CG: no convergence in " + std::to_string(p.max_iter) + " iterations");
}

}  // namespace cmx
