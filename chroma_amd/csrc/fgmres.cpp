// FGMRES-DR: flexible (right-preconditioned) GMRES with deflated restarts, after
// LinOpSysSolverFGMRESDR (syssolver_linop_fgmres_dr.cc:78-774, Givens rotations in
// syssolver_linop_fgmres_dr.h:32-140).
//
// Device side: the Krylov basis V and the preconditioned basis Z live in the sloppy precision; one
// Arnoldi step is the preconditioner, one operator application, one pass of all inner products
// <V_i, w>, one pass w -= sum_i h_i V_i with |w|^2 (classical Gram-Schmidt, optionally repeated),
// and the normalisation.  Host side: the small Hessenberg / Givens / least-squares / deflation
// algebra (one read-back of the new column per step).  After every cycle the residual is recomputed
// in FP64 (as the reference does, :740-745), which makes the sloppy basis safe.
//
// Differences from the reference, each deliberate:
//  * deflated cycles apply only the Givens rotations of the current cycle to a new column (the
//    reference also applies the previous cycle's rotations of the deflated columns, :151-153, which
//    does not match the Q_k^H transformation it has already applied to those rows);
//  * each solve starts with Z = 0 (the reference keeps Z from the previous solve as the
//    preconditioner's initial guesses, because it zeroes Z only at construction, :197-225);
//  * the identity preconditioner is offered (the reference needs a factory solver);
//  * a deflated cycle that fails to halve the true residual is followed by a plain restart (in FP64
//    arithmetic this does not trigger: the oracle's cycles are reproduced exactly).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>

#include "dense.h"
#include "fgmres.h"
#include "mg.h"
#include "op.h"

namespace cmx {

int kry_blocks(int n);

namespace {

using dense::cd;
using dense::Mat;

// Givens rotation with a real, non-negative sine (syssolver_linop_fgmres_dr.h:32-140)
struct Givens {
  int col = 0;
  cd c = 1, s = 0, r = 0;
  Givens() = default;
  Givens(int col_, cd f, cd g) : col(col_) {
    if (f == cd(0)) {
      c = 0;
      s = 1;
      r = g;
    } else if (g == cd(0)) {
      s = 0;
      c = std::conj(f) / std::abs(f);
      r = std::abs(f);
    } else {
      const double t = std::sqrt(std::norm(f) + std::norm(g));
      r = t;
      c = f / t;
      s = g / t;
    }
  }
  // rows (col, col+1) of a column vector
  void apply(std::vector<cd>& v) const {
    const cd a = v[col], b = v[col + 1];
    v[col] = std::conj(c) * a + std::conj(s) * b;
    v[col + 1] = -s * a + c * b;
  }
};

// Device helpers over a basis of sloppy vectors
struct KryDev {
  Op& op;
  int n;
  DevBuf& part;  // reduction partials, kept on the operator across solves
  static DevBuf& partials_of(Op& o) {
    if (!o.kry_part.p) o.kry_part.alloc(sizeof(double) * KRY_BLOCKS * 2 * KRY_MAXV);
    return o.kry_part;
  }
  explicit KryDev(Op& o, int n_) : op(o), n(n_), part(partials_of(o)) {}
  MultiLaunch base(const Vec& w, const std::vector<Vec*>& vs, int m) const {
    MultiLaunch L;
    std::memset(&L, 0, sizeof(L));
    L.n = n;
    L.m = m;
    L.w = w.view();
    for (int i = 0; i < m; ++i) L.v[i] = vs[i]->view();
    L.S = op.state_S.as<double>();
    L.partials = part.as<double>();
    return L;
  }
  // S[slot + 2i, 2i+1] = <vs[i], w>, i < m
  void mdot(const std::vector<Vec*>& vs, int m, const Vec& w, int slot) {
    MultiLaunch L = base(w, vs, m);
    launch_mdot(L, vs[0]->prec, w.prec, op.stream);
    launch_colsum(part.as<double>(), kry_blocks(n), 2 * m, op.state_S.as<double>() + slot, 0, op.stream);
  }
  // w (+)= sgn sum_i S[slot+2i] vs[i]; if nslot >= 0: S[nslot] = |w|^2, S[nslot+1] = 1/|w|
  void maxpy(Vec& w, const std::vector<Vec*>& vs, int m, int slot, double sgn, bool overwrite, int nslot) {
    MultiLaunch L = base(w, vs, m);
    L.coef = slot;
    L.sgn = sgn;
    L.overwrite = overwrite ? 1 : 0;
    launch_maxpy(L, vs[0]->prec, w.prec, nslot >= 0, op.stream);
    if (nslot >= 0) launch_colsum(part.as<double>(), kry_blocks(n), 1, op.state_S.as<double>() + nslot, 1, op.stream);
  }
  // The copy reads the pinned host mirror when the stream reaches it, so the mirror must not be
  // rewritten before then: put() waits for the copy (callers issue several puts back to back).
  void put(int slot, const std::vector<cd>& c) {
    for (size_t i = 0; i < c.size(); ++i) {
      op.hS[slot + 2 * i] = c[i].real();
      op.hS[slot + 2 * i + 1] = c[i].imag();
    }
    CMX_HIP(hipMemcpyAsync(static_cast<double*>(op.state_S.p) + slot, op.hS + slot, sizeof(double) * 2 * c.size(),
                           hipMemcpyHostToDevice, op.stream));
    op.sync();
  }
  void put1(int slot, double v) {
    op.hS[slot] = v;
    CMX_HIP(hipMemcpyAsync(static_cast<double*>(op.state_S.p) + slot, op.hS + slot, sizeof(double),
                           hipMemcpyHostToDevice, op.stream));
    op.sync();
  }
  void get(int slot, int count) {
    get_async(slot, count);
    op.sync();
  }
  // the copy only: the values are in op.hS once the stream has reached it
  void get_async(int slot, int count) {
    CMX_HIP(hipMemcpyAsync(op.hS + slot, static_cast<double*>(op.state_S.p) + slot, sizeof(double) * count,
                           hipMemcpyDeviceToHost, op.stream));
  }
};

// Krylov-area layout of the scalar array
constexpr int K_H = S_KRY;                 // new Hessenberg column (2 * KRY_MAXV)
constexpr int K_H2 = K_H + 2 * KRY_MAXV;   // second Gram-Schmidt pass
constexpr int K_WN = K_H2 + 2 * KRY_MAXV;  // |w|^2, 1/|w|
constexpr int K_C = K_WN + 2;              // host-supplied coefficients (2 * KRY_MAXV)
static_assert(K_C + 2 * KRY_MAXV <= S_MR, "Krylov slots overlap the MR slots");


}  // namespace

struct FgmresDR::Impl {
  Op& op;
  const cmx_solver_params p;
  Vec& psi;
  const Vec& chi;
  int n, ps, pb, nk, nd_max, total, reorth, nb;
  // the bases, the spare vectors of the deflated restart and the scratch vectors live in the
  // operator's workspace (slots 100-306), so repeated solves allocate nothing; V and Z map basis
  // indices to those vectors and are permuted, not copied, when the deflated basis is formed
  std::vector<Vec*> V, Z, spareV, spareZ, vs;
  Vec *r = nullptr, *tmp = nullptr;
  Vec* rs = nullptr;                  // the FP64 residual in the basis precision (deflated restarts)
  Vec *xs = nullptr, *ys = nullptr;   // operator-precision staging when pb != ps
  KryDev kd;
  std::unique_ptr<MRPrecond> mr;
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  double norm_rhs = 0, target = 0, r_norm = 0, r_prev = 0;
  Mat H, R, Qk;  // Qk: Q of the deflated block (nd+1 x nd+1)
  std::vector<cd> g, c, eta;
  std::vector<Givens> rot;
  int iters = 0, cycles = 0, prev_dim = 0;
  bool done = false;
  bool plain_next = true;  // the first cycle, and any cycle after a deflated one that stalled
  // the cycle in progress
  int nd = 0, j = 0, dim = 0;
  double cycle_stop = 0;

  Impl(Op& o, const cmx_solver_params& p_, Vec& psi_, const Vec& chi_)
      : op(o), p(p_), psi(psi_), chi(chi_), n(o.subset_sites()), ps(p_.sloppy_prec), kd(o, o.subset_sites()) {
    CMX_REQUIRE(ps == P64 || ps == P32 || ps == P16, CMX_ERR_ARG, "FGMRES-DR: bad sloppy precision");
    nk = p.n_krylov, nd_max = p.n_defl;
    CMX_REQUIRE(nk >= 1 && nd_max >= 0 && nd_max < nk && nk + nd_max + 1 <= KRY_MAXV, CMX_ERR_ARG,
                "FGMRES-DR: need 1 <= NKrylov, 0 <= NDefl < NKrylov, NKrylov + NDefl < 48");
    CMX_REQUIRE(p.precond >= CMX_PRECOND_NONE && p.precond <= CMX_PRECOND_MG, CMX_ERR_ARG,
                "FGMRES-DR: unknown preconditioner");
    // basis precision pb (V, Z) may be wider than the precision ps of the operator and preconditioner
    pb = p.basis_prec < 0 ? ps : p.basis_prec;
    CMX_REQUIRE(pb == P64 || pb == P32 || pb == P16, CMX_ERR_ARG, "FGMRES-DR: bad basis precision");
    total = nk + nd_max;
    prev_dim = nk;
    V.resize(total + 1), Z.resize(total), spareV.resize(nd_max + 1), spareZ.resize(nd_max), vs.resize(total + 1);
    for (int i = 0; i <= total; ++i) V[i] = &op.work(100 + i, pb, n);
    for (int i = 0; i < total; ++i) Z[i] = &op.work(150 + i, pb, n);
    for (int i = 0; i <= nd_max; ++i) spareV[i] = &op.work(200 + i, pb, n);
    for (int i = 0; i < nd_max; ++i) spareZ[i] = &op.work(250 + i, pb, n);
    for (Vec* z : Z) op.blas(B_ZERO, {z});
    r = &op.work(300, P64, n);
    tmp = &op.work(301, P64, n);
    rs = nd_max > 0 ? (pb == P64 ? r : &op.work(306, pb, n)) : nullptr;
    if (pb != ps) {
      xs = &op.work(302, ps, n);
      ys = &op.work(303, ps, n);
    }
    if (p.precond == CMX_PRECOND_MR)
      mr.reset(new MRPrecond(op, ps, n, std::max(0, p.precond_iters), p.precond_rsd,
                             p.precond_omega != 0 ? p.precond_omega : 1.0, op.work(304, ps, n), op.work(305, ps, n)));
    reorth = p.reorth >= 2 ? 2 : 1;
    nb = blas_blocks(n);
    CMX_HIP(hipEventCreate(&ev_a));
    CMX_HIP(hipEventCreate(&ev_b));
    CMX_HIP(hipEventRecord(ev_a, op.stream));

    // |chi|, target (:463-465); r = chi - A psi, |r| (:468-475)
    op.blas(B_NORM2, {&chi});
    op.finalize(nb, 1, PG_STORE, K_WN);
    kd.get(K_WN, 1);
    norm_rhs = std::sqrt(op.hS[K_WN]);
    target = norm_rhs * p.rsd_target;
    // a zero initial guess (the usual case) needs no operator application for the first residual
    op.blas(B_NORM2, {&psi});
    op.finalize(nb, 1, PG_STORE, K_WN);
    kd.get(K_WN, 1);
    r_norm = true_residual(op.hS[K_WN] == 0.0);
    r_prev = r_norm;
    H = Mat(total + 1, total);
    R = Mat(total + 1, total);
    g.assign(total + 1, cd(0));
    c.assign(total + 1, cd(0));
    eta.assign(total, cd(0));
    rot.assign(total, Givens());
    done = r_norm <= target;
  }
  ~Impl() {
    if (ev_a) (void)hipEventDestroy(ev_a);
    if (ev_b) (void)hipEventDestroy(ev_b);
  }

  double true_residual(bool psi_zero = false) {
    if (psi_zero)  // A 0 = 0 exactly: the same r without the FP64 operator
      op.blas(B_ZERO, {tmp});
    else
      op.apply(P64, tmp->view(), psi.view(), +1, nullptr);
    op.blas(B_XMY_NORM, {r, &chi, tmp});
    op.finalize(nb, 1, PG_STORE, K_WN);
    kd.get(K_WN, 1);
    return std::sqrt(op.hS[K_WN]);
  }
  void precondition_s(Vec& z, const Vec& v) {  // both in precision ps
    switch (p.precond) {
      case CMX_PRECOND_NONE: op.blas(B_COPY, {&z, &v}); break;
      case CMX_PRECOND_MR: mr->apply(z, v); break;
      case CMX_PRECOND_MG: mg_precondition(op, ps, v, z); break;
      default: learned_precond_apply(op, ps, v, z); break;
    }
  }
  void precondition(Vec& z, const Vec& v) {  // z, v in the basis precision
    if (pb == ps || p.precond == CMX_PRECOND_NONE) {
      if (pb == ps) precondition_s(z, v);
      else op.blas(B_COPY, {&z, &v});
      return;
    }
    op.blas(B_COPY, {xs, &v});
    if (p.precond == CMX_PRECOND_MR) op.blas(B_COPY, {ys, &z});  // MR starts from z
    precondition_s(*ys, *xs);
    op.blas(B_COPY, {&z, ys});
  }
  void apply_a(Vec& w, const Vec& z) {  // w = A z at the operator precision
    if (pb == ps) {
      op.apply(ps, w.view(), z.view(), +1, nullptr);
      return;
    }
    op.blas(B_COPY, {xs, &z});
    op.apply(ps, ys->view(), xs->view(), +1, nullptr);
    op.blas(B_COPY, {&w, ys});
  }

  void begin_cycle() {
    ++cycles;
    // a cycle that ended early (cycle_reduction, a converged estimate, an exact breakdown) may have
    // fewer than NDefl Arnoldi vectors: deflate at most that many harmonic Ritz vectors
    nd = std::min(nd_max, prev_dim);
    if (plain_next || nd == 0) {
      // c = g = [|r|, 0, ...]; V_0 = r / |r|  (:491-513)
      std::fill(c.begin(), c.end(), cd(0));
      std::fill(g.begin(), g.end(), cd(0));
      c[0] = g[0] = r_norm;
      H = Mat(total + 1, total);
      R = Mat(total + 1, total);
      kd.put1(K_WN + 1, 1.0 / r_norm);
      op.blas(B_SCALE, {V[0], r}, K_WN + 1);
      nd = 0;
    } else {
      deflated_restart();
    }
    // flexible Arnoldi (:78-183); with cycle_reduction > 0 the cycle also ends once its estimate has
    // fallen by that factor below the residual it started from
    cycle_stop = std::max(target, p.cycle_reduction > 0 ? p.cycle_reduction * r_norm : 0.0);
    dim = nd;
    j = nd;
  }

  // deflated restart (:514-688): harmonic Ritz vectors of the previous cycle
  void deflated_restart() {
    const int m = prev_dim;
    std::vector<cd> hm(m), f(m);
    Mat Hm(m, m);
    for (int i = 0; i < m; ++i)
      for (int k = 0; k < m; ++k) Hm(i, k) = H(i, k);
    for (int k = 0; k < m; ++k) hm[k] = std::conj(H(m, k));
    f = dense::lu_solve(dense::adjoint(Hm), hm);  // H_m^H f = h_m
    Mat Ht(m, m);
    for (int i = 0; i < m; ++i)
      for (int k = 0; k < m; ++k) Ht(i, k) = H(i, k) + f[i] * std::conj(hm[k]);
    std::vector<cd> evals;
    Mat evecs;
    dense::eig(Ht, evals, evecs);
    std::vector<int> order(m);
    for (int i = 0; i < m; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return std::norm(evals[a]) < std::norm(evals[b]); });
    // G_{k+1} = [ [G_k; 0] | c - H eta ]  ((m+1) x (nd+1)), then Q_{k+1} from its QR
    Mat G(m + 1, nd + 1);
    for (int col = 0; col < nd; ++col)
      for (int row = 0; row < m; ++row) G(row, col) = evecs(row, order[col]);
    for (int row = 0; row <= m; ++row) {
      cd s = c[row];
      for (int col = 0; col < m; ++col) s -= H(row, col) * eta[col];
      G(row, nd) = s;
    }
    Mat Q, Rq;
    dense::qr(G, Q, Rq);
    // H_new = Q_{k+1}^H H Q_k   ((nd+1) x nd)
    Mat Hfull(m + 1, m), Qk_(m, nd);
    for (int i = 0; i <= m; ++i)
      for (int k = 0; k < m; ++k) Hfull(i, k) = H(i, k);
    for (int i = 0; i < m; ++i)
      for (int k = 0; k < nd; ++k) Qk_(i, k) = Q(i, k);
    Mat Hn = dense::matmul(dense::adjoint(Q), dense::matmul(Hfull, Qk_));
    // new V_i = sum_k V_k Q(k, i) (i <= nd); new Z_i = sum_k Z_k Q(k, i) (i < nd), formed in the
    // spare vectors (basis precision), which then take the places of V_i, Z_i
    for (int k = 0; k <= m; ++k) vs[k] = V[k];
    for (int i = 0; i <= nd; ++i) {
      std::vector<cd> col(m + 1);
      for (int k = 0; k <= m; ++k) col[k] = Q(k, i);
      kd.put(K_C, col);
      kd.maxpy(*spareV[i], vs, m + 1, K_C, 1.0, true, -1);
    }
    for (int k = 0; k < m; ++k) vs[k] = Z[k];
    for (int i = 0; i < nd; ++i) {
      std::vector<cd> col(m);
      for (int k = 0; k < m; ++k) col[k] = Q(k, i);
      kd.put(K_C, col);
      kd.maxpy(*spareZ[i], vs, m, K_C, 1.0, true, -1);
    }
    for (int i = 0; i <= nd; ++i) std::swap(V[i], spareV[i]);
    for (int i = 0; i < nd; ++i) std::swap(Z[i], spareZ[i]);
    for (int i = nd; i < total; ++i) op.blas(B_ZERO, {Z[i]});
    // reset H, R, c, g; c = V_{k+1}^H r (:640-650)
    H = Mat(total + 1, total);
    R = Mat(total + 1, total);
    std::fill(c.begin(), c.end(), cd(0));
    std::fill(g.begin(), g.end(), cd(0));
    {
      // <V_i, r> with r in FP64: convert r once to the basis precision
      if (rs != r) op.blas(B_COPY, {rs, r});
      for (int i = 0; i <= nd; ++i) vs[i] = V[i];
      kd.mdot(vs, nd + 1, *rs, K_H);
      kd.get(K_H, 2 * (nd + 1));
      for (int i = 0; i <= nd; ++i) c[i] = g[i] = cd(op.hS[K_H + 2 * i], op.hS[K_H + 2 * i + 1]);
    }
    for (int i = 0; i <= nd; ++i)
      for (int k = 0; k < nd; ++k) H(i, k) = Hn(i, k);
    // QR of H_k ((nd+1) x nd): R block and g = Q_k^H c (:652-670)
    Mat Hk(nd + 1, nd);
    for (int i = 0; i <= nd; ++i)
      for (int k = 0; k < nd; ++k) Hk(i, k) = Hn(i, k);
    Mat Rk;
    dense::qr(Hk, Qk, Rk, true);
    for (int i = 0; i < nd; ++i)
      for (int k = i; k < nd; ++k) R(i, k) = Rk(i, k);
    std::vector<cd> gq(nd + 1, cd(0));
    for (int i = 0; i <= nd; ++i)
      for (int k = 0; k <= nd; ++k) gq[i] += std::conj(Qk(k, i)) * g[k];
    for (int i = 0; i <= nd; ++i) g[i] = gq[i];
  }

  bool learned_step_io(const Vec*& v, Vec*& z) const {
    if (p.precond != CMX_PRECOND_LEARNED || pb != ps) return false;
    v = V[j];
    z = Z[j];
    return true;
  }

  bool mg_step_io(const Vec*& v, Vec*& z) const {
    if (p.precond != CMX_PRECOND_MG || pb != ps) return false;
    v = V[j];
    z = Z[j];
    return true;
  }

  bool apply_step_io(const Vec*& z, Vec*& w) const {
    if (pb != ps) return false;
    z = Z[j];
    w = V[j + 1];
    return true;
  }

  void issue_step(bool preconditioned, bool applied) {
    if (!preconditioned) precondition(*Z[j], *V[j]);  // z_j = M v_j
    if (!applied) apply_a(*V[j + 1], *Z[j]);          // w = A z_j
    for (int i = 0; i <= j; ++i) vs[i] = V[i];
    kd.mdot(vs, j + 1, *V[j + 1], K_H);  // h_i = <v_i, w>
    kd.maxpy(*V[j + 1], vs, j + 1, K_H, -1.0, false, reorth == 1 ? K_WN : -1);
    if (reorth == 2) {
      kd.mdot(vs, j + 1, *V[j + 1], K_H2);
      kd.maxpy(*V[j + 1], vs, j + 1, K_H2, -1.0, false, K_WN);
    }
    kd.get_async(K_H, K_WN + 2 - K_H);  // the new column (both passes), |w|^2 and 1/|w| in one read-back
  }

  bool finish_step() {
    for (int i = 0; i <= j; ++i) {
      H(i, j) = cd(op.hS[K_H + 2 * i], op.hS[K_H + 2 * i + 1]);
      if (reorth == 2) H(i, j) += cd(op.hS[K_H2 + 2 * i], op.hS[K_H2 + 2 * i + 1]);
    }
    const double wnorm = std::sqrt(op.hS[K_WN]);
    H(j + 1, j) = wnorm;
    if (std::fabs(wnorm) < 1e-14) {  // exact convergence (:117-127)
      dim = j;
      return false;
    }
    op.blas(B_SCALE, {V[j + 1], V[j + 1]}, K_WN + 1);  // v_{j+1} = w / |w|
    std::vector<cd> col(j + 2);
    for (int i = 0; i <= j + 1; ++i) col[i] = H(i, j);
    if (nd > 0) {  // R_col <- Q_k^H R_col on the deflated rows (:133-152)
      std::vector<cd> t(nd + 1, cd(0));
      for (int i = 0; i <= nd; ++i)
        for (int k = 0; k <= nd; ++k) t[i] += std::conj(Qk(k, i)) * col[k];
      for (int i = 0; i <= nd; ++i) col[i] = t[i];
    }
    for (int i = nd; i < j; ++i) rot[i].apply(col);
    rot[j] = Givens(j, col[j], col[j + 1]);
    col[j] = rot[j].r;
    col[j + 1] = 0;
    rot[j].apply(g);
    for (int i = 0; i <= j + 1; ++i) R(i, j) = col[i];
    dim = j + 1;
    const double accum = std::fabs(g[j + 1].real());  // (:172)
    if (p.verbose) std::printf("FGMRESDR: iter %d |r| ~ %g target %g\n", iters + j - nd + 1, accum, target);
    if (accum <= cycle_stop) return false;
    ++j;
    return j < nd + nk;
  }

  void end_cycle() {
    // eta = R^-1 g by back substitution (:240-262); psi += sum eta_j Z_j (:716-724)
    if (dim > 0) {
      for (int row = dim - 1; row >= 0; --row) {
        cd s = g[row];
        for (int col = row + 1; col < dim; ++col) s -= R(row, col) * eta[col];
        eta[row] = s / R(row, row);
      }
      for (int k = 0; k < dim; ++k) vs[k] = Z[k];
      kd.put(K_C, std::vector<cd>(eta.begin(), eta.begin() + dim));
      kd.maxpy(psi, vs, dim, K_C, 1.0, false, -1);
    }
    r_norm = true_residual();  // (:727-736)
    iters += dim - nd;
    // With a basis narrower than FP64 the true residual is not exactly in the span of the deflated
    // basis, and a deflated cycle can stall on the part outside it; the next cycle then restarts
    // from the true residual (a plain FGMRES cycle).
    plain_next = nd > 0 && r_norm > 0.5 * r_prev;
    r_prev = r_norm;
    done = r_norm <= target || iters >= p.max_iter || (dim == nd && !plain_next);
    prev_dim = dim;
    if (p.verbose)
      std::printf("FGMRESDR: cycle %d: %d iterations, |r| = %g target %g\n", cycles, dim - nd, r_norm, target);
  }

  void result(cmx_solver_result& res) {
    CMX_HIP(hipEventRecord(ev_b, op.stream));
    CMX_HIP(hipEventSynchronize(ev_b));
    float ms = 0;
    CMX_HIP(hipEventElapsedTime(&ms, ev_a, ev_b));
    res.n_count = iters;
    res.resid = r_norm;
    res.rel_resid = norm_rhs > 0 ? r_norm / norm_rhs : r_norm;
    res.n_reliable = cycles;
    res.solve_seconds = ms * 1e-3;
    res.gflops = 0;
    if (r_norm > target)
      throw Error(CMX_ERR_NOCONV, "FGMRES-DR: no convergence in " + std::to_string(p.max_iter) + " iterations");
  }
};

FgmresDR::FgmresDR(Op& op, const cmx_solver_params& p, Vec& psi, const Vec& chi) : m_(new Impl(op, p, psi, chi)) {}
void FgmresDR::reserve(Op& op, const cmx_solver_params& p) {
  const int n = op.subset_sites(), ps = p.sloppy_prec, pb = p.basis_prec < 0 ? ps : p.basis_prec;
  CMX_REQUIRE(ps >= P64 && ps <= P16 && pb >= P64 && pb <= P16 && p.n_krylov >= 1 && p.n_defl >= 0 &&
                  p.n_krylov + p.n_defl + 1 <= KRY_MAXV, CMX_ERR_ARG, "FGMRES-DR: bad parameters");
  const int nd = p.n_defl, total = p.n_krylov + nd;
  for (int i = 0; i <= total; ++i) op.work(100 + i, pb, n);
  for (int i = 0; i < total; ++i) op.work(150 + i, pb, n);
  for (int i = 0; i <= nd; ++i) op.work(200 + i, pb, n);
  for (int i = 0; i < nd; ++i) op.work(250 + i, pb, n);
  op.work(300, P64, n);
  op.work(301, P64, n);
  if (nd > 0 && pb != P64) op.work(306, pb, n);
  if (pb != ps) op.work(302, ps, n), op.work(303, ps, n);
  if (p.precond == CMX_PRECOND_MR) op.work(304, ps, n), op.work(305, ps, n);
  KryDev::partials_of(op);
}
FgmresDR::~FgmresDR() = default;
bool FgmresDR::finished() const { return m_->done; }
void FgmresDR::begin_cycle() { m_->begin_cycle(); }
void FgmresDR::issue_step(bool preconditioned, bool applied) { m_->issue_step(preconditioned, applied); }
bool FgmresDR::apply_step_io(const Vec*& z, Vec*& w) const { return m_->apply_step_io(z, w); }
void FgmresDR::issue_precondition() { m_->precondition(*m_->Z[m_->j], *m_->V[m_->j]); }
bool FgmresDR::learned_step_io(const Vec*& v, Vec*& z) const { return m_->learned_step_io(v, z); }
bool FgmresDR::mg_step_io(const Vec*& v, Vec*& z) const { return m_->mg_step_io(v, z); }
bool FgmresDR::finish_step() { return m_->finish_step(); }
void FgmresDR::end_cycle() { m_->end_cycle(); }
void FgmresDR::result(cmx_solver_result& res) { m_->result(res); }

void solve_fgmresdr(Op& op, const cmx_solver_params& p, Vec& psi, const Vec& chi, cmx_solver_result& res) {
  FgmresDR run(op, p, psi, chi);
  while (!run.finished()) {
    run.begin_cycle();
    do {
      run.issue_step();
      op.sync();
    } while (run.finish_step());
    run.end_cycle();
  }
  run.result(res);
}

}  // namespace cmx
