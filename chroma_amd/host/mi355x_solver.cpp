// C++ host side over the C-ABI: see mi355x_solver.h.
#include "mi355x_solver.h"

#include <algorithm>
#include <cmath>

namespace cmx_chroma {

void check(int status, const char* call) {
  if (status != CMX_OK) throw SolverAbort(std::string(call) + ": " + cmx_last_error());
}

namespace {
int precision_code(const std::string& s) {
  if (s == "DOUBLE") return CMX_DOUBLE;
  if (s == "SINGLE") return CMX_SINGLE;
  if (s == "HALF") return CMX_HALF;
  throw SolverAbort("unknown precision " + s);
}
int solver_code(const std::string& s) {
  if (s == "CG") return CMX_SOLVER_CG;
  if (s == "BICGSTAB") return CMX_SOLVER_BICGSTAB;
  if (s == "RELIABLE_CG") return CMX_SOLVER_RELIABLE_CG;
  if (s == "RELIABLE_BICGSTAB") return CMX_SOLVER_RELIABLE_BICGSTAB;
  if (s == "FGMRESDR") return CMX_SOLVER_FGMRESDR;
  throw SolverAbort("unknown SolverType " + s);  // the factory's StringFactoryError role
}
double norm2(const double* v, long n) {
  double s = 0;
  for (long k = 0; k < n; ++k) s += v[k] * v[k];
  return s;
}
}  // namespace

GaugeField::GaugeField(const int nrow[4], const std::vector<double>& links) {
  long V = 1;
  for (int m = 0; m < 4; ++m) nrow_[m] = nrow[m], V *= nrow[m];
  if ((long)links.size() != 4 * V * 18) throw SolverAbort("GaugeField: expected 4*V*18 doubles");
  check(cmx_gauge_create(nrow_, links.data(), &g_), "cmx_gauge_create");
}
GaugeField::~GaugeField() {
  if (g_) cmx_gauge_destroy(g_);
}

EvenOddPrecLinOp::EvenOddPrecLinOp(const GaugeField& u, const FermActParams& p) {
  V_ = 1;
  for (int m = 0; m < 4; ++m) V_ *= u.nrow()[m];
  cmx_op_params op{};
  if (p.FermAct == "WILSON")
    op.kind = CMX_OP_EO_WILSON;
  else if (p.FermAct == "CLOVER")
    op.kind = CMX_OP_EO_CLOVER;
  else
    throw SolverAbort("unknown FermAct " + p.FermAct);
  op.mass = p.Mass;
  op.clov_coeff_r = p.clovCoeffR;
  op.clov_coeff_t = p.clovCoeffT;
  op.sloppy_prec = precision_code(p.SloppyPrecision);
  op.sloppy_recon = p.Reconstruct;
  op.autotune_dslash = p.AutotuneDslash ? 1 : 0;
  check(cmx_op_create(u.handle(), &op, &op_), "cmx_op_create");
}
EvenOddPrecLinOp::~EvenOddPrecLinOp() {
  if (op_) cmx_op_destroy(op_);
}

void EvenOddPrecLinOp::operator()(LatticeFermion& chi, const LatticeFermion& psi, int isign) const {
  if ((long)chi.size() != 24 * V_ || (long)psi.size() != 24 * V_) throw SolverAbort("LinOp: field size");
  const long off = 24 * Vh();
  check(cmx_op_apply(op_, chi.data() + off, psi.data() + off, isign, CMX_DOUBLE), "cmx_op_apply");
}

void EvenOddPrecLinOp::piece(int which, LatticeFermion& chi, const LatticeFermion& psi, int isign) const {
  if ((long)chi.size() != 24 * V_ || (long)psi.size() != 24 * V_) throw SolverAbort("LinOp: field size");
  check(cmx_op_eo_piece(op_, which, chi.data(), psi.data(), isign), "cmx_op_eo_piece");
}

namespace {
cmx_solver_params solver_params(const SysSolverMI355XParams& p) {
  cmx_solver_params sp{};
  sp.solver = solver_code(p.SolverType);
  sp.max_iter = p.MaxIter;
  sp.rsd_target = p.RsdTarget;
  sp.delta = p.Delta;
  sp.sloppy_prec = precision_code(p.SloppyPrecision);
  sp.check_every = 0;
  sp.verbose = p.Verbose ? 1 : 0;
  sp.true_residual = p.TrueResidual ? 1 : 0;
  sp.n_krylov = p.NKrylov;
  sp.n_defl = p.NDefl;
  sp.precond = CMX_PRECOND_NONE;
  sp.reorth = 1;
  sp.basis_prec = -1;
  return sp;
}
}  // namespace

LinOpSysSolverMI355X::LinOpSysSolverMI355X(const EvenOddPrecLinOp& A, const SysSolverMI355XParams& p)
    : A_(A), sp_(solver_params(p)), silent_(p.SilentFail) {}

MdagMSysSolverMI355X::MdagMSysSolverMI355X(const EvenOddPrecLinOp& A, const SysSolverMI355XParams& p)
    : A_(A), sp_(solver_params(p)), silent_(p.SilentFail) {}

SystemSolverResults_t MdagMSysSolverMI355X::operator()(LatticeFermion& psi, const LatticeFermion& chi) const {
  const long V = A_.V(), off = 24 * A_.Vh();
  if ((long)chi.size() != 24 * V || (long)psi.size() != 24 * V) throw SolverAbort("solver: field size");
  cmx_solver_result r{};
  const int st = cmx_solve_mdagm(A_.handle(), &sp_, psi.data() + off, chi.data() + off, 1, &r);
  if (st != CMX_OK && !(silent_ && st == CMX_ERR_NOCONV)) check(st, "cmx_solve_mdagm");
  SystemSolverResults_t res;
  res.n_count = r.n_count;
  res.resid = r.resid;
  return res;
}

SystemSolverResults_t MdagMMultiSysSolverCGMI355X::operator()(std::vector<LatticeFermion>& psi,
                                                             const std::vector<double>& shifts,
                                                             const LatticeFermion& chi) const {
  const long V = A_.V(), Vh = A_.Vh(), off = 24 * Vh;
  const int ns = (int)shifts.size();
  if ((long)chi.size() != 24 * V) throw SolverAbort("multishift: field size");
  std::vector<double> rsd(ns);
  if (rsd_.size() == 1)
    rsd.assign(ns, rsd_[0]);
  else if ((int)rsd_.size() == ns)
    rsd = rsd_;
  else
    throw SolverAbort("MdagMMultiSysSolverCG: shifts incompatible");  // :62-75
  std::vector<double> buf((size_t)ns * 24 * Vh);
  std::vector<cmx_solver_result> r(ns);
  cmx_solver_params sp = solver_params(SysSolverMI355XParams{});
  sp.max_iter = maxcg_;
  check(cmx_solve_multishift(A_.handle(), &sp, ns, shifts.data(), rsd.data(), buf.data(), chi.data() + off, r.data()),
        "cmx_solve_multishift");
  psi.assign(ns, LatticeFermion(24 * V, 0.0));
  resid_.assign(ns, 0.0);
  for (int s = 0; s < ns; ++s) {
    std::copy(buf.begin() + (size_t)s * 24 * Vh, buf.begin() + (size_t)(s + 1) * 24 * Vh, psi[s].begin() + off);
    resid_[s] = r[s].resid;
  }
  SystemSolverResults_t res;
  res.n_count = r[0].n_count;
  for (int s = 0; s < ns; ++s) res.resid = std::max(res.resid, r[s].resid);
  return res;
}

namespace {
// <a, b> on rb[1] (innerProduct over the odd subset)
Complex odd_dot(const LatticeFermion& a, const LatticeFermion& b, long Vh) {
  double re = 0, im = 0;
  for (long k = 24 * Vh; k < 48 * Vh; k += 2) {
    re += a[k] * b[k] + a[k + 1] * b[k + 1];
    im += a[k] * b[k + 1] - a[k + 1] * b[k];
  }
  return Complex(re, im);
}
// x_s += c y_s on rb[1]
void odd_axpy(LatticeFermion& x, Complex c, const LatticeFermion& y, long Vh) {
  for (long k = 24 * Vh; k < 48 * Vh; k += 2) {
    x[k] += c.real() * y[k] - c.imag() * y[k + 1];
    x[k + 1] += c.real() * y[k + 1] + c.imag() * y[k];
  }
}
// LUSolve (lu_solve.cc): a = G^-1 b, Gaussian elimination with partial pivoting
std::vector<Complex> lu_solve(std::vector<Complex> G, std::vector<Complex> b, int n) {
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::abs(G[r * n + c]) > std::abs(G[piv * n + c])) piv = r;
    if (std::abs(G[piv * n + c]) == 0) throw SolverAbort("MRE predictor: singular G");
    if (piv != c) {
      for (int k = 0; k < n; ++k) std::swap(G[c * n + k], G[piv * n + k]);
      std::swap(b[c], b[piv]);
    }
    for (int r = c + 1; r < n; ++r) {
      const Complex f = G[r * n + c] / G[c * n + c];
      for (int k = c; k < n; ++k) G[r * n + k] -= f * G[c * n + k];
      b[r] -= f * b[c];
    }
  }
  std::vector<Complex> a(n);
  for (int r = n - 1; r >= 0; --r) {
    Complex s = b[r];
    for (int k = r + 1; k < n; ++k) s -= G[r * n + k] * a[k];
    a[r] = s / G[r * n + r];
  }
  return a;
}
}  // namespace

void MREChronoPredictor::newVector(const LatticeFermion& psi) {
  buf_.push_front(psi);
  while ((int)buf_.size() > max_) buf_.pop_back();
}

void MREChronoPredictor::operator()(LatticeFermion& psi, const EvenOddPrecLinOp& M, const LatticeFermion& chi) const {
  const long V = M.V(), Vh = M.Vh();
  psi.assign(24 * V, 0.0);
  const int n = (int)buf_.size();
  if (n == 0) return;
  // orthonormal basis of the buffer (GramSchm, then normalise)
  std::vector<LatticeFermion> v;
  for (int i = 0; i < n; ++i) {
    LatticeFermion t(24 * V, 0.0);
    std::copy(buf_[i].begin() + 24 * Vh, buf_[i].end(), t.begin() + 24 * Vh);
    for (int j = 0; j < i; ++j) odd_axpy(t, -odd_dot(v[j], t, Vh), v[j], Vh);
    const double nrm = std::sqrt(odd_dot(t, t, Vh).real());
    if (nrm == 0) throw SolverAbort("MRE predictor: linearly dependent buffer");
    for (long k = 24 * Vh; k < 48 * Vh; ++k) t[k] /= nrm;
    v.push_back(std::move(t));
  }
  // G_nm = <v_n, M^dag M v_m>, b_n = <v_n, chi>
  std::vector<Complex> G((size_t)n * n), b(n);
  LatticeFermion mv(24 * V, 0.0), mmv(24 * V, 0.0);
  for (int m = 0; m < n; ++m) {
    M(mv, v[m], CMX_PLUS);
    M(mmv, mv, CMX_MINUS);
    for (int r = 0; r < n; ++r) G[(size_t)r * n + m] = odd_dot(v[r], mmv, Vh);
  }
  for (int r = 0; r < n; ++r) b[r] = odd_dot(v[r], chi, Vh);
  const std::vector<Complex> a = lu_solve(G, b, n);
  for (int r = 0; r < n; ++r) odd_axpy(psi, a[r], v[r], Vh);
}

SystemSolverResults_t MdagMSysSolverMI355X::operator()(LatticeFermion& psi, const LatticeFermion& chi,
                                                       MREChronoPredictor& chrono) const {
  chrono(psi, A_, chi);
  const SystemSolverResults_t res = (*this)(psi, chi);
  chrono.newVector(psi);
  return res;
}

std::vector<SystemSolverResults_t> LinOpSysSolverMI355X::operator()(std::vector<LatticeFermion>& psi,
                                                                    const std::vector<LatticeFermion>& chi) const {
  const long V = A_.V(), off = 24 * A_.Vh();
  const int n = (int)chi.size();
  if ((int)psi.size() != n || n < 1) throw SolverAbort("batched solver: psi and chi counts differ");
  std::vector<cmx_op> ops(n, A_.handle());
  std::vector<double*> pp(n);
  std::vector<const double*> cp(n);
  for (int k = 0; k < n; ++k) {
    if ((long)chi[k].size() != 24 * V || (long)psi[k].size() != 24 * V) throw SolverAbort("solver: field size");
    pp[k] = psi[k].data() + off;
    cp[k] = chi[k].data() + off;
  }
  std::vector<cmx_solver_result> r(n);
  cmx_batch_params bp{/*tune*/ 0, /*graph*/ 0, /*max_concurrent*/ 0};
  cmx_batch_result br{};
  const int st = cmx_solve_batch(ops.data(), n, &sp_, &bp, pp.data(), cp.data(), 0, r.data(), &br);
  if (st != CMX_OK && !(silent_ && st == CMX_ERR_NOCONV)) check(st, "cmx_solve_batch");
  std::vector<SystemSolverResults_t> res(n);
  for (int k = 0; k < n; ++k) res[k].n_count = r[k].n_count, res[k].resid = r[k].resid;
  last_ = r.back();
  return res;
}

SystemSolverResults_t LinOpSysSolverMI355X::operator()(LatticeFermion& psi, const LatticeFermion& chi) const {
  const long V = A_.V(), off = 24 * A_.Vh();
  if ((long)chi.size() != 24 * V || (long)psi.size() != 24 * V) throw SolverAbort("solver: field size");
  last_ = cmx_solver_result{};
  const int st = cmx_solve(A_.handle(), &sp_, psi.data() + off, chi.data() + off, 1, &last_);
  if (st != CMX_OK && !(silent_ && st == CMX_ERR_NOCONV)) check(st, "cmx_solve");
  SystemSolverResults_t res;
  res.n_count = last_.n_count;
  res.resid = last_.resid;  // |chi - A psi| on rb[1], recomputed in FP64 on the device
  return res;
}

// (i) chi_tmp_o = chi_o - D_oe A_ee^-1 chi_e
void PrecFermActQprop::prepare(LatticeFermion& chi_tmp, const LatticeFermion& chi) const {
  const long V = A_.V(), Vh = A_.Vh();
  LatticeFermion tmp1(24 * V, 0.0), tmp2(24 * V, 0.0);
  A_.evenEvenInvLinOp(tmp1, chi, CMX_PLUS);
  A_.oddEvenLinOp(tmp2, tmp1, CMX_PLUS);
  chi_tmp.assign(24 * V, 0.0);
  for (long k = 24 * Vh; k < 24 * V; ++k) chi_tmp[k] = chi[k] - tmp2[k];
}

SystemSolverResults_t PrecFermActQprop::operator()(LatticeFermion& psi, const LatticeFermion& chi) const {
  LatticeFermion chi_tmp;
  prepare(chi_tmp, chi);
  SystemSolverResults_t res = invA_(psi, chi_tmp);
  reconstruct(psi, chi, res);
  return res;
}

std::vector<SystemSolverResults_t> PrecFermActQprop::operator()(std::vector<LatticeFermion>& psi,
                                                                const std::vector<LatticeFermion>& chi) const {
  std::vector<LatticeFermion> chi_tmp(chi.size());
  for (size_t k = 0; k < chi.size(); ++k) prepare(chi_tmp[k], chi[k]);
  std::vector<SystemSolverResults_t> res = invA_(psi, chi_tmp);
  for (size_t k = 0; k < chi.size(); ++k) reconstruct(psi[k], chi[k], res[k]);
  return res;
}

// (ii) psi_e = A_ee^-1 (chi_e - D_eo psi_o), then the residual of the unpreconditioned system
void PrecFermActQprop::reconstruct(LatticeFermion& psi, const LatticeFermion& chi, SystemSolverResults_t& res) const {
  const long V = A_.V(), Vh = A_.Vh();
  LatticeFermion tmp1(24 * V, 0.0), tmp2(24 * V, 0.0);
  A_.evenOddLinOp(tmp1, psi, CMX_PLUS);
  for (long k = 0; k < 24 * Vh; ++k) tmp2[k] = chi[k] - tmp1[k];
  LatticeFermion pe(24 * V, 0.0);
  A_.evenEvenInvLinOp(pe, tmp2, CMX_PLUS);
  for (long k = 0; k < 24 * Vh; ++k) psi[k] = pe[k];
  // residual of the unpreconditioned system
  LatticeFermion r(24 * V, 0.0);
  A_.unprecLinOp(r, psi, CMX_PLUS);
  for (long k = 0; k < 24 * V; ++k) r[k] -= chi[k];
  res.resid = std::sqrt(norm2(r.data(), 24 * V));
}

int quarkProp4(std::vector<Complex>& prop, const PrecFermActQprop& qprop, long V, long src_site,
               std::vector<SystemSolverResults_t>* results, bool batched) {
  prop.assign((size_t)V * 144, Complex(0, 0));
  // the 12 sources (colour outer, spin inner, as quarkprop4_w.cc:70-117 loops), normalised
  // (quarkprop4_w.cc:86-92, 107-109)
  std::vector<LatticeFermion> chi(12, LatticeFermion(24 * V, 0.0)), psi(12, LatticeFermion(24 * V, 0.0));
  std::vector<double> fact(12);
  for (int color_source = 0; color_source < 3; ++color_source)
    for (int spin_source = 0; spin_source < 4; ++spin_source) {
      const int j = color_source * 4 + spin_source;
      chi[j][(size_t)src_site * 24 + (spin_source * 3 + color_source) * 2] = 1.0;
      const double nrm = std::sqrt(norm2(chi[j].data(), 24 * V));
      fact[j] = nrm != 0 ? 1.0 / nrm : 1.0;
      for (double& v : chi[j]) v *= fact[j];
    }
  std::vector<SystemSolverResults_t> res;
  if (batched) {
    res = qprop(psi, chi);
  } else {
    for (int j = 0; j < 12; ++j) res.push_back(qprop(psi[j], chi[j]));
  }
  int ncg_had = 0;
  for (int color_source = 0; color_source < 3; ++color_source)
    for (int spin_source = 0; spin_source < 4; ++spin_source) {
      const int j = color_source * 4 + spin_source;
      ncg_had += res[j].n_count;
      if (results) results->push_back(res[j]);
      for (long s = 0; s < V; ++s)
        for (int k = 0; k < 12; ++k)
          prop[((size_t)s * 12 + k) * 12 + spin_source * 3 + color_source] =
              Complex(psi[j][(size_t)s * 24 + 2 * k], psi[j][(size_t)s * 24 + 2 * k + 1]) / fact[j];
    }
  return ncg_had;
}

std::vector<double> propCorr(const int nrow[4], const std::vector<Complex>& prop) {
  const long V = (long)nrow[0] * nrow[1] * nrow[2] * nrow[3], Vh = V / 2;
  std::vector<double> corr(nrow[3], 0.0);
  for (long s = 0; s < V; ++s) {
    const long t = (s % Vh) / ((long)(nrow[0] / 2) * nrow[1] * nrow[2]);
    double n = 0;
    for (int k = 0; k < 144; ++k) n += std::norm(prop[(size_t)s * 144 + k]);
    corr[t] += n;
  }
  return corr;
}

}  // namespace cmx_chroma
