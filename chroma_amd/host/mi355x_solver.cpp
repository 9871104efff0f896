// C++ host side over the C-ABI: see mi355x_solver.h.
#include "mi355x_solver.h"

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

LinOpSysSolverMI355X::LinOpSysSolverMI355X(const EvenOddPrecLinOp& A, const SysSolverMI355XParams& p)
    : A_(A), sp_{}, silent_(p.SilentFail) {
  sp_.solver = solver_code(p.SolverType);
  sp_.max_iter = p.MaxIter;
  sp_.rsd_target = p.RsdTarget;
  sp_.delta = p.Delta;
  sp_.sloppy_prec = precision_code(p.SloppyPrecision);
  sp_.check_every = 0;
  sp_.verbose = p.Verbose ? 1 : 0;
  sp_.true_residual = p.TrueResidual ? 1 : 0;
  sp_.n_krylov = p.NKrylov;
  sp_.n_defl = p.NDefl;
  sp_.precond = CMX_PRECOND_NONE;
  sp_.reorth = 1;
  sp_.basis_prec = -1;
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

SystemSolverResults_t PrecFermActQprop::operator()(LatticeFermion& psi, const LatticeFermion& chi) const {
  const long V = A_.V(), Vh = A_.Vh();
  // (i) chi_tmp_o = chi_o - D_oe A_ee^-1 chi_e
  LatticeFermion tmp1(24 * V, 0.0), tmp2(24 * V, 0.0), chi_tmp(24 * V, 0.0);
  A_.evenEvenInvLinOp(tmp1, chi, CMX_PLUS);
  A_.oddEvenLinOp(tmp2, tmp1, CMX_PLUS);
  for (long k = 24 * Vh; k < 24 * V; ++k) chi_tmp[k] = chi[k] - tmp2[k];
  SystemSolverResults_t res = invA_(psi, chi_tmp);
  // (ii) psi_e = A_ee^-1 (chi_e - D_eo psi_o)
  std::fill(tmp1.begin(), tmp1.end(), 0.0);
  std::fill(tmp2.begin(), tmp2.end(), 0.0);
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
  return res;
}

int quarkProp4(std::vector<Complex>& prop, const PrecFermActQprop& qprop, long V, long src_site,
               std::vector<SystemSolverResults_t>* results) {
  prop.assign((size_t)V * 144, Complex(0, 0));
  int ncg_had = 0;
  for (int color_source = 0; color_source < 3; ++color_source)
    for (int spin_source = 0; spin_source < 4; ++spin_source) {
      LatticeFermion chi(24 * V, 0.0), psi(24 * V, 0.0);
      chi[(size_t)src_site * 24 + (spin_source * 3 + color_source) * 2] = 1.0;
      // source normalisation (quarkprop4_w.cc:86-92, 107-109)
      const double nrm = std::sqrt(norm2(chi.data(), 24 * V));
      const double fact = nrm != 0 ? 1.0 / nrm : 1.0;
      for (double& v : chi) v *= fact;
      const SystemSolverResults_t res = qprop(psi, chi);
      ncg_had += res.n_count;
      if (results) results->push_back(res);
      for (long s = 0; s < V; ++s)
        for (int k = 0; k < 12; ++k)
          prop[((size_t)s * 12 + k) * 12 + spin_source * 3 + color_source] =
              Complex(psi[(size_t)s * 24 + 2 * k], psi[(size_t)s * 24 + 2 * k + 1]) / fact;
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
