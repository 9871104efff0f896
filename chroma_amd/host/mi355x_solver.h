// C++ host side of the MI355X Wilson-clover library: the Chroma interfaces of the fermion-matrix
// path, restated over the C-ABI (chroma_mi355x.h), so that a caller written against them runs on the
// device library.  Names, argument meaning and error behaviour follow the reference:
//
//   SystemSolverResults_t, SystemSolver<T>::operator()(psi, chi)     lib/syssolver.h:16-46
//   EvenOddPrecLinearOperator blocks (evenEvenLinOp, evenEvenInvLinOp, evenOddLinOp, oddEvenLinOp,
//     oddOddLinOp, unprecLinOp) and operator() on the odd subset    lib/eoprec_linop.h:98-165
//   LinOpSysSolverMI355X: the inverter a Chroma build registers in place of QUDA_CLOVER_INVERTER
//     (syssolver_linop_clover_quda_w.h:41-168; parameters as SysSolverQUDACloverParams,
//     syssolver_quda_clover_params.cc:12-131, and SysSolverCGParams)
//   MdagMSysSolverMI355X: MdagMSystemSolver<T>::operator()           lib/syssolver.h:72-84
//   MdagMMultiSysSolverCGMI355X: MdagMMultiSystemSolver<T>::operator()(psi, shifts, chi), the RHMC
//     multishift solve                                               multi_syssolver_mdagm_cg.h:58-100
//   MREChronoPredictor: MinimalResidualExtrapolation4DChronoPredictor, the HMC initial guess of the
//     M^dag M solves                                                 mre_extrap_predictor.h:104-230
//   PrecFermActQprop::operator()                                     eoprec_fermact_qprop.cc:41-80
//   quarkProp4 (QUARK_SPIN_TYPE_FULL point sources)                  quarkprop4_w.cc:31-117
//   the propagator correlator of the propagator task                 inline_propagator_w.cc:383-395
//
// Fields are host arrays in the library's host layout: a LatticeFermion is V*24 doubles,
// checkerboard-major (rb[0] = sites [0, Vh), rb[1] = [Vh, V)), 24 doubles per site [spin][colour][re,im].
// As in the reference, a solver on an even-odd operator acts on rb[1] of full-lattice fields.
#pragma once

#include <complex>
#include <deque>
#include <stdexcept>
#include <string>
#include <vector>

#include "chroma_mi355x.h"

namespace cmx_chroma {

using LatticeFermion = std::vector<double>;
using Complex = std::complex<double>;

// lib/syssolver.h:16-27
struct SystemSolverResults_t {
  int n_count = 0;    // iterations
  double resid = 0;   // true residual |rhs - A soln|
};

// The reference aborts the job (QDP_abort(1)) when the device solver fails and SilentFail is not
// set (syssolver_linop_clover_quda_w.h:638-644); this host layer throws instead.
class SolverAbort : public std::runtime_error {
 public:
  explicit SolverAbort(const std::string& what) : std::runtime_error(what) {}
};

// Throws SolverAbort with the library's message for a non-zero status.
void check(int status, const char* call);

// Mass from kappa (param_io.cc:12-15): Mass = 1/(2 Kappa) - Nd
inline double kappaToMass(double kappa) { return 1.0 / (2.0 * kappa) - 4.0; }

// Links with the fermion boundary conditions applied (what FermState::getLinks() hands to the
// operator, simple_fermbc.h:87-104, twisted_fermbc_w.h:96-119); 4*V*18 doubles.
class GaugeField {
 public:
  GaugeField(const int nrow[4], const std::vector<double>& links);
  ~GaugeField();
  GaugeField(const GaugeField&) = delete;
  GaugeField& operator=(const GaugeField&) = delete;
  cmx_gauge handle() const { return g_; }
  const int* nrow() const { return nrow_; }

 private:
  cmx_gauge g_ = nullptr;
  int nrow_[4];
};

// Fermion-action parameters of the operator (WilsonFermActParams / CloverFermActParams,
// clover_fermact_params_w.cc:27-99), and the device storage of its sloppy copy.
struct FermActParams {
  std::string FermAct = "WILSON";   // WILSON | CLOVER
  double Mass = 0;                   // kappaToMass(Kappa)
  double clovCoeffR = 0, clovCoeffT = 0;
  std::string SloppyPrecision = "SINGLE";  // SINGLE | HALF: precision of the inner operator copy
  int Reconstruct = 12;              // 12 | 18 (QUDA's cuda_reconstruct)
  bool AutotuneDslash = false;
};

// EvenOddPrecWilsonLinOp / EvenOddPrecCloverLinOp (eoprec_wilson_linop_w.cc:48-166,
// eoprec_clover_linop_w.cc:19-187) on the device.
class EvenOddPrecLinOp {
 public:
  EvenOddPrecLinOp(const GaugeField& u, const FermActParams& p);
  ~EvenOddPrecLinOp();
  EvenOddPrecLinOp(const EvenOddPrecLinOp&) = delete;
  EvenOddPrecLinOp& operator=(const EvenOddPrecLinOp&) = delete;

  // chi = M(isign) psi on rb[1] (lib/eoprec_linop.h:125-141), FP64
  void operator()(LatticeFermion& chi, const LatticeFermion& psi, int isign) const;
  void evenEvenLinOp(LatticeFermion& chi, const LatticeFermion& psi, int isign) const { piece(0, chi, psi, isign); }
  void evenEvenInvLinOp(LatticeFermion& chi, const LatticeFermion& psi, int isign) const { piece(1, chi, psi, isign); }
  void evenOddLinOp(LatticeFermion& chi, const LatticeFermion& psi, int isign) const { piece(2, chi, psi, isign); }
  void oddEvenLinOp(LatticeFermion& chi, const LatticeFermion& psi, int isign) const { piece(3, chi, psi, isign); }
  void oddOddLinOp(LatticeFermion& chi, const LatticeFermion& psi, int isign) const { piece(4, chi, psi, isign); }
  // the unpreconditioned operator on the full lattice (eoprec_linop.h:147-165)
  void unprecLinOp(LatticeFermion& chi, const LatticeFermion& psi, int isign) const { piece(5, chi, psi, isign); }

  cmx_op handle() const { return op_; }
  long V() const { return V_; }
  long Vh() const { return V_ / 2; }

 private:
  void piece(int which, LatticeFermion& chi, const LatticeFermion& psi, int isign) const;
  cmx_op op_ = nullptr;
  long V_ = 0;
};

// Solver parameters, named as the reference's InvertParam XML names them (SysSolverCGParams,
// SysSolverQUDACloverParams, SysSolverFGMRESDRParams)
struct SysSolverMI355XParams {
  std::string SolverType = "CG";  // CG | BICGSTAB | RELIABLE_CG | RELIABLE_BICGSTAB | FGMRESDR
  int MaxIter = 1000;
  double RsdTarget = 1e-8;
  double Delta = 0.1;             // reliable-update delta
  std::string SloppyPrecision = "SINGLE";  // DOUBLE | SINGLE | HALF (inner precision)
  bool TrueResidual = false;      // restart until |chi - A psi| <= RsdTarget |chi| in FP64
  int NKrylov = 16, NDefl = 0;    // FGMRESDR
  bool SilentFail = false;        // return the results of a failed solve instead of aborting
  bool Verbose = false;
};

// LinOpSystemSolver<LatticeFermion> on the odd checkerboard of an even-odd operator
// (syssolver_linop.h:20-23): solves A psi = chi on rb[1], psi is the initial guess.
class LinOpSysSolverMI355X {
 public:
  LinOpSysSolverMI355X(const EvenOddPrecLinOp& A, const SysSolverMI355XParams& p);
  SystemSolverResults_t operator()(LatticeFermion& psi, const LatticeFermion& chi) const;
  // Several right-hand sides on this operator advanced together (cmx_solve_batch: one launch per
  // kernel for all of them); each solution equals its single solve's.  CG and RELIABLE_CG.
  std::vector<SystemSolverResults_t> operator()(std::vector<LatticeFermion>& psi,
                                                const std::vector<LatticeFermion>& chi) const;
  // the device solver's own figures of the last solve (iterations, FP64 residual, device seconds)
  const cmx_solver_result& lastResult() const { return last_; }

 private:
  const EvenOddPrecLinOp& A_;
  cmx_solver_params sp_;
  bool silent_;
  mutable cmx_solver_result last_{};
};

// MdagMSystemSolver<LatticeFermion> (lib/syssolver.h:72-84): solves M^dag M psi = chi on rb[1]
// (MdagMSysSolverCG, syssolver_mdagm_cg.h:56-90, and the reliable and two-step BiCGStab forms);
// resid = |chi - M^dag M psi|.
class MREChronoPredictor;
class MdagMSysSolverMI355X {
 public:
  MdagMSysSolverMI355X(const EvenOddPrecLinOp& A, const SysSolverMI355XParams& p);
  SystemSolverResults_t operator()(LatticeFermion& psi, const LatticeFermion& chi) const;
  // with a chronological predictor (the MdagM solvers' predictor argument, syssolver_mdagm.h:24-36):
  // psi starts from the predictor's guess and the solution is added to its buffer
  SystemSolverResults_t operator()(LatticeFermion& psi, const LatticeFermion& chi, MREChronoPredictor& chrono) const;

 private:
  const EvenOddPrecLinOp& A_;
  cmx_solver_params sp_;
  bool silent_;
};

// MinimalResidualExtrapolation4DChronoPredictor (mre_extrap_predictor.h:104-230): keeps the last
// MaxChrono solutions; the guess for A psi = chi (A = M^dag M on rb[1]) is sum_n a_n v_n, with v the
// Gram-Schmidt-orthonormalised buffer (newest first) and G a = b, G_nm = <v_n, A v_m>, b_n = <v_n, chi>,
// solved by LU.  An empty buffer predicts zero.
class MREChronoPredictor {
 public:
  explicit MREChronoPredictor(int MaxChrono) : max_(MaxChrono) {}
  void operator()(LatticeFermion& psi, const EvenOddPrecLinOp& M, const LatticeFermion& chi) const;
  void newVector(const LatticeFermion& psi);
  void reset() { buf_.clear(); }
  int size() const { return (int)buf_.size(); }

 private:
  int max_;
  std::deque<LatticeFermion> buf_;  // newest first (CircularBuffer order)
};

// MdagMMultiSysSolverCG (multi_syssolver_mdagm_cg.h:58-100 -> MInvCG2, minvcg2.cc:72-355): psi[s] =
// (M^dag M + shifts[s])^-1 chi on rb[1] for every shift with one Krylov space, FP64.  RsdCG holds one
// target or one per shift, as the reference's parameter does (:62-75); psi is resized to the shifts.
class MdagMMultiSysSolverCGMI355X {
 public:
  MdagMMultiSysSolverCGMI355X(const EvenOddPrecLinOp& A, std::vector<double> RsdCG, int MaxCG)
      : A_(A), rsd_(std::move(RsdCG)), maxcg_(MaxCG) {}
  SystemSolverResults_t operator()(std::vector<LatticeFermion>& psi, const std::vector<double>& shifts,
                                   const LatticeFermion& chi) const;
  // |chi - (M^dag M + shift_s) psi_s| of the last solve, per shift
  const std::vector<double>& shiftResiduals() const { return resid_; }

 private:
  const EvenOddPrecLinOp& A_;
  std::vector<double> rsd_;
  int maxcg_;
  mutable std::vector<double> resid_;
};

// EvenOddPrecWilsonTypeFermAct::qprop's solver object (eoprec_fermact_qprop.cc:16-80): even-odd
// source preparation, the odd solve, the even reconstruction and the unpreconditioned residual.
class PrecFermActQprop {
 public:
  PrecFermActQprop(const EvenOddPrecLinOp& A, const LinOpSysSolverMI355X& invA) : A_(A), invA_(invA) {}
  SystemSolverResults_t operator()(LatticeFermion& psi, const LatticeFermion& chi) const;
  // the same on several sources, with one batched odd solve
  std::vector<SystemSolverResults_t> operator()(std::vector<LatticeFermion>& psi,
                                                const std::vector<LatticeFermion>& chi) const;

 private:
  void prepare(LatticeFermion& chi_tmp, const LatticeFermion& chi) const;
  void reconstruct(LatticeFermion& psi, const LatticeFermion& chi, SystemSolverResults_t& res) const;
  const EvenOddPrecLinOp& A_;
  const LinOpSysSolverMI355X& invA_;
};

// quarkProp4_a (quarkprop4_w.cc:31-117) for point sources at site src_site: the 12 spin x colour
// solves with source normalisation.  prop[(site*12 + sink)*12 + source] with index spin*3 + colour.
// Returns the summed iteration count (ncg_had).
// batched = true solves the 12 sources together (one batched odd solve).
int quarkProp4(std::vector<Complex>& prop, const PrecFermActQprop& qprop, long V, long src_site,
               std::vector<SystemSolverResults_t>* results = nullptr, bool batched = false);

// sum over each time slice of localNorm2(prop) (inline_propagator_w.cc:383-395)
std::vector<double> propCorr(const int nrow[4], const std::vector<Complex>& prop);

}  // namespace cmx_chroma
