// FGMRES-DR as a stepping object, so that several solves can advance together on one stream with one
// host read-back per Arnoldi step for all of them (the batched solver), while each performs exactly the
// kernels and host arithmetic of a single solve.
#pragma once
#include <memory>

#include "op.h"

namespace cmx {

// MR preconditioner (LinOpSysSolverMR -> InvMR, invmr.cc:55-160) in precision ps: z is the initial
// guess on entry (the flexible solver passes Z_j) and the result on exit.
struct MRPrecond {
  Op& op;
  int ps, n, maxit;
  double rsd, omega;
  Vec &r, &mr;
  MRPrecond(Op& o, int ps_, int n_, int maxit_, double rsd_, double omega_, Vec& r_, Vec& mr_)
      : op(o), ps(ps_), n(n_), maxit(maxit_), rsd(rsd_), omega(omega_), r(r_), mr(mr_) {}
  void apply(Vec& z, const Vec& v) {
    const int nb = blas_blocks(n);
    int* F = op.state_F.as<int>();
    op.hS[S_MR + 9] = omega;
    op.hS[S_MR + 10] = rsd;
    CMX_HIP(hipMemcpyAsync(static_cast<double*>(op.state_S.p) + S_MR + 9, op.hS + S_MR + 9, 2 * sizeof(double),
                           hipMemcpyHostToDevice, op.stream));
    op.blas(B_NORM2, {&v});
    op.finalize(nb, 1, PG_STORE, S_MR + 7);  // |chi|^2
    op.apply(ps, mr.view(), z.view(), +1, nullptr);
    op.blas(B_XMY_NORM, {&r, &v, &mr});      // r = chi - M psi
    op.finalize(nb, 1, PG_MR_INIT);
    for (int k = 0; k < maxit; ++k) {
      if (op.eo()) {  // M r with |Mr|^2 and <Mr, r> from the operator's second kernel
        const int nh = op.apply(ps, mr.view(), r.view(), +1, op.partials.as<double>(), F + F_MRDONE, nullptr, true);
        op.finalize(nh, 3, PG_MR_A);                                   // a = omega c / d
      } else {
        op.apply(ps, mr.view(), r.view(), +1, nullptr, F + F_MRDONE);  // M r
        op.blas(B_TNORM_TDOT, {&mr, &r}, 0, 0, F_MRDONE);               // |Mr|^2, <Mr, r>
        op.finalize(nb, 3, PG_MR_A);                                     // a = omega c / d
      }
      op.blas(B_MR_UPDATE, {&z, &r, &mr}, S_MR + 3, S_MR + 5, F_MRDONE);  // psi += a r ; r -= a M r
      op.finalize(nb, 1, PG_MR_R);
    }
  }
};

class FgmresDR {
 public:
  // Sets up the workspace of `op` and computes the initial residual (synchronises op's stream)
  FgmresDR(Op& op, const cmx_solver_params& p, Vec& psi, const Vec& chi);
  ~FgmresDR();
  FgmresDR(const FgmresDR&) = delete;
  FgmresDR& operator=(const FgmresDR&) = delete;
  // allocates the workspace a solve with these parameters uses on op (the constructor does so too)
  static void reserve(Op& op, const cmx_solver_params& p);
  bool finished() const;
  // the restart: plain (V_0 = r / |r|) or deflated (harmonic Ritz vectors of the previous cycle)
  void begin_cycle();
  // the next step's preconditioner input and output when that step applies the learned network directly
  // to basis vectors (precond LEARNED, basis precision = operator precision): a batched caller applies
  // it for several solves at once and then calls issue_step(true)
  bool learned_step_io(const Vec*& v, Vec*& z) const;
  // the same for the multigrid preconditioner (precond MG, basis precision = operator precision)
  bool mg_step_io(const Vec*& v, Vec*& z) const;
  // z_j = M v_j of the next step alone (then issue_step(true, ...))
  void issue_precondition();
  // the next step's operator input z_j and output w when the operator acts on basis vectors directly
  // (basis precision = operator precision): a batched caller applies it for several solves at once
  bool apply_step_io(const Vec*& z, Vec*& w) const;
  // issues one Arnoldi step on op's stream, ending with an asynchronous read-back of the new column;
  // preconditioned: z_j = M v_j has been formed already; applied: also w = A z_j
  void issue_step(bool preconditioned = false, bool applied = false);
  // after op's stream has reached the read-back: the Givens update; false when the cycle has ended
  bool finish_step();
  // psi += Z eta, the FP64 residual, the stopping decision
  void end_cycle();
  // fills res; throws CMX_ERR_NOCONV when the target was not reached
  void result(cmx_solver_result& res);

 private:
  struct Impl;
  std::unique_ptr<Impl> m_;
};

// one solve (LinOpSysSolverFGMRESDR::operator(), syssolver_linop_fgmres_dr.cc:448-774)
void solve_fgmresdr(Op& op, const cmx_solver_params& p, Vec& psi, const Vec& chi, cmx_solver_result& res);

}  // namespace cmx
