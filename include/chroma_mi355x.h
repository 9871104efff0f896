/*
 * chroma_mi355x.h -- C-ABI of the MI355X-native Wilson-clover fermion-matrix library.
 *
 * This is the drop-in boundary that a Chroma-side LinOpSystemSolver (registered in the InvertParam
 * factory the way QUDA_CLOVER_INVERTER is) binds to.  Plain C types only: int status codes, host
 * pointers and sizes, opaque handles.  Nothing throws across this boundary; on failure a call
 * returns a non-zero CMX_ERR_* code and cmx_last_error() describes it (the Chroma adapter maps a
 * failure to QDP_abort(1) unless SilentFail is set, as syssolver_linop_clover_quda_w.h:638-644 does).
 *
 * Host layout of every host array crossing this boundary (bit-exact permutation of the device
 * layout; see cmx_layout_roundtrip):
 *   lattice dims L = {Lx, Ly, Lz, Lt}, all even;  V = Lx*Ly*Lz*Lt,  Vh = V/2
 *   checkerboard cb(x) = (x+y+z+t) mod 2;  full site index = cb*Vh + (((t*Lz+z)*Ly+y)*Lx + x)/2
 *   fermion : per site 24 doubles, [spin 0..3][colour 0..2][re,im]   (QDP++ Fermion elem(s).elem(c))
 *   link    : per site 18 doubles, 3x3 complex row-major            (QDP++ ColorMatrix elem(i,j))
 *   clover  : per site 72 doubles, {diag[2][6], offd[2][15][re,im]}  (PrimitiveClovTriang,
 *             clover_term_qdp_w.h:19-24)
 * An even-odd operator acts on the odd checkerboard: its fermion arguments are the Vh*24 doubles
 * that start at full site index Vh, i.e. &chi.elem(rb[1].start()) as syssolver_linop_clover_quda_w.cc:76,85
 * passes them.
 */
#ifndef CHROMA_MI355X_H
#define CHROMA_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------------- */
#define CMX_OK 0
#define CMX_ERR_ARG 1          /* invalid argument or inconsistent sizes                          */
#define CMX_ERR_HIP 2          /* HIP runtime failure                                             */
#define CMX_ERR_STATE 3        /* library not initialised / handle in wrong state                 */
#define CMX_ERR_NOCONV 4       /* solver reached MaxIter (reliable_cg.cc:180-183 aborts here)     */
#define CMX_ERR_BREAKDOWN 5    /* BiCGStab breakdown (invbicgstab.cc:80-83,109-112,133-136)      */
#define CMX_ERR_UNSUPPORTED 6  /* option not supported by this build                              */

/* ---- enums ----------------------------------------------------------------------------------- */
#define CMX_DOUBLE 0  /* FP64 storage and arithmetic                                                */
#define CMX_SINGLE 1  /* FP32 storage and arithmetic                                                */
#define CMX_HALF 2    /* FP16 storage (per-site FP32 norm for fermions), FP32 arithmetic            */

#define CMX_RECONS_18 18 /* 3x3 complex links                                                        */
#define CMX_RECONS_12 12 /* two rows stored, third rebuilt in registers (QUDA_RECONSTRUCT_12 role)  */

#define CMX_OP_UNPREC_WILSON 0 /* UnprecWilsonLinOp, unprec_wilson_linop_w.cc:58-77                 */
#define CMX_OP_EO_WILSON 1     /* EvenOddPrecWilsonLinOp, eoprec_wilson_linop_w.cc:126-166           */
#define CMX_OP_EO_CLOVER 2     /* EvenOddPrecCloverLinOp, eoprec_clover_linop_w.cc:142-187           */

#define CMX_SOLVER_CG 0           /* CG_INVERTER: LinOpSysSolverCG + InvCG2 (syssolver_linop_cg.h:57-96) */
#define CMX_SOLVER_BICGSTAB 1     /* BICGSTAB_INVERTER: InvBiCGStab (invbicgstab.cc:10-202)             */
#define CMX_SOLVER_RELIABLE_CG 2  /* RELIABLE_CG_MP_CLOVER_INVERTER: RelInvCG (reliable_cg.cc:10-188)    */
#define CMX_SOLVER_RELIABLE_BICGSTAB 3 /* RELIABLE_BICGSTAB_MP_CLOVER_INVERTER: RelInvBiCGStab
                                          (reliable_bicgstab.cc:14-290)                            */
#define CMX_SOLVER_FGMRESDR 4     /* FGMRESDR_INVERTER: LinOpSysSolverFGMRESDR
                                     (syssolver_linop_fgmres_dr.cc:459-774), right-preconditioned  */

/* Preconditioners of the flexible solver (its PrecondParams, syssolver_fgmres_dr_params.cc:21) */
#define CMX_PRECOND_NONE 0        /* identity                                                       */
#define CMX_PRECOND_MR 1          /* MR_INVERTER: InvMR (invmr.cc:55-160) in the sloppy precision     */
#define CMX_PRECOND_LEARNED 2     /* gauge-equivariant network loaded with cmx_op_set_learned_precond */
#define CMX_PRECOND_MG 3          /* two-level multigrid V-cycle set up with cmx_mg_setup              */

#define CMX_PLUS 1
#define CMX_MINUS (-1)

typedef struct cmx_gauge_s* cmx_gauge;
typedef struct cmx_op_s* cmx_op;

/* ---- lifecycle: initQuda / endQuda hooks (lib/init/chroma_init.cc:215-250, 308-312) ----------- */
int cmx_init(int device);
int cmx_finalize(void);
const char* cmx_last_error(void);
int cmx_version(void);

/* ---- gauge field: replaces loadGaugeQuda (syssolver_linop_clover_quda_w.h:510-523) -------------
 * links: 4*V*18 doubles, [mu][site][3][3][re,im], with the fermion boundary conditions already
 * applied (state->getLinks(), simple_fermbc.h:87-104 / twisted_fermbc_w.h:96-119).  Copied to the
 * device at creation; the host array may be freed afterwards. */
int cmx_gauge_create(const int dims[4], const double* links, cmx_gauge* out);
/* The same from four per-direction arrays, each V*18 doubles [site][3][3][re,im]: the gauge argument
 * of loadGaugeQuda in QUDA_QDP_GAUGE_ORDER, void* gauge[Nd] with gauge[mu] = &links[mu].elem(0)
 * (syssolver_linop_clover_quda_w.h:510-523), so the caller needs no contiguous copy of the field. */
int cmx_gauge_create_dirs(const int dims[4], const double* const links[4], cmx_gauge* out);
int cmx_gauge_destroy(cmx_gauge g);

/* ---- operator: EvenOddPrecCloverLinOp::create (eoprec_clover_linop_w.cc:19-39) ------------------
 * The clover term is built on the device from the links (mesField + makeClov, FP64) and its even
 * block is LDL^dag-inverted in FP64 (choles(0)); loadCloverQuda's role (:536-552). */
typedef struct {
  int kind;            /* CMX_OP_*                                                                  */
  double mass;         /* Mass = 1/(2 Kappa) - Nd  (param_io.cc:12-15)                             */
  double clov_coeff_r; /* CloverFermActParams clovCoeffR (clover_fermact_params_w.cc:65-77)         */
  double clov_coeff_t; /* CloverFermActParams clovCoeffT                                            */
  int sloppy_prec;     /* CMX_SINGLE or CMX_HALF: the inner (sloppy) operator copy                  */
  int sloppy_recon;    /* CMX_RECONS_12 or CMX_RECONS_18                                            */
  int autotune_dslash; /* 1: time the candidate launch shapes at creation and keep the fastest
                          (AutotuneDslash, syssolver_quda_clover_params.cc:93-99)                  */
} cmx_op_params;

int cmx_op_create(cmx_gauge g, const cmx_op_params* p, cmx_op* out);
int cmx_op_destroy(cmx_op op);
/* Number of doubles of one fermion on the operator's subset (Vh*24 for EO, V*24 unpreconditioned) */
long cmx_op_subset_doubles(cmx_op op);

/* chi = M(isign) psi on the operator's subset, computed on the device at precision prec
 * (LinearOperator::operator(), lib/linearop.h:33).  Host FP64 arrays of cmx_op_subset_doubles(). */
int cmx_op_apply(cmx_op op, double* chi, const double* psi, int isign, int prec);

/* Raw hopping term chi_cb = D(isign) psi_{1-cb} (QDPWilsonDslashT::apply, lwldslash_w.h:252-295).
 * chi: Vh*24 doubles of checkerboard cb, psi: Vh*24 doubles of checkerboard 1-cb. */
int cmx_dslash_apply(cmx_op op, double* chi, const double* psi, int isign, int cb, int prec);
/* The hopping term on nrhs sources at once (the 12 spin-colour sources of quarkProp4_a,
 * quarkprop4_w.cc:70-117): chi[k] = D psi[k], consecutive Vh*24-double arrays.  At FP16 storage with
 * 12-real links and real boundary phases each link is loaded once per group of 4 sources; otherwise the
 * single-source kernel runs per source.  Results equal nrhs calls of cmx_dslash_apply. */
int cmx_dslash_multi(cmx_op op, int nrhs, double* chi, const double* psi, int isign, int cb, int prec);

/* Even-odd blocks on full-lattice host arrays (V*24 doubles), FP64
 * (EvenOddPrecLinearOperator, lib/eoprec_linop.h:102-165).  which: 0 evenEven, 1 evenEvenInv,
 * 2 evenOdd, 3 oddEven, 4 oddOdd, 5 unprecLinOp.  Only the target checkerboard of chi is written. */
int cmx_op_eo_piece(cmx_op op, int which, double* chi, const double* psi, int isign);

/* Link derivative of the hopping term (WilsonDslashBase::deriv, lwldslash_base_w.h:132-227), FP64, the
 * fermion-force building block of HMC: for x on checkerboard cb and every mu,
 *   ds_u[mu](x) = traceSpin( [(1 - isign gamma_mu) psi](x+mu) outer chi(x)^dag ),
 * zero on the other checkerboard; cb = -1 gives the sum over both (:110-124).  chi, psi: full-lattice
 * fermions (V*24 doubles); ds_u: 4*V*18 doubles in the link layout (no link factor: the force is
 * taproj(U ds_u) as the reference forms it). */
int cmx_dslash_deriv(cmx_op op, double* ds_u, const double* chi, const double* psi, int isign, int cb);
/* Clover force terms, FP64 on the device, ds_u as for cmx_dslash_deriv ([4][V][3][3] complex, host order):
 *   cmx_clover_deriv      CloverTermBase::deriv (clover_term_base_w.h:691-745): ds of chi^dag A' psi over the
 *                         sites of cb (-1: both), chi and psi full-lattice host arrays (V*24 doubles); the
 *                         term is Hermitian, so isign only selects the reference's overload;
 *   cmx_clover_deriv_trln CloverTermBase::derivTrLn (:825-870) of the inverted even-even term, i.e.
 *                         EvenOddPrecCloverLinOp::derivLogDetEvenEvenLinOp (eoprec_clover_linop_w.cc:225-235);
 *   cmx_op_deriv          EvenOddPrecLogDetLinearOperator::deriv (eoprec_logdet_linop.h:114-157) of the
 *                         even-odd clover operator: ds of chi^dag M'(isign) psi, chi and psi odd-checkerboard
 *                         arrays (Vh*24 doubles).
 * ds is the derivative with respect to the occurrences of U (not U^dag), as WilsonDslashBase::deriv's:
 * for U_mu(y) -> exp(eps X) U_mu(y), d<chi, M psi>/d eps = tr(X U ds(chi, psi, isign)) +
 * conj tr(X U ds(psi, chi, -isign)).  Clover operators only. */
int cmx_clover_deriv(cmx_op op, double* ds_u, const double* chi, const double* psi, int isign, int cb);
int cmx_clover_deriv_trln(cmx_op op, double* ds_u);
int cmx_op_deriv(cmx_op op, double* ds_u, const double* chi, const double* psi, int isign);

/* Clover term download (FP64): which 0 = A on all sites, 1 = A_ee^-1 on the even sites.
 * tri: V*72 doubles in PrimitiveClovTriang layout; sites not covered are left untouched. */
int cmx_clover_download(cmx_op op, int which, double* tri);
/* Sum over the even sites of log|d_i| of the LDL^dag pivots (cholesDet(0), clover_term_qdp_w.h:580-606;
 * logDetEvenEvenLinOp, eoprec_clover_linop_w.cc:310-312) and the count of sites with negative pivots. */
int cmx_clover_logdet(cmx_op op, double* logdet, long* n_neg_sites);
/* QUDAPackedClovSite wire format (clover_term_qdp_w.h:26-32) of checkerboard cb, as packForQUDA
 * (:2171-2228) writes it: out holds V*72 doubles indexed by the full-lattice site (the host order
 * above); the sites of checkerboard cb are written, the others left as they are, so two calls (cb 0,
 * then cb 1) fill the array the way syssolver_linop_clover_quda_w.h:539-548 does.  which 0 = A,
 * 1 = A^-1 (for cb 1 the LDL^dag inverse of A_oo is formed for the call, choles(1)). */
int cmx_clover_pack_quda(cmx_op op, int which, int cb, double* out);

/* ---- solver: replaces invertQuda (syssolver_linop_clover_quda_w.cc:98) --------------------------
 * Solves A psi = chi on the operator's subset (SystemSolver::operator(), lib/syssolver.h:46).
 * psi is the initial guess on entry.  nrhs right-hand sides are consecutive subset vectors.
 * resid is |chi - A psi| recomputed in FP64 on the device (syssolver_linop_cg.h:80-87). */
typedef struct {
  int solver;        /* CMX_SOLVER_*                                                               */
  int max_iter;      /* MaxCG / MaxIter                                                            */
  double rsd_target; /* RsdCG / RsdTarget (relative)                                               */
  double delta;      /* reliable-update Delta (syssolver_rel_bicgstab_clover_params.cc:11-19)      */
  int sloppy_prec;   /* inner precision for the reliable solvers: CMX_SINGLE or CMX_HALF (reliable
                        BiCGStab also CMX_DOUBLE, its pure-double instance reliable_bicgstab.cc:309) */
  int check_every;   /* iterations between host convergence checks (0 = library default)          */
  int verbose;
  int true_residual; /* 0: stop as the reference solver does (CG family: on the normal-equation
                        residual, invcg2.cc:107,194).  1: additionally restart from the current psi
                        with a tightened inner tolerance until |chi - A psi| <= rsd_target |chi| in
                        FP64 (the generalisation of CHROMA_DO_ONE_CG_RESTART, syssolver_linop_cg.h:69-76) */
  /* FGMRES-DR (SysSolverFGMRESDRParams, syssolver_fgmres_dr_params.cc:13-21); the Krylov basis is
     kept in sloppy_prec, the residual is recomputed in FP64 after every cycle */
  int n_krylov;      /* NKrylov >= 1                                                               */
  int n_defl;        /* NDefl < NKrylov, deflated (harmonic Ritz) vectors kept across restarts;
                        NKrylov + NDefl <= 47                                                      */
  int precond;       /* CMX_PRECOND_*                                                              */
  int precond_iters; /* MR: MaxMR                                                                  */
  double precond_rsd;   /* MR: RsdMR                                                               */
  double precond_omega; /* MR: MROver (over-relaxation)                                            */
  int reorth;        /* Gram-Schmidt passes per Arnoldi step: 1 (classical) or 2                   */
  int basis_prec;    /* FGMRES-DR storage of the Krylov bases V, Z (CMX_DOUBLE/SINGLE/HALF); -1: the
                        sloppy precision.  The operator and preconditioner run at sloppy_prec.     */
  double cycle_reduction; /* FGMRES-DR: end a cycle once its residual estimate has fallen by this
                        factor (then the FP64 residual is recomputed and the next cycle starts from
                        it, as a reliable update); 0: only at RsdTarget or NKrylov, as the reference */
  int graph;         /* 1: CG, BiCGStab and reliable CG replay their iterations from captured HIP graphs
                        (fewer launch overheads on small volumes; results are identical)           */
} cmx_solver_params;

typedef struct {
  int n_count;           /* iterations (SystemSolverResults_t::n_count), summed over restarts      */
  double resid;          /* |chi - A psi|, FP64                                                    */
  double rel_resid;      /* resid / |chi|                                                          */
  int n_reliable;        /* reliable updates performed                                             */
  double solve_seconds;  /* device time of the solve (HIP events), fields already resident         */
  double total_seconds;  /* wall time including host<->device transfers                            */
  double gflops;         /* Chroma flop convention (invcg2.cc:67, eoprec_clover_linop_w.cc:299-306) */
  int n_restarts;        /* true_residual restarts performed                                       */
} cmx_solver_result;

int cmx_solve(cmx_op op, const cmx_solver_params* p, double* psi, const double* chi, int nrhs,
              cmx_solver_result* res);

/* MdagMSystemSolver::operator() (lib/syssolver.h:72-84): solves M^dag M psi = chi on the subset, as
 * MdagMSysSolverCG (syssolver_mdagm_cg.h:56-90), MdagMSysSolverBiCGStab (syssolver_mdagm_bicgstab.h:
 * 64-100, two solves: M^dag Y = chi from Y = M psi, then M psi = Y), MdagMSysSolverReliableCGClover
 * (syssolver_mdagm_rel_cg_clover.h:104-150) and MdagMSysSolverReliableBiCGStabClover
 * (syssolver_mdagm_rel_bicgstab_clover.h:104-176).  resid = |chi - M^dag M psi| in FP64; n_count sums
 * both solves of the two-step forms. */
int cmx_solve_mdagm(cmx_op op, const cmx_solver_params* p, double* psi, const double* chi, int nrhs,
                    cmx_solver_result* res);

/* MdagMMultiSystemSolver with CG (multi_syssolver_mdagm_cg.h:58-100 -> MInvCG2, minvcg2.cc:72-355):
 * solves (M^dag M + shifts[s]) psi_s = chi for s < nshift (1..64) with one Krylov space, in FP64 --
 * the rational-approximation solves of RHMC.  Every psi_s starts at zero, as MInvCG2 requires (the
 * entry values of psi are ignored); psi holds nshift consecutive subset vectors.  rsd[s] is the
 * relative target of shift s (RsdCG per shift); shift s stops updating once its shifted residual
 * estimate is below it, and the solve ends when all have (the smallest shift decides the first test).
 * p: max_iter and verbose are used.  res[s]: n_count the common iteration count, resid =
 * |chi - (M^dag M + shift_s) psi_s| recomputed in FP64.  CMX_ERR_NOCONV at MaxCG iterations
 * (QDP_error_exit in the reference). */
int cmx_solve_multishift(cmx_op op, const cmx_solver_params* p, int nshift, const double* shifts, const double* rsd,
                         double* psi, const double* chi, cmx_solver_result* res);

/* ---- batched solves: no single counterpart in the reference -------------------------------------
 * n independent solves advance together: every kernel of an iteration is launched once over a grid
 * that covers all n problems (work-group -> problem, work-group of that problem), so small volumes
 * share launches and fill the device.  Problem i solves A_i psi_i = chi_i (mdagm: M_i^dag M_i psi_i =
 * chi_i) on ops[i]; operators may differ in lattice, gauge field and mass, and may repeat (the 12
 * sources of quarkProp4_a, quarkprop4_w.cc:70-117, on one operator).  Solver CMX_SOLVER_CG or
 * CMX_SOLVER_RELIABLE_CG; each problem's iterations, stopping rule and result are those of cmx_solve
 * with the same launch shapes.  Solver CMX_SOLVER_FGMRESDR (any preconditioner; a repeated operator's
 * solves share its network): the solves advance one Arnoldi step per tick with one host read-back for
 * all of them, each with the kernels and host arithmetic of cmx_solve, so each result equals cmx_solve's
 * bit for bit; as many run at once as their workspaces fit in free device memory.  psi[i] is the initial
 * guess on entry; res[i].solve_seconds is the device time of the whole batch. */
typedef struct {
  int tune;   /* 1: time candidate launch shapes for the batch (plane-grid or site-linear hopping
                 kernel, 64/128/256 threads) and keep the fastest; 0: the operators' own shapes, which
                 must select the same kernel on every problem */
  int graph;  /* 1: replay the batched iterations from a captured HIP graph */
  int max_concurrent; /* FGMRES-DR: at most this many solves in flight (0: as many as fit in free device
                         memory); a waiting solve takes over the context of a finished one on the same
                         operator */
} cmx_batch_params;
typedef struct {
  double solve_seconds;  /* device time of the batched iterations (HIP events)                   */
  double total_seconds;  /* wall time of the call                                                 */
  int n_iter;            /* batched iterations launched                                           */
  int n_stages;          /* kernels per batched iteration                                         */
  long blocks_per_iter;  /* work-groups per batched iteration                                     */
  int hop_block;         /* hopping-kernel launch shape used: threads per work-group              */
  int hop_fast;          /* 0 site-linear, 1 plane-grid; +2 with streamed read-once loads        */
} cmx_batch_result;
int cmx_solve_batch(const cmx_op* ops, int n, const cmx_solver_params* p, const cmx_batch_params* bp,
                    double* const* psi, const double* const* chi, int mdagm, cmx_solver_result* res,
                    cmx_batch_result* bres);

/* ---- learned preconditioner (CMX_PRECOND_LEARNED): no counterpart in the reference ------------------
 * A gauge-equivariant parallel-transport network on the odd checkerboard, a chain of 1..4 blocks
 * P = B_{n-1} o ... o B_0, each
 *   B(v)(x) = W0 v(x) + sum_q U_q(x) Pi_q g_q(x+q),  g_q(y) = sum_c W2[q][c] sum_p W1[c][p] U_p(y) Pi_p v(y+p),
 * with one-hop paths p, q in {+-mu}, their spin projectors Pi = (1 -+ gamma_mu)/2, 16 hidden channels and
 * 4x4 complex spin matrices W0, W1, W2 (colour identity).  Every block carries its own transports, so the
 * receptive field grows by two hops per block.  It is the right preconditioner of CMX_SOLVER_FGMRESDR (the
 * slot of PrecondParams, syssolver_linop_fgmres_dr.h:175-185).  A network is independent of the gauge
 * field: train it on some operators, attach it to others. */
typedef struct cmx_net_s* cmx_net;
int cmx_net_create(cmx_net* out);  /* one block */
int cmx_net_destroy(cmx_net net);
/* 1..4 blocks; blocks added start as the identity (W0 = 1, W2 = 0, W1 channel c < 8 carrying path c) */
int cmx_net_set_depth(cmx_net net, int nblocks);
int cmx_net_depth(cmx_net net, int* nblocks);
/* block 0: W0 = 1/diag, W2 W1 = (1/(4 diag^3)) (1 -+ gamma_q)(1 -+ gamma_p), the two-term Neumann series
 * of the even-odd operator with the clover term replaced by its diagonal diag = Nd + m; the other blocks
 * the identity */
int cmx_net_init_neumann(cmx_net net, double diag);
/* block weights as interleaved (re, im) doubles: W1[(c*4+s)*32 + p*4+t] (64 x 32), W2[(q*4+s)*64 + c*4+t]
 * (32 x 64), W0[s*4+t] (4 x 4); path index p = 2 mu + (0 forward, 1 backward).  get/set_weights: block 0 */
int cmx_net_get_block(cmx_net net, int blk, double* w1, double* w2, double* w0);
int cmx_net_set_block(cmx_net net, int blk, const double* w1, const double* w2, const double* w0);
int cmx_net_get_weights(cmx_net net, double* w1, double* w2, double* w0);
int cmx_net_set_weights(cmx_net net, const double* w1, const double* w2, const double* w0);
int cmx_op_set_net(cmx_op op, cmx_net net);
/* kernel path of the network at precision prec: 0 fused gather and mix (MFMA operands in registers),
 * 1 fused (transports staged in LDS), 2 separate gather and mix kernels, 3 and 4 as 0 and 1 with the links
 * of the gather and the scatter streamed (non-temporal loads); -1 time all five at the next application
 * and keep the fastest (the default).  The paths agree to the network's BF16 precision but not bit for
 * bit (the compiler contracts their FP32 arithmetic differently); each is deterministic. */
int cmx_op_set_net_variant(cmx_op op, int prec, int variant);
/* out = P(in) on the odd checkerboard at storage precision prec (host arrays, Vh*24 doubles) */
int cmx_net_apply(cmx_op op, int prec, double* out, const double* in);
/* loss = |P(v) - e|^2; per block b (nblocks of them, in order): g_wc[b] = d loss / d conj(W2 W1) of
 * block b (32 x 32 complex), g_w0[b] = d loss / d conj(W0) of block b (4 x 4 complex) */
int cmx_net_grad(cmx_op op, int prec, const double* v, const double* e, double* loss, double* g_wc, double* g_w0);
typedef struct {
  int n_sources;      /* Gaussian sources solved by plain FP64 CGNE                                 */
  int snapshot_every; /* CG iterations between (residual, error) pairs                              */
  double solve_rsd;   /* CGNE target defining the "exact" solution x*                               */
  int epochs;         /* passes over the pairs (one Adam step per pair on |M (P r - e)|^2 / |r|^2)  */
  double lr;          /* Adam step size                                                             */
  unsigned long long seed;
  int max_pairs;
  int prec;           /* storage precision of the pairs and the network input: CMX_SINGLE/CMX_HALF  */
  int verbose;
  int pairs_per_source; /* cap on the pairs taken from one source's iterates (0: no cap)          */
  int normalize;        /* 1: each pair's loss divided by |r|^2; 0: raw |M (P r - e)|^2, so early,
                           large residuals weigh more                                               */
} cmx_train_params;
typedef struct {
  int n_pairs, n_steps;
  double loss_first, loss_last; /* the objective as optimised (mean weighted loss over the pairs)
                                   before and after training                                     */
  double gen_seconds, train_seconds;
} cmx_train_result;
int cmx_net_train(cmx_op op, const cmx_train_params* p, cmx_train_result* res);

/* ---- multigrid (CMX_PRECOND_MG) ----------------------------------------------------------------------
 * The two-level form of MULTIGRIDSolverParams (quda_multigrid_params.h:12-113, read by
 * quda_multigrid_params.cc:44-235), behind the same role as QUDA_MULTIGRID_CLOVER_INVERTER's solver: the
 * right preconditioner of CMX_SOLVER_FGMRESDR (the outer flexible solver, OuterGCRNKrylov ~ NKrylov).
 * Null vectors: nvec random starts relaxed by setup_maxit MR steps on A x = 0 (A the even-odd operator,
 * at precision prec), extended to the full lattice by the even-odd reconstruction, block-orthonormalised per block and
 * chirality.  Coarse operator: the Galerkin product P^dag M P of the unpreconditioned operator M (FP64
 * probes, FP32 storage), a 9-point block stencil, 2 nvec degrees of freedom per block.  V-cycle on the
 * odd checkerboard, A^-1 = (M^-1)_oo: nu_pre MR steps, restriction of the odd residual, GCR(coarse_nkrylov)
 * on the coarse system to coarse_tol or coarse_maxit steps, prolongation, nu_post MR steps; smoother and
 * V-cycle at the solver's sloppy precision.  Defaults (cmx_mg_params_default) follow the reader's. */
typedef struct {
  int blocking[4];       /* Blocking/elem (first level): even extents dividing the lattice (4 4 4 4)    */
  int nvec;              /* NullVectors (first level), 1..16 (16)                                       */
  int nu_pre, nu_post;   /* Pre-SmootherApplications, Post-SmootherApplications (2, 2)                  */
  double omega;          /* RelaxationOmegaMG: the MR smoother's over-relaxation (1.0)                 */
  double coarse_tol;     /* CoarseResidual (coarse level): relative residual of the coarse GCR (1e-4)   */
  int coarse_maxit;      /* MaxCoarseIterations (coarse level) (12)                                     */
  int coarse_nkrylov;    /* PrecondGCRNKrylov: GCR directions before a restart, <= 32 (12)             */
  int setup_maxit;       /* MaxIterSubspaceCreate: MR relaxation steps per null vector (500)           */
  unsigned long long seed;
  int verbose;           /* Verbosity                                                                    */
  int prec;              /* Precision: storage precision of the null-space relaxation, CMX_SINGLE        */
                         /* (default) or CMX_HALF (half the bytes per MR step) or CMX_DOUBLE              */
  int smoother;          /* post-smoother: 0 nu_post MR steps (default); 1 the operator's learned network */
                         /* (cmx_net_attach) applied once to the residual of the coarse correction, then */
                         /* nu_post MR steps.  No counterpart in the reference.                           */
} cmx_mg_params;
typedef struct {
  int n_blocks, coarse_dof;        /* coarse sites, degrees of freedom per coarse site (2 nvec)       */
  int n_probes;                    /* operator applications that built the coarse operator            */
  double setup_seconds, nullspace_seconds, coarse_op_seconds;
  double nullvec_rel_residual_max; /* max over the null vectors of |A x_o| / |x_o| after relaxation    */
} cmx_mg_result;
int cmx_mg_params_default(cmx_mg_params* p);
int cmx_mg_setup(cmx_op op, const cmx_mg_params* p, cmx_mg_result* res);
/* Changes the V-cycle of a set-up multigrid without rebuilding it: nu_pre, nu_post, omega, coarse_tol,
 * coarse_maxit, coarse_nkrylov, smoother, verbose (the role of updateMultigridQuda, which Chroma's QUDA multigrid
 * solver calls when it recovers a saved subspace, syssolver_linop_clover_quda_multigrid_w.h:500-508).  blocking and
 * nvec must equal the set-up's (CMX_ERR_ARG otherwise); setup_maxit, seed and prec are ignored. */
int cmx_mg_update(cmx_op op, const cmx_mg_params* p);
/* Inspection and test access: the null vectors after block orthonormalisation (nvec * V * 24 doubles,
 * host order), the coarse operator (n_blocks * 9 * D * D complex: block, stencil entry -- the block,
 * then +x, -x, +y, -y, +z, -z, +t, -t --, row, column), restriction R = P^dag of a full-lattice field
 * (V * 24 doubles -> n_blocks * D complex), prolongation (the reverse) and the coarse operator. */
int cmx_mg_get_nullvecs(cmx_op op, double* out);
int cmx_mg_get_coarse_op(cmx_op op, double* out);
int cmx_mg_restrict(cmx_op op, const double* full_in, double* coarse_out);
int cmx_mg_prolong(cmx_op op, const double* coarse_in, double* full_out);
int cmx_mg_coarse_apply(cmx_op op, const double* in, double* out);

/* ---- launch shape of the hopping kernels ---------------------------------------------------------
 * cfg = {threads per work-group (64, 128, 256), tiled work-group order (0/1), z planes per tile,
 * double-buffered direction loads (0/1), kernel (0 site-linear, 1 plane-grid -- 12-real links, real
 * boundary phases, fields below 2 GiB; +2 with the loads of data read once per launch -- links,
 * clover, xpay input -- streamed non-temporally), target (0 both, 1 the EO operator's fused kernels,
 * 2 the plain hopping term; for get, 2 reads the plain-hop shape and 0/1 the operator's)}.  No counterpart in the
 * reference beyond QUDA's AutotuneDslash flag (syssolver_quda_clover_params.cc:93-99). */
int cmx_op_get_hop_config(cmx_op op, int prec, int cfg[6]);
int cmx_op_set_hop_config(cmx_op op, int prec, const int cfg[6]);
int cmx_op_autotune(cmx_op op);
/* Reals stored per link at precision prec (12 or 18).  FP64 links are stored as 12 reals when every
 * link is SU(3) times its boundary phase to 1e-14, else as 18; FP32 follows cmx_op_params.sloppy_recon;
 * FP16 is always 12.  Like QUDA's cuda_reconstruct (syssolver_quda_clover_params.cc:71-88). */
int cmx_op_link_recon(cmx_op op, int prec, int* recon);
/* Kernel of cmx_dslash_multi: 1..4 sources per lane on the vector units (the links are loaded once per
 * lane for all of them), 5 the link products of all sources on the matrix cores (one site per 4 lanes,
 * v_mfma_f32_4x4x4_16b_f16), 0 timed at first use (default).  A tuning knob like
 * cmx_op_set_hop_config; no counterpart in the reference. */
int cmx_op_set_multi_group(cmx_op op, int group);

/* ---- measurement -------------------------------------------------------------------------------
 * Average device milliseconds per call over n back-to-back launches on the operator's stream,
 * bracketed by HIP events.  what: 0 hopping term (cb 1), 1 EO operator kernel pair (PLUS),
 * 2 fused EO kernel 2 only, 3 streaming copy of one odd fermion (bandwidth reference), 4 the hopping
 * term on 12 sources (cmx_dslash_multi; one call covers all 12), 5 a streaming kernel with the hopping
 * term's bytes and read/write mix (five read streams, one write stream of 16-byte elements, together
 * FP16 296, FP32 576, FP64 1152 B per site: the rate the hop could reach without its gathers). */
int cmx_op_time(cmx_op op, int what, int prec, int n, double* ms_per_call);

/* ---- layout ------------------------------------------------------------------------------------
 * Uploads one checkerboard fermion (Vh*24 doubles) into the device SoA layout at precision prec and
 * downloads it back.  For CMX_DOUBLE the result is bit-identical to the input. */
int cmx_layout_roundtrip(cmx_op op, int prec, const double* in, double* out);

/* ---- synthetic inputs (host): counter-based RNG keyed by (seed, site, mu, component) -------------
 * kind 0: unit links; 1: U = exp(i eps H), H traceless Hermitian Gaussian, then reunitarised;
 * 2: Haar-random SU(3).  Then nstout isotropic stout steps with parameter rho
 * (stout_utils.cc:105-180,936-995).  links: 4*V*18 doubles (host layout). */
int cmx_gauge_generate(const int dims[4], int kind, unsigned long long seed, double eps, int nstout,
                       double rho, double* links);
/* Fermion boundary conditions applied in place: SimpleFermBC::modify (simple_fermbc.h:87-104) with
 * boundary[mu] = +1/-1, then TwistedFermBC::modify (twisted_fermbc_w.h:96-119) when twist != NULL:
 * links in direction twist_dir[i] times exp(i pi twist_by_pi[i] / L[dir]). */
int cmx_gauge_apply_bc(const int dims[4], double* links, const int boundary[4], const double* twist_by_pi,
                       const int* twist_dir);
/* Isotropic stout smearing on the device, in place: nstep applications of Stouting::smear_links
 * (stout_utils.cc:936-962) with every direction smeared and rho(mu,nu) = rho; staple sum and Q as in
 * getQsandCs (:105-180), exp(iQ) to rounding, each new link re-projected onto SU(3) at rounding level.
 * links: 4*V*18 doubles (host layout).  Needs cmx_init; FP64 throughout. */
int cmx_gauge_stout_smear(const int dims[4], double* links, int nstep, double rho);
/* The same over a subset of directions, StoutLinkSmearing's smear_dirs / no_smear_dir
 * (stout_link_smearing.cc:62-120): bit mu of smear_dirs set = direction mu smeared.  Links of the other
 * directions are left as they are and contribute no staples (rho(mu,nu) = 0 when either is unsmeared);
 * 15 is cmx_gauge_stout_smear, 7 the usual spatial smearing with no_smear_dir = 3. */
int cmx_gauge_stout_smear_dirs(const int dims[4], double* links, int nstep, double rho, int smear_dirs);
/* Meson two-point functions on the device: mesons() (mesons_w.cc:97-180) for the 16 insertions
 * Gamma(n) = gamma_1^n0 gamma_2^n1 gamma_3^n2 gamma_4^n3 with the sink-momentum projection of
 * SftMom::sft (sftmom.cc:405-460), decay direction 3:
 *   corr[n][m][t_eff] = sum_{x in slice t} exp(i 2 pi sum_k moms[3m+k] (x_k - origin[k]) / L_k)
 *                       tr[adj(gamma_5 prop2(x) gamma_5) Gamma(n) prop1(x) Gamma(n)],
 * t_eff = (t - origin[3]) mod L_t.  prop1, prop2: V*144 complex (2*V*144 doubles), site in the host
 * order, then sink spin*3+colour, then source spin*3+colour; prop2 may equal prop1.  corr: 16*nmom*L_t
 * complex (mesons() writes the real part).  Needs cmx_init; FP64 throughout. */
int cmx_meson_2pt(const int dims[4], const double* prop1, const double* prop2, const int origin[4], int nmom,
                  const int* moms, double* corr);
/* The same contraction without allocating: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * prop1, prop2, moms (3*nmom ints) and corr are DEVICE pointers with the layouts of cmx_meson_2pt; the
 * kernels are queued on `stream` (a hipStream_t, NULL for the null stream) and the call returns without
 * waiting, so corr is valid once the stream has drained.  The work space is a fixed 64 MiB device array
 * of the library (per GPU, held until the process ends): the spatial volume Lx*Ly*Lz must not exceed
 * 262144 sites (64^3); larger lattices are rejected with CMX_ERR_ARG.  Calls that overlap on one GPU share
 * that array and race; serialise them.  The results equal cmx_meson_2pt's bit for bit. */
int cmx_meson_2pt_noalloc(const int dims[4], const double* prop1, const double* prop2, const int origin[4],
                          int nmom, const int* moms, double* corr, void* stream);
/* SciDAC single-file output (the file QIO_WRITE_NAMED_OBJECT writes through QDPFileWriter with
 * QDPIO_SINGLEFILE, inline_qio_write_obj.cc:150-180, qio_write_obj_funcmap.cc:30-100) and input (the
 * QIO_READ_NAMED_OBJECT path).  Host-only: no device work, no cmx_init needed.  kind 0: a propagator
 * (field: V*144 complex, sites in the host order, then sink spin*3+colour, then source spin*3+colour),
 * datatype QDP_D3_DiracPropagator / QDP_F3_DiracPropagator; kind 1: a fermion (V*12 complex,
 * QDP_D3_DiracFermion / QDP_F3_DiracFermion).  prec CMX_DOUBLE or CMX_SINGLE (the file's
 * precision; the host array is always double).  The LIME records: scidac-private-file-xml,
 * scidac-file-xml (file_xml), scidac-private-record-xml, scidac-record-xml (record_xml),
 * scidac-binary-data (lexicographic sites, big-endian), scidac-checksum (the SciDAC CRC-32 sums).
 * The reader checks dims, datatype and checksum, and returns the user XML records truncated to the
 * capacities given (NULL: not returned). */
int cmx_field_write_scidac(const char* path, const int dims[4], int kind, int prec, const double* field,
                           const char* file_xml, const char* record_xml);
int cmx_field_read_scidac(const char* path, const int dims[4], int kind, double* field, char* file_xml,
                          int file_xml_cap, char* record_xml, int record_xml_cap);
/* Average plaquette Re tr(U_p)/3 over all sites and planes (MesPlq normalisation). */
double cmx_plaquette(const int dims[4], const double* links);
/* Z2 noise source: each real and imaginary part +-1/sqrt(2) from (seed, site, component). */
int cmx_fill_z2(long n_doubles, unsigned long long seed, double* out);

#ifdef __cplusplus
}
#endif
#endif /* CHROMA_MI355X_H */
