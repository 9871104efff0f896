// Device-resident operator state behind the cmx_gauge / cmx_op handles.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "blas.h"
#include "cmx_internal.h"
#include "learned.h"

namespace cmx {

// Owning device allocation
// Owning device allocation (or, after alias(), a non-owning view of another one)
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool own = true;
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), own(o.own) { o.p = nullptr, o.bytes = 0, o.own = true; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p, bytes = o.bytes, own = o.own;
      o.p = nullptr, o.bytes = 0, o.own = true;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n) {
    release();
    if (n) CMX_HIP(hipMalloc(&p, n));
    bytes = n;
  }
  // refers to o's memory without owning it (o must outlive this buffer)
  void alias(const DevBuf& o) {
    release();
    p = o.p, bytes = o.bytes, own = false;
  }
  void release() {
    if (p && own) (void)hipFree(p);
    p = nullptr, bytes = 0, own = true;
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// A fermion vector over n sites (one checkerboard, or a full lattice stored even-then-odd)
struct Vec {
  DevBuf buf;
  int prec = P64;
  int n = 0;
  Vec() = default;
  Vec(int prec_, int n_) : prec(prec_), n(n_) { buf.alloc(spinor_bytes(prec_, n_)); }
  float* nrm() const {
    return prec == P16 ? reinterpret_cast<float*>(buf.as<char>() + (size_t)spinor_chunks(prec) * 16 * n) : nullptr;
  }
  // view of the sites [off, off + count) with this vector's stride
  VRef view(int off = 0) const {
    VRef r;
    r.v = buf.as<char>() + (size_t)off * 16;
    r.nrm = prec == P16 ? nrm() + off : nullptr;
    r.stride = n;
    return r;
  }
};

struct GaugeData {
  Geom g;
  long V = 0, Vh = 0;
  DevBuf g64[2];           // FP64, per parity [mu][9 x double2][Vh]
  double ph[4][2][2] = {}; // det(U) per direction: [bulk | last slice][re, im]
  double su3_dev = 0;      // max deviation from (phase x SU(3)) over all links
};

struct Op {
  std::shared_ptr<GaugeData> gd;
  cmx_op_params par{};
  Geom g;
  long V = 0, Vh = 0;
  double fact = 0;  // Nd + Mass
  int recon[3] = {18, 12, 12};
  // hopping-kernel launch shape per precision: [0] the EO operator's fused kernels, [1] the plain
  // hopping term (tuned separately by autotune(): their best shapes differ)
  HopConfig hcfg[3][2];
  int multi_group = 0;  // sources per multi-source hop kernel; 0: timed at first use
  // per precision: links (per parity) and the clover blocks used by the EO operator
  DevBuf gauge[3][2];
  DevBuf clov_odd[3];     // A_oo
  DevBuf clovinv_even[3]; // A_ee^-1
  DevBuf clov64_even;     // A_ee (FP64, for evenEvenLinOp / unprecLinOp)
  double logdet_even = 0;
  long nneg_even = 0;
  hipStream_t stream = nullptr;
  bool shared_stream = false;  // a clone: the stream and operator fields belong to another Op
  DevBuf partials;
  int npartials = 0;
  DevBuf state_S, state_F;
  double* hS = nullptr;  // pinned mirrors
  int* hF = nullptr;
  DevBuf staging;        // V*24 doubles
  std::vector<std::unique_ptr<Vec>> tmp_even[3];  // scratch even-checkerboard vectors
  // solver workspace kept across solves (so that captured iteration graphs stay valid), and the
  // captured graphs: key = (solver tag, batch length, solution and source buffers); an entry is
  // re-captured when the workspace generation changed
  std::vector<std::unique_ptr<Vec>> ws;
  long ws_gen = 0;
  struct GraphEntry {
    long tag, batch, gen;
    const void *psi, *chi;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;
  DevBuf kry_part;  // reduction partials of the Krylov-basis kernels (FGMRES-DR)
  Vec& work(int slot, int prec, int n);
  // learned preconditioner: the attached network, the weight version last uploaded, device buffers
  std::shared_ptr<LearnedNet> net;
  long net_version = -1;
  std::shared_ptr<void> ptc;
  // multigrid preconditioner state (mg.cpp)
  std::shared_ptr<void> mg;

  ~Op();
  bool eo() const { return par.kind != CMX_OP_UNPREC_WILSON; }
  bool clover() const { return par.kind == CMX_OP_EO_CLOVER; }
  int subset_sites() const { return eo() ? (int)Vh : (int)V; }
  void ensure_prec(int prec);
  Vec& tmp(int prec, int k);
  HopLaunch hop_base(int prec, int which = 0) const;

  // chi = M(isign) psi on the subset; partials (optional) receives |chi|^2 per work-group, or with ndot
  // (even-odd operators) |chi|^2, Re <chi, psi>, Im <chi, psi> per work-group (three partials each);
  // returns the number of work-groups whose partials were written.
  int apply(int prec, VRef out, VRef in, int isign, double* partials, const int* skip = nullptr,
            const int* run = nullptr, bool ndot = false);
  // raw hopping term on output checkerboard cb (views of one checkerboard each), scaled by b
  void dslash(int prec, VRef out, VRef in, int isign, int cb, double b = 1.0);
  // the same on n sources at once (FP16 storage: links loaded once per group of sources); false when
  // the multi-source kernel does not apply (then the caller uses dslash per source)
  bool dslash_multi(int prec, const VRef* out, const VRef* in, int n, int isign, int cb, double b = 1.0);
  // stage a host array (n doubles) into the device staging buffer and back
  void upload(const double* host, size_t n);
  void download(double* host, size_t n);
  // host fermion (n sites x 24 doubles) <-> device vector view
  void pack(const double* host, int nsites, const Vec& v, int off = 0);
  void unpack(double* host, int nsites, const Vec& v, int off = 0);
  // BLAS helper
  void blas(int op, std::initializer_list<const Vec*> vs, int ia = 0, int ib = 0, int skip = -1, int run = -1,
            int n = -1, int off = 0);
  void finalize(int nparts, int nred, int prog, int slot = -1);
  void sync();
  // Times the candidate launch shapes of the EO operator at each precision in use and keeps the
  // fastest (QUDA's AutotuneDslash role).  Results are cached per (lattice, kind, precision).
  void autotune();
};

std::shared_ptr<GaugeData> make_gauge(const int dims[4], const double* const links[4]);
Op* make_op(std::shared_ptr<GaugeData> gd, const cmx_op_params& p);
// A second solver context on the operator o: shares o's links, clover fields, staging buffer and
// stream (o must outlive it) and owns its solver state and workspace.  o's fields for precision
// prec are created first.
Op* clone_op(Op& o, int prec);

// Solvers (solvers.cpp)
// Learned preconditioner (learned.cpp)
void net_init_neumann(LearnedNet& net, double diag);
// 1..4 blocks; blocks added start as the identity
void net_set_depth(LearnedNet& net, int nblocks);
void net_apply_host(Op& op, int prec, double* out, const double* in);
void net_set_variant(Op& op, int prec, int variant);
void net_grad_host(Op& op, int prec, const double* v, const double* e, double* loss, double* gwc, double* gw0);
void net_train(Op& op, const cmx_train_params& tp, cmx_train_result& res);
// out = P(in): the attached network applied to one vector at precision prec (the learned preconditioner)
void learned_precond_apply(Op& op, int prec, const Vec& in, Vec& out);
// P(in[i]) -> out[i] for several sources: one launch per kernel and block for all of them, each source's
// result equal to learned_precond_apply's bit for bit (fgmres uses it for the batched solves)
void learned_precond_apply_batch(Op& op, int prec, const std::vector<const Vec*>& in, const std::vector<Vec*>& out);
// `to` (a solver context on the same operator, clone_op) uses from's network and its device buffers; the
// two must issue their work on one stream
void learned_share(Op& from, Op& to);

// mdagm: solve M^dag M psi = chi instead of M psi = chi
void solve(Op& op, const cmx_solver_params& p, double* psi, const double* chi, bool mdagm, cmx_solver_result& res);
// WilsonDslashBase::deriv (lwldslash_base_w.h:132-227) for the sites of cb, FP64 (deriv.hip): chi on cb,
// psi on 1 - cb, ds = 4 * Vh * 9 double2 (direction, site of cb, 3x3 row-major)
void hop_deriv(const Geom& g, int cb, int isign, VRef chi, VRef psi, void* ds, hipStream_t s);
// Clover force (force.hip), FP64, accumulated into ds_full = [4][V][9] double2 (host site order) times
// scale: trln = false: CloverTermBase::deriv on cb from chi, psi (views of cb); trln = true: derivTrLn
// of the inverted even-even term (cb = 0; chi, psi unused).  force_axpy adds f times a hopping-term
// derivative of one checkerboard (hop_deriv's layout) into ds_full.
void clover_force(Op& op, bool trln, int cb, VRef chi, VRef psi, double scale, void* ds_full);
void force_axpy(Op& op, int cb, double f, const void* ds_cb, void* ds_full);
// nstep stout steps on the device over the directions in the bit mask dirs, in place on host-layout
// links (stout.hip)
void stout_smear(const int dims[4], double* links, int nstep, double rho, int dirs = 15);
// meson two-point functions of two host-layout propagators for the 16 gamma insertions (meson.hip)
void meson_2pt(const int dims[4], const double* prop1, const double* prop2, const int origin[4], int nmom,
               const int* moms, double* corr);
// the same with device arrays and no allocation: the work space is a fixed device array of meson.hip,
// set up once per device (meson_scratch_setup, with that device current)
void meson_scratch_setup();
void meson_2pt_noalloc(const int dims[4], const double* prop1, const double* prop2, const int origin[4], int nmom,
                       const int* moms, double* corr, hipStream_t stream);
// SciDAC single-file fields (qio.cpp, host only): kind 0 propagator, 1 fermion; prec P64 / P32 in the file
void scidac_write(const char* path, const int dims[4], int kind, int prec, const double* host, const char* file_xml,
                  const char* record_xml);
void scidac_read(const char* path, const int dims[4], int kind, double* host, std::string* file_xml,
                 std::string* record_xml);
// MInvCG2 (minvcg2.cc:72-355): (M^dag M + shifts[s]) psi[s] = chi, s < ns, FP64 (multishift.cpp)
void solve_multishift(Op& op, const cmx_solver_params& p, int ns, const double* shifts, const double* rsd,
                      double* psi, const double* chi, cmx_solver_result* res);
// The same for n problems at once (batch.h): problem i solves on ops[i] (distinct solver contexts,
// one stream); solver RELIABLE_CG or CG
void solve_batch(const std::vector<Op*>& ops, const cmx_solver_params& p, const cmx_batch_params& bp,
                 double* const* psi, const double* const* chi, bool mdagm, cmx_solver_result* res,
                 cmx_batch_result& bres);

}  // namespace cmx
