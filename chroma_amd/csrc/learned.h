// Gauge-equivariant parallel-transport network (the learned preconditioner): weights, device
// buffers and kernel launchers.  The network and its kernels are described in learned.hip.
#pragma once
#include <complex>
#include <vector>

#include "cmx_internal.h"

namespace cmx {

constexpr int PTC_CH = 16;    // hidden channels
constexpr int PTC_PATHS = 8;  // one-hop paths +-mu

struct PtcLaunch {
  Geom g;
  int prec, recon;
  const void* gauge[2];  // links per parity in the operator's layout for `prec`
  double ph[4][2][2];    // det(U) per direction (bulk / last slice) for 12-real links
  const void* in;        // odd-checkerboard input (gather: v or delta; scatter: v; w0 grad: v)
  const float* in_nrm;
  int in_stride;
  void* out;             // odd-checkerboard output (scatter) / delta (w0 grad)
  float* out_nrm;
  int out_stride;
  unsigned short* T;     // transported half spinors, BF16 MFMA tiles (Vh/32 x 3 x 2 KiB)
  unsigned short* dT;    // transported loss gradient, same layout
  unsigned short* G;     // mixed field gh_q (half spinors), BF16 (8 x 3 x Vh x 8 B)
  const unsigned short* Afr;  // real embedding of Wh in MFMA A-operand order (4 x 64 x 4 BF16)
  const float* W0;       // 4x4 complex, row-major (re, im)
  float* gpart;          // gradient partials (waves x 1024 floats)
  double* w0part;        // W0 gradient partials (blocks x 32 doubles)
};

// Several sources through one block in one launch per kernel (inference only): per-source input,
// output and mixed field G; strides and everything else from the PtcLaunch
constexpr int PTC_MAXSRC = 16;
struct PtcBatch {
  int ns;
  const void* in[PTC_MAXSRC];
  const float* in_nrm[PTC_MAXSRC];
  void* out[PTC_MAXSRC];
  float* out_nrm[PTC_MAXSRC];
  unsigned short* G[PTC_MAXSRC];
};

int ptc_waves();
void launch_ptc_gather(const PtcLaunch& L, bool adj, hipStream_t s);
void launch_ptc_mix(const PtcLaunch& L, hipStream_t s);
// inference: variant 0 operands in registers, 1 transports staged in LDS; 3, 4 the same with the links
// streamed (non-temporal loads)
void launch_ptc_gather_mix(const PtcLaunch& L, int variant, hipStream_t s);
// variant 0 or 1 of launch_ptc_gather_mix, and the scatter, on B.ns sources
void launch_ptc_gather_mix_batch(const PtcLaunch& L, const PtcBatch& B, int variant, hipStream_t s);
void launch_ptc_scatter_batch(const PtcLaunch& L, const PtcBatch& B, bool nt, hipStream_t s);
// adj: the block's adjoint W0^dag delta + G^dag(L.G) (L.W0 holding W0^dag, L.G the mixed adjoint)
// nt: the links streamed (inference variants 3, 4)
void launch_ptc_scatter(const PtcLaunch& L, bool adj, hipStream_t s, bool nt = false);
void launch_ptc_grad(const PtcLaunch& L, double* out1024, hipStream_t s);
int launch_w0_grad(const PtcLaunch& L, hipStream_t s);

constexpr int PTC_MAX_BLOCKS = 4;

// One block of the network (host, FP32 complex) with its Adam state.  Component indices: path/channel-
// major, spin-minor: W1[(c*4 + s) * 32 + (p*4 + t)], W2[(q*4 + s) * 64 + (c*4 + t)], W0[s*4 + t].
//   B(v)(x) = W0 v(x) + sum_q U_q(x) Pi_q g_q(x+q),  g_q(y) = sum_p Wc[q][p] U_p(y) (Pi_p v)(y+p),  Wc = W2 W1
struct PtcBlock {
  using cf = std::complex<float>;
  std::vector<cf> W1, W2, W0;
  std::vector<float> m1, v1, m2, v2, m0, v0;  // Adam moments over (re, im)
  PtcBlock();
  // Wc = W2 W1 (32 x 32 complex), row-major over (q*4+s, p*4+t); the network applies it through the
  // half-spinor projectors of the paths (Pi_q Wc Pi_p)
  std::vector<cf> contracted() const;
  // B = 1: W0 the identity, W2 = 0, and W1 channel c < 8 carrying path c with its projector (so that the
  // gradient reaches W2 from the start)
  void set_identity();
  void reset_moments();
};

// The network: blocks applied one after another, P = B_{n-1} o ... o B_1 o B_0.  Every block carries
// its own transports, so the receptive field grows by two hops per block.
struct LearnedNet {
  std::vector<PtcBlock> blk;
  long step = 0;
  long version = 0;  // bumped on every weight change; operators re-upload when it differs
  LearnedNet() : blk(1) {}
};

}  // namespace cmx
