// Gauge-equivariant parallel-transport network used as the right preconditioner of the flexible
// solver.  On odd-checkerboard vectors v:
//
//   P(v)(x) = W0 v(x) + sum_q U_q(x) g_q(x+q),                       x odd,  q in {+-mu}
//   g_q(y)  = sum_{c<16} W2[q][c] h_c(y),  h_c(y) = sum_p W1[c][p] U_p(y) v(y+p),   y even
//
// U_{+mu}(y) = U_mu(y), U_{-mu}(y) = U_mu(y-mu)^dag (parallel transport to y); W0, W1, W2 are 4x4 complex
// spin matrices acting identically on colour, so P commutes with gauge transformations.  The two
// linear channel layers contract to one 32x32 complex matrix Wc = W2 W1 over (path, spin), applied on
// MFMA in BF16 with FP32 accumulation; the parallel transports run on the vector units.
//
// Kernels:
//   k_ptc_gather   (even sites)  t_p(y) = U_p(y) v(y+p) for the 8 paths, stored in BF16 tiles laid out as
//                  the B operand of the 32x32x8 MFMA (or, with ADJ, the transport of the loss gradient
//                  dg_q(y) = t_{-q}[delta](y), the adjoint of the scatter)
//   k_ptc_mix      (MFMA)        g = Wc t per (site, colour) column
//   k_ptc_gather_mix             both of the above in one kernel, transports packed straight into the
//                                MFMA operands (inference path)
//   k_ptc_gather_mix_lds         the same with the transports staged in LDS tiles
//   k_ptc_scatter  (odd sites)   P(v) = W0 v + sum_q U_q g_q
//   k_ptc_grad     (MFMA)        S = sum_{y,c} dg(y,c) t(y,c)^dag, the gradient of |P(v) - e|^2 w.r.t. conj(Wc)
//   k_w0_grad                    sum_x delta(x) v(x)^dag, the gradient w.r.t. conj(W0)
#include "hop_kernel.h"
#include "learned.h"

namespace cmx {

namespace {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short f2bf(float x) {  // round to nearest even
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }

// site coordinates of checkerboard index i on parity cb
__device__ __forceinline__ void ptc_coords(const Geom& g, int cb, int i, int (&x)[4]) {
  int r = i;
  const int xh = r % g.Lxh;
  r /= g.Lxh;
  x[1] = r % g.L[1];
  r /= g.L[1];
  x[2] = r % g.L[2];
  x[3] = r / g.L[2];
  x[0] = 2 * xh + ((cb + x[1] + x[2] + x[3]) & 1);
}
__device__ __forceinline__ int ptc_index(const Geom& g, const int (&x)[4]) {
  return ((x[3] * g.L[2] + x[2]) * g.L[1] + x[1]) * g.Lxh + (x[0] >> 1);
}

// o[s] = U v[s] (DAG: U^dag v[s]) for 4 spins; v, o as [spin][colour][re,im]
template <bool DAG, typename R>
__device__ __forceinline__ void transport4(const R (&u)[18], const R (&v)[24], R (&o)[24]) {
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      R re = 0, im = 0;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const R vr = v[2 * (3 * s + j)], vi = v[2 * (3 * s + j) + 1];
        if constexpr (!DAG) {
          const R ur = u[2 * (3 * i + j)], ui = u[2 * (3 * i + j) + 1];
          re += ur * vr - ui * vi;
          im += ur * vi + ui * vr;
        } else {
          const R ur = u[2 * (3 * j + i)], ui = u[2 * (3 * j + i) + 1];
          re += ur * vr + ui * vi;
          im += ur * vi - ui * vr;
        }
      }
      o[2 * (3 * s + i)] = re;
      o[2 * (3 * s + i) + 1] = im;
    }
}

// neighbour y +- mu of site x (coordinates), its checkerboard index, and whether the link crosses
// the last slice (for the 12-real reconstruction phase)
__device__ __forceinline__ int ptc_neighbour(const Geom& g, const int (&x)[4], int mu, bool fwd, bool& last) {
  int y[4] = {x[0], x[1], x[2], x[3]};
  const int L = g.L[mu];
  if (fwd) {
    last = (x[mu] == L - 1);
    y[mu] = last ? 0 : x[mu] + 1;
  } else {
    last = (x[mu] == 0);
    y[mu] = last ? L - 1 : x[mu] - 1;
  }
  return ptc_index(g, y);
}

// T tile layout (the B operand of v_mfma_f32_32x32x8bf16_1k for 32 sites x one colour):
//   T[((blk * 3 + c) * 8 + ks) * 256 + (h * 32 + n) * 4 + j]  (bf16)
// holds real-embedding row k = 8 ks + 4 h + j of column n (site blk*32+n, colour c); rows 0..31 are the
// real parts of the components p*4+s (path p, spin s), rows 32..63 their imaginary parts.
__device__ __forceinline__ size_t tile_off(int blk, int c, int ks, int h, int n) {
  return ((size_t)((blk * 3 + c) * 8 + ks) * 64 + h * 32 + n) * 4;
}

template <int P, int RECON, bool ADJ>
__global__ void __launch_bounds__(256) k_ptc_gather(PtcLaunch L) {
  const int yi = blockIdx.x * blockDim.x + threadIdx.x;
  const Geom& g = L.g;
  if (yi >= g.Vh) return;
  using R = typename SpinorIO<P>::R;
  constexpr int LCH = link_chunks(P, RECON), LCB = link_chunk_bytes(P, RECON);
  int x[4];
  ptc_coords(g, 0, yi, x);
  const int blk = yi >> 5, n = yi & 31;
  unsigned short* T = L.T;
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      const bool fwd = d == 0;
      bool last;
      const int yn = ptc_neighbour(g, x, mu, fwd, last);
      R v[24], u[18], t[24];
      SpinorIO<P>::load(L.in, L.in_nrm, L.in_stride, yn, v);
      if (fwd) {  // U_mu(y) from the even-parity links
        const char* gb = static_cast<const char*>(L.gauge[0]) + (size_t)mu * LCH * LCB * g.Vh;
        LinkIO<P, RECON>::load(gb, g.Vh, yi, u, R(L.ph[mu][last][0]), R(L.ph[mu][last][1]));
        transport4<false>(u, v, t);
      } else {    // U_mu(y-mu)^dag from the odd-parity links
        const char* gb = static_cast<const char*>(L.gauge[1]) + (size_t)mu * LCH * LCB * g.Vh;
        LinkIO<P, RECON>::load(gb, g.Vh, yn, u, R(L.ph[mu][last][0]), R(L.ph[mu][last][1]));
        transport4<true>(u, v, t);
      }
      const int p = 2 * mu + d;
      const int slot = ADJ ? (p ^ 1) : p;  // dg_q = t_{-q}[delta]
      const int ks = slot >> 1, h = slot & 1;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ushort4 re, im;
        re.x = f2bf((float)t[2 * (0 + c)]);
        re.y = f2bf((float)t[2 * (3 + c)]);
        re.z = f2bf((float)t[2 * (6 + c)]);
        re.w = f2bf((float)t[2 * (9 + c)]);
        im.x = f2bf((float)t[2 * (0 + c) + 1]);
        im.y = f2bf((float)t[2 * (3 + c) + 1]);
        im.z = f2bf((float)t[2 * (6 + c) + 1]);
        im.w = f2bf((float)t[2 * (9 + c) + 1]);
        *reinterpret_cast<ushort4*>(T + tile_off(blk, c, ks, h, n)) = re;
        *reinterpret_cast<ushort4*>(T + tile_off(blk, c, 4 + ks, h, n)) = im;
      }
    }
  }
}

// g = Wc t: one wave per (32-site block, colour) tile, grid-stride over tiles.  A operand (the real
// embedding [[Wr, -Wi], [Wi, Wr]] of Wc, 64x64) is read once per wave into registers.
// G layout: G[((q * 3 + c) * Vh + y) * 8 + 2 s + ri]  (bf16), one 16-byte group per (q, c, y).
__global__ void __launch_bounds__(256) k_ptc_mix(PtcLaunch L) {
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const int nblk = (L.g.Vh + 31) >> 5;
  const int ntiles = nblk * 3;
  bf4 a[2][8];
  const bf4* A = reinterpret_cast<const bf4*>(L.Afr);
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) a[mt][ks] = A[(mt * 8 + ks) * 64 + lane];
  for (int tile = wave; tile < ntiles; tile += nwaves) {
    const int blk = tile / 3, c = tile - 3 * blk;
    const bf4* B = reinterpret_cast<const bf4*>(L.T + tile_off(blk, c, 0, 0, 0));
    f16v acc0 = {0}, acc1 = {0};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const bf4 b = B[ks * 64 + lane];
      acc0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[0][ks], b, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[1][ks], b, acc1, 0, 0, 0);
    }
    const int y = blk * 32 + (lane & 31);
    if (y < L.g.Vh) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int q = 2 * j + (lane >> 5);
        ushort4 lo, hi;  // (s0 re, s0 im, s1 re, s1 im), (s2 re, s2 im, s3 re, s3 im)
        lo.x = f2bf(acc0[4 * j + 0]);
        lo.y = f2bf(acc1[4 * j + 0]);
        lo.z = f2bf(acc0[4 * j + 1]);
        lo.w = f2bf(acc1[4 * j + 1]);
        hi.x = f2bf(acc0[4 * j + 2]);
        hi.y = f2bf(acc1[4 * j + 2]);
        hi.z = f2bf(acc0[4 * j + 3]);
        hi.w = f2bf(acc1[4 * j + 3]);
        ushort4* dst = reinterpret_cast<ushort4*>(L.G + ((size_t)(q * 3 + c) * L.g.Vh + y) * 8);
        dst[0] = lo;
        dst[1] = hi;
      }
    }
  }
}

// Fused gather + mix for inference, no LDS.  A wave owns 32 even sites; lane l takes site l % 32 and
// the forward (l < 32) or backward (l >= 32) paths of it.  In the B-operand layout of the 32x32x8 MFMA
// lane l holds rows 8 ks + 4 (l / 32) + j of column l % 32, which for the real-embedded components
// p*4 + s are exactly path p = 2 mu + (l / 32) (mu = ks for the real rows ks < 4, mu = ks - 4 for the
// imaginary rows): each lane computes its 4 transports and packs them straight into its B fragments.
template <int P, int RECON>
__global__ void __launch_bounds__(256) k_ptc_gather_mix(PtcLaunch L) {
  const Geom& g = L.g;
  using R = typename SpinorIO<P>::R;
  constexpr int LCH = link_chunks(P, RECON), LCB = link_chunk_bytes(P, RECON);
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5;  // 0: forward paths, 1: backward paths
  const int y = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 32 + (lane & 31);
  const bool live = y < g.Vh;
  bf4 b[3][8];
  if (live) {
    int x[4];
    ptc_coords(g, 0, y, x);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      bool last;
      const int yn = ptc_neighbour(g, x, mu, h == 0, last);
      R v[24], u[18], t[24];
      SpinorIO<P>::load(L.in, L.in_nrm, L.in_stride, yn, v);
      if (h == 0) {
        const char* gb = static_cast<const char*>(L.gauge[0]) + (size_t)mu * LCH * LCB * g.Vh;
        LinkIO<P, RECON>::load(gb, g.Vh, y, u, R(L.ph[mu][last][0]), R(L.ph[mu][last][1]));
        transport4<false>(u, v, t);
      } else {
        const char* gb = static_cast<const char*>(L.gauge[1]) + (size_t)mu * LCH * LCB * g.Vh;
        LinkIO<P, RECON>::load(gb, g.Vh, yn, u, R(L.ph[mu][last][0]), R(L.ph[mu][last][1]));
        transport4<true>(u, v, t);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          b[c][mu][s] = (short)f2bf((float)t[2 * (3 * s + c)]);
          b[c][4 + mu][s] = (short)f2bf((float)t[2 * (3 * s + c) + 1]);
        }
      }
    }
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) b[c][ks] = (bf4){0, 0, 0, 0};
  }
  const bf4* A = reinterpret_cast<const bf4*>(L.Afr);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    f16v acc0 = {0}, acc1 = {0};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      acc0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A[ks * 64 + lane], b[c][ks], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A[(8 + ks) * 64 + lane], b[c][ks], acc1, 0, 0, 0);
    }
    if (live) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int q = 2 * j + h;
        ushort4 lo, hi;
        lo.x = f2bf(acc0[4 * j + 0]);
        lo.y = f2bf(acc1[4 * j + 0]);
        lo.z = f2bf(acc0[4 * j + 1]);
        lo.w = f2bf(acc1[4 * j + 1]);
        hi.x = f2bf(acc0[4 * j + 2]);
        hi.y = f2bf(acc1[4 * j + 2]);
        hi.z = f2bf(acc0[4 * j + 3]);
        hi.w = f2bf(acc1[4 * j + 3]);
        ushort4* dst = reinterpret_cast<ushort4*>(L.G + ((size_t)(q * 3 + c) * g.Vh + y) * 8);
        dst[0] = lo;
        dst[1] = hi;
      }
    }
  }
}

// Fused gather + mix for inference: a work-group of 128 threads transports its 128 even sites into an
// LDS copy of their four T tiles (48 KiB), then its two waves run the MFMA mix of those 12 (block,
// colour) tiles and write G.  T never goes to HBM.
template <int P, int RECON>
__global__ void __launch_bounds__(128) k_ptc_gather_mix_lds(PtcLaunch L) {
  __shared__ bf4 lds[4 * 3 * 8 * 64];
  const Geom& g = L.g;
  using R = typename SpinorIO<P>::R;
  constexpr int LCH = link_chunks(P, RECON), LCB = link_chunk_bytes(P, RECON);
  const int yi = blockIdx.x * 128 + threadIdx.x;
  const int lb = threadIdx.x >> 5, n = threadIdx.x & 31;  // local block, site within the block
  auto lds_at = [&](int c, int ks, int h) -> bf4& { return lds[((lb * 3 + c) * 8 + ks) * 64 + h * 32 + n]; };
  if (yi < g.Vh) {
    int x[4];
    ptc_coords(g, 0, yi, x);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const bool fwd = d == 0;
        bool last;
        const int yn = ptc_neighbour(g, x, mu, fwd, last);
        R v[24], u[18], t[24];
        SpinorIO<P>::load(L.in, L.in_nrm, L.in_stride, yn, v);
        if (fwd) {
          const char* gb = static_cast<const char*>(L.gauge[0]) + (size_t)mu * LCH * LCB * g.Vh;
          LinkIO<P, RECON>::load(gb, g.Vh, yi, u, R(L.ph[mu][last][0]), R(L.ph[mu][last][1]));
          transport4<false>(u, v, t);
        } else {
          const char* gb = static_cast<const char*>(L.gauge[1]) + (size_t)mu * LCH * LCB * g.Vh;
          LinkIO<P, RECON>::load(gb, g.Vh, yn, u, R(L.ph[mu][last][0]), R(L.ph[mu][last][1]));
          transport4<true>(u, v, t);
        }
        const int p = 2 * mu + d;
        const int ks = p >> 1, h = p & 1;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          bf4 re, im;
          re[0] = (short)f2bf((float)t[2 * (0 + c)]);
          re[1] = (short)f2bf((float)t[2 * (3 + c)]);
          re[2] = (short)f2bf((float)t[2 * (6 + c)]);
          re[3] = (short)f2bf((float)t[2 * (9 + c)]);
          im[0] = (short)f2bf((float)t[2 * (0 + c) + 1]);
          im[1] = (short)f2bf((float)t[2 * (3 + c) + 1]);
          im[2] = (short)f2bf((float)t[2 * (6 + c) + 1]);
          im[3] = (short)f2bf((float)t[2 * (9 + c) + 1]);
          lds_at(c, ks, h) = re;
          lds_at(c, 4 + ks, h) = im;
        }
      }
    }
  } else {  // sites past the end of the lattice contribute zero columns
    const bf4 z = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int ks = 0; ks < 8; ++ks)
#pragma unroll
        for (int h = 0; h < 2; ++h) lds_at(c, ks, h) = z;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  bf4 a[2][8];
  const bf4* A = reinterpret_cast<const bf4*>(L.Afr);
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) a[mt][ks] = A[(mt * 8 + ks) * 64 + lane];
  for (int tile = wave; tile < 12; tile += 2) {  // (local block, colour)
    const int tb = tile / 3, c = tile - 3 * tb;
    f16v acc0 = {0}, acc1 = {0};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const bf4 b = lds[((tb * 3 + c) * 8 + ks) * 64 + lane];
      acc0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[0][ks], b, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[1][ks], b, acc1, 0, 0, 0);
    }
    const int y = blockIdx.x * 128 + tb * 32 + (lane & 31);
    if (y < g.Vh) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int q = 2 * j + (lane >> 5);
        ushort4 lo, hi;
        lo.x = f2bf(acc0[4 * j + 0]);
        lo.y = f2bf(acc1[4 * j + 0]);
        lo.z = f2bf(acc0[4 * j + 1]);
        lo.w = f2bf(acc1[4 * j + 1]);
        hi.x = f2bf(acc0[4 * j + 2]);
        hi.y = f2bf(acc1[4 * j + 2]);
        hi.z = f2bf(acc0[4 * j + 3]);
        hi.w = f2bf(acc1[4 * j + 3]);
        ushort4* dst = reinterpret_cast<ushort4*>(L.G + ((size_t)(q * 3 + c) * g.Vh + y) * 8);
        dst[0] = lo;
        dst[1] = hi;
      }
    }
  }
}

template <int P, int RECON>
__global__ void __launch_bounds__(256) k_ptc_scatter(PtcLaunch L) {
  const int xi = blockIdx.x * blockDim.x + threadIdx.x;
  const Geom& g = L.g;
  if (xi >= g.Vh) return;
  using R = typename SpinorIO<P>::R;
  constexpr int LCH = link_chunks(P, RECON), LCB = link_chunk_bytes(P, RECON);
  int x[4];
  ptc_coords(g, 1, xi, x);
  float acc[24];
  {  // W0 v(x)
    R v[24];
    SpinorIO<P>::load(L.in, L.in_nrm, L.in_stride, xi, v);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float re = 0, im = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const float wr = L.W0[2 * (4 * s + t)], wi = L.W0[2 * (4 * s + t) + 1];
          const float vr = (float)v[2 * (3 * t + c)], vi = (float)v[2 * (3 * t + c) + 1];
          re += wr * vr - wi * vi;
          im += wr * vi + wi * vr;
        }
        acc[2 * (3 * s + c)] = re;
        acc[2 * (3 * s + c) + 1] = im;
      }
  }
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      const bool fwd = d == 0;
      bool last;
      const int yn = ptc_neighbour(g, x, mu, fwd, last);
      const int q = 2 * mu + d;
      float gv[24];  // g_q(y) as [spin][colour][re,im]
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const ushort4* src = reinterpret_cast<const ushort4*>(L.G + ((size_t)(q * 3 + c) * g.Vh + yn) * 8);
        const ushort4 lo = src[0], hi = src[1];
        gv[2 * (0 + c)] = bf2f(lo.x);
        gv[2 * (0 + c) + 1] = bf2f(lo.y);
        gv[2 * (3 + c)] = bf2f(lo.z);
        gv[2 * (3 + c) + 1] = bf2f(lo.w);
        gv[2 * (6 + c)] = bf2f(hi.x);
        gv[2 * (6 + c) + 1] = bf2f(hi.y);
        gv[2 * (9 + c)] = bf2f(hi.z);
        gv[2 * (9 + c) + 1] = bf2f(hi.w);
      }
      R u[18];
      float uf[18], o[24];
      if (fwd) {  // U_mu(x), odd-parity links
        const char* gb = static_cast<const char*>(L.gauge[1]) + (size_t)mu * LCH * LCB * g.Vh;
        LinkIO<P, RECON>::load(gb, g.Vh, xi, u, R(L.ph[mu][last][0]), R(L.ph[mu][last][1]));
#pragma unroll
        for (int k = 0; k < 18; ++k) uf[k] = (float)u[k];
        transport4<false>(uf, gv, o);
      } else {    // U_mu(x-mu)^dag, even-parity links
        const char* gb = static_cast<const char*>(L.gauge[0]) + (size_t)mu * LCH * LCB * g.Vh;
        LinkIO<P, RECON>::load(gb, g.Vh, yn, u, R(L.ph[mu][last][0]), R(L.ph[mu][last][1]));
#pragma unroll
        for (int k = 0; k < 18; ++k) uf[k] = (float)u[k];
        transport4<true>(uf, gv, o);
      }
#pragma unroll
      for (int k = 0; k < 24; ++k) acc[k] += o[k];
    }
  }
  typename SpinorIO<P>::R out[24];
#pragma unroll
  for (int k = 0; k < 24; ++k) out[k] = (typename SpinorIO<P>::R)acc[k];
  SpinorIO<P>::store(L.out, L.out_nrm, L.out_stride, xi, out);
}

// S_real (64x64) = sum over tiles of dG_tile (64 x 32 sites) T_tile^T (32 sites x 64), per wave into
// partials[wave][4][64][16] (four 32x32 output blocks (mt, nt) in the MFMA accumulator layout).
__global__ void __launch_bounds__(256) k_ptc_grad(PtcLaunch L) {
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const int nblk = (L.g.Vh + 31) >> 5;
  const int ntiles = nblk * 3;
  f16v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (f16v){0};
  const int row = lane & 31, half = lane >> 5;
  for (int tile = wave; tile < ntiles; tile += nwaves) {
    const int blk = tile / 3, c = tile - 3 * blk;
    const unsigned short* dG = L.dT + tile_off(blk, c, 0, 0, 0);
    const unsigned short* T = L.T + tile_off(blk, c, 0, 0, 0);
#pragma unroll
    for (int kq = 0; kq < 4; ++kq) {  // 8 sites per step: lanes 0-31 sites 8kq..+3, lanes 32-63 8kq+4..+7
      const int s0 = 8 * kq + 4 * half;
      bf4 a[2], b[2];
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        const int r = 32 * m + row;  // real-embedding row: (ks, h, j) = (r / 8, (r % 8) / 4, r % 4)
        const size_t base = ((size_t)(r >> 3) * 64 + ((r >> 2) & 1) * 32) * 4 + (r & 3);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          a[m][j] = (short)dG[base + (size_t)(s0 + j) * 4];
          b[m][j] = (short)T[base + (size_t)(s0 + j) * 4];
        }
      }
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
    }
  }
  float* out = L.gpart + (size_t)wave * 4 * 64 * 16;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
      for (int r = 0; r < 16; ++r) out[((mt * 2 + nt) * 64 + lane) * 16 + r] = acc[mt][nt][r];
}

// out[e] = sum over waves of partials[wave][e], e < 4*64*16 (FP64 accumulation, fixed order)
__global__ void __launch_bounds__(256) k_ptc_grad_sum(const float* part, int nwaves, double* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 4 * 64 * 16) return;
  double s = 0;
  for (int w = 0; w < nwaves; ++w) s += part[(size_t)w * 4 * 64 * 16 + e];
  out[e] = s;
}

// partials[b][2(4s+t) + ri] = sum_x sum_c delta(x)[s][c] conj(v(x)[t][c]) over work-group b
template <int P>
__global__ void __launch_bounds__(256) k_w0_grad(PtcLaunch L) {
  const int xi = blockIdx.x * blockDim.x + threadIdx.x;
  double red[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) red[k] = 0;
  if (xi < L.g.Vh) {
    typename SpinorIO<P>::R d[24], v[24];
    SpinorIO<P>::load(L.out, L.out_nrm, L.out_stride, xi, d);
    SpinorIO<P>::load(L.in, L.in_nrm, L.in_stride, xi, v);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double dr = d[2 * (3 * s + c)], di = d[2 * (3 * s + c) + 1];
          const double vr = v[2 * (3 * t + c)], vi = v[2 * (3 * t + c) + 1];
          red[2 * (4 * s + t)] += dr * vr + di * vi;
          red[2 * (4 * s + t) + 1] += di * vr - dr * vi;
        }
  }
  block_sum<32>(red);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 32; ++k) L.w0part[(size_t)blockIdx.x * 32 + k] = red[k];
}

template <typename F> void dispatch_pr(int prec, int recon, F&& f) {
  if (prec == P64) f(std::integral_constant<int, P64>(), std::integral_constant<int, 18>());
  else if (prec == P32 && recon == 18) f(std::integral_constant<int, P32>(), std::integral_constant<int, 18>());
  else if (prec == P32) f(std::integral_constant<int, P32>(), std::integral_constant<int, 12>());
  else if (prec == P16 && recon == 12) f(std::integral_constant<int, P16>(), std::integral_constant<int, 12>());
  else throw Error(CMX_ERR_UNSUPPORTED, "learned preconditioner: precision / reconstruction not supported");
}

}  // namespace

int ptc_waves() { return 1024; }

void launch_ptc_gather(const PtcLaunch& L, bool adj, hipStream_t s) {
  const dim3 grid((L.g.Vh + 255) / 256), block(256);
  dispatch_pr(L.prec, L.recon, [&](auto P, auto RC) {
    if (adj)
      hipLaunchKernelGGL((k_ptc_gather<decltype(P)::value, decltype(RC)::value, true>), grid, block, 0, s, L);
    else
      hipLaunchKernelGGL((k_ptc_gather<decltype(P)::value, decltype(RC)::value, false>), grid, block, 0, s, L);
  });
}

void launch_ptc_gather_mix(const PtcLaunch& L, int variant, hipStream_t s) {
  if (variant == 1) {  // transports staged in LDS
    const dim3 grid((L.g.Vh + 127) / 128), block(128);
    dispatch_pr(L.prec, L.recon, [&](auto P, auto RC) {
      hipLaunchKernelGGL((k_ptc_gather_mix_lds<decltype(P)::value, decltype(RC)::value>), grid, block, 0, s, L);
    });
    return;
  }
  const dim3 grid((L.g.Vh + 127) / 128), block(256);  // 4 waves x 32 sites, operands in registers
  dispatch_pr(L.prec, L.recon, [&](auto P, auto RC) {
    hipLaunchKernelGGL((k_ptc_gather_mix<decltype(P)::value, decltype(RC)::value>), grid, block, 0, s, L);
  });
}

void launch_ptc_mix(const PtcLaunch& L, hipStream_t s) {
  const int ntiles = ((L.g.Vh + 31) / 32) * 3;
  const int waves = std::min(ntiles, ptc_waves());
  hipLaunchKernelGGL(k_ptc_mix, dim3((waves + 3) / 4), dim3(256), 0, s, L);
}

void launch_ptc_scatter(const PtcLaunch& L, hipStream_t s) {
  const dim3 grid((L.g.Vh + 255) / 256), block(256);
  dispatch_pr(L.prec, L.recon, [&](auto P, auto RC) {
    hipLaunchKernelGGL((k_ptc_scatter<decltype(P)::value, decltype(RC)::value>), grid, block, 0, s, L);
  });
}

void launch_ptc_grad(const PtcLaunch& L, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_ptc_grad, dim3(ptc_waves() / 4), dim3(256), 0, s, L);
  hipLaunchKernelGGL(k_ptc_grad_sum, dim3(4 * 64 * 16 / 256), dim3(256), 0, s, L.gpart, ptc_waves(), out);
}

int launch_w0_grad(const PtcLaunch& L, hipStream_t s) {
  const int nb = (L.g.Vh + 255) / 256;
  switch (L.prec) {
    case P64: hipLaunchKernelGGL(k_w0_grad<P64>, dim3(nb), dim3(256), 0, s, L); break;
    case P32: hipLaunchKernelGGL(k_w0_grad<P32>, dim3(nb), dim3(256), 0, s, L); break;
    default: hipLaunchKernelGGL(k_w0_grad<P16>, dim3(nb), dim3(256), 0, s, L); break;
  }
  return nb;
}

}  // namespace cmx
