// Gauge-equivariant parallel-transport network used as the right preconditioner of the flexible
// solver.  On odd-checkerboard vectors v:
//
//   P(v)(x) = W0 v(x) + sum_q U_q(x) Pi_q g_q(x+q),                       x odd,  q in {+-mu}
//   g_q(y)  = sum_p Wc[q][p] U_p(y) (Pi_p v)(y+p),   Wc = W2 W1,             y even
//
// U_{+mu}(y) = U_mu(y), U_{-mu}(y) = U_mu(y-mu)^dag; Pi_p = (1 + sigma_p gamma_mu)/2 (sigma = -1 forward,
// +1 backward) is the hopping term's rank-2 spin projector, so every transport moves a half spinor, as
// in the hopping term.  W0 and the 4x4 spin blocks of Wc act identically on colour, so P commutes with
// gauge transformations.  In half-spinor coordinates (Pi = R Phat, Phat = (1/2)[1, sigma B_mu],
// R = [1; sigma B_mu^dag], B_mu the upper-right block of gamma_mu in the DeGrand-Rossi basis) the network
// is one 16x16 complex matrix Wh = Phat_q Wc R_p over (path, half-spin):
//
//   th_p(y) = U_p(y) (Phat_p v)(y+p)          gh = Wh th          P(v)(x) = W0 v(x) + sum_q U_q(x) R_q gh_q(x+q)
//
// Wh is applied on MFMA (v_mfma_f32_32x32x8bf16_1k, its real embedding is one 32x32 A operand) in BF16
// with FP32 accumulation; the transports run on the vector units.
//
// Kernels:
//   k_ptc_gather   (even sites)  th_p(y) for the 8 paths, stored in BF16 tiles laid out as the B operand
//                  of the MFMA (or, with ADJ, the loss gradient dgh_q(y) = R_q^dag U_q(y-q)^dag delta(y-q),
//                  the adjoint of the scatter, gathered along -q)
//   k_ptc_mix      (MFMA)        gh = Wh th per (site, colour) column
//   k_ptc_gather_mix             both of the above in one kernel, transports packed straight into the
//                                MFMA operands (inference path)
//   k_ptc_gather_mix_lds         the same with the transports staged in LDS tiles
//   k_ptc_scatter  (odd sites)   P(v) = W0 v + sum_q U_q R_q gh_q (with ADJ: a block's adjoint,
//                                W0^dag delta + G^dag(dth), from the mixed adjoint dth = Wh^dag dgh)
//   k_ptc_grad     (MFMA)        Sh = sum_{y,c} dgh(y,c) th(y,c)^dag, the gradient w.r.t. conj(Wh)
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

// h[s'] = scale (v_up[s'] + sigma (B_mu v_lo)[s']), s' = 0, 1; v as [spin][colour][re,im], h as
// [s'][colour][re,im].  B_0 = [[0, i], [i, 0]], B_1 = [[0, -1], [1, 0]], B_2 = [[i, 0], [0, -i]], B_3 = 1.
template <int MU, typename R>
__device__ __forceinline__ void half_project(const R (&v)[24], float sigma, float scale, float (&h)[12]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float u0r = (float)v[2 * c], u0i = (float)v[2 * c + 1];
    const float u1r = (float)v[2 * (3 + c)], u1i = (float)v[2 * (3 + c) + 1];
    const float l0r = (float)v[2 * (6 + c)], l0i = (float)v[2 * (6 + c) + 1];
    const float l1r = (float)v[2 * (9 + c)], l1i = (float)v[2 * (9 + c) + 1];
    float b0r, b0i, b1r, b1i;  // B v_lo
    if constexpr (MU == 0) {
      b0r = -l1i, b0i = l1r, b1r = -l0i, b1i = l0r;
    } else if constexpr (MU == 1) {
      b0r = -l1r, b0i = -l1i, b1r = l0r, b1i = l0i;
    } else if constexpr (MU == 2) {
      b0r = -l0i, b0i = l0r, b1r = l1i, b1i = -l1r;
    } else {
      b0r = l0r, b0i = l0i, b1r = l1r, b1i = l1i;
    }
    h[2 * c] = scale * (u0r + sigma * b0r);
    h[2 * c + 1] = scale * (u0i + sigma * b0i);
    h[2 * (3 + c)] = scale * (u1r + sigma * b1r);
    h[2 * (3 + c) + 1] = scale * (u1i + sigma * b1i);
  }
}

// acc += R o = (o, sigma B_mu^dag o) for a half spinor o [s'][colour][re,im]
template <int MU>
__device__ __forceinline__ void reconstruct_add(const float (&o)[12], float sigma, float (&acc)[24]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float o0r = o[2 * c], o0i = o[2 * c + 1], o1r = o[2 * (3 + c)], o1i = o[2 * (3 + c) + 1];
    float l0r, l0i, l1r, l1i;  // B^dag o
    if constexpr (MU == 0) {  // B_0^dag = [[0, -i], [-i, 0]]
      l0r = o1i, l0i = -o1r, l1r = o0i, l1i = -o0r;
    } else if constexpr (MU == 1) {  // [[0, 1], [-1, 0]]
      l0r = o1r, l0i = o1i, l1r = -o0r, l1i = -o0i;
    } else if constexpr (MU == 2) {  // [[-i, 0], [0, i]]
      l0r = o0i, l0i = -o0r, l1r = -o1i, l1i = o1r;
    } else {
      l0r = o0r, l0i = o0i, l1r = o1r, l1i = o1i;
    }
    acc[2 * c] += o0r;
    acc[2 * c + 1] += o0i;
    acc[2 * (3 + c)] += o1r;
    acc[2 * (3 + c) + 1] += o1i;
    acc[2 * (6 + c)] += sigma * l0r;
    acc[2 * (6 + c) + 1] += sigma * l0i;
    acc[2 * (9 + c)] += sigma * l1r;
    acc[2 * (9 + c) + 1] += sigma * l1i;
  }
}

// o[s'] = U h[s'] (DAG: U^dag h[s']) for the two half-spin components
template <bool DAG, typename R>
__device__ __forceinline__ void transport2(const R (&u)[18], const float (&h)[12], float (&o)[12]) {
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      float re = 0, im = 0;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float hr = h[2 * (3 * s + j)], hi = h[2 * (3 * s + j) + 1];
        if constexpr (!DAG) {
          const float ur = (float)u[2 * (3 * i + j)], ui = (float)u[2 * (3 * i + j) + 1];
          re += ur * hr - ui * hi;
          im += ur * hi + ui * hr;
        } else {
          const float ur = (float)u[2 * (3 * j + i)], ui = (float)u[2 * (3 * j + i) + 1];
          re += ur * hr + ui * hi;
          im += ur * hi - ui * hr;
        }
      }
      o[2 * (3 * s + i)] = re;
      o[2 * (3 * s + i) + 1] = im;
    }
}

// The transported half spinors of the two paths 2 mu (forward) and 2 mu + 1 (backward) at even site y
// (coordinates x): th[d][s'][c].  ADJ: the adjoint gather of the loss gradient, dgh_{q}(y) for q = p^1,
// i.e. the projection of the opposite path with a factor 2 (R^dag = 2 Phat), stored under slot p^1.
template <int P, int RECON, int MU, bool ADJ, int LCP = 0>
__device__ __forceinline__ void gather_mu(const PtcLaunch& L, const int (&x)[4], int y, float (&th)[2][12],
                                          const void* in, const float* in_nrm) {
  using R = typename SpinorIO<P>::R;
  constexpr int LCH = link_chunks(P, RECON), LCB = link_chunk_bytes(P, RECON);
  const Geom& g = L.g;
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const bool fwd = d == 0;
    bool last;
    const int yn = ptc_neighbour(g, x, MU, fwd, last);
    R v[24], u[18];
    SpinorIO<P>::load(in, in_nrm, L.in_stride, yn, v);
    const float sigma_p = fwd ? -1.f : 1.f;
    float h[12];
    half_project<MU>(v, ADJ ? -sigma_p : sigma_p, ADJ ? 1.f : 0.5f, h);
    const int slot = ADJ ? (d ^ 1) : d;
    if (fwd) {  // U_mu(y) from the even-parity links
      const char* gb = static_cast<const char*>(L.gauge[0]) + (size_t)MU * LCH * LCB * g.Vh;
      LinkIO<P, RECON>::template load<LCP>(gb, g.Vh, y, u, R(L.ph[MU][last][0]), R(L.ph[MU][last][1]));
      transport2<false>(u, h, th[slot]);
    } else {    // U_mu(y-mu)^dag from the odd-parity links
      const char* gb = static_cast<const char*>(L.gauge[1]) + (size_t)MU * LCH * LCB * g.Vh;
      LinkIO<P, RECON>::template load<LCP>(gb, g.Vh, yn, u, R(L.ph[MU][last][0]), R(L.ph[MU][last][1]));
      transport2<true>(u, h, th[slot]);
    }
  }
}

// B-operand components of direction mu for colour c: real parts (rows 2p + s', p = 2 mu, 2 mu + 1) and
// imaginary parts (rows 16 + 2p + s'); rows 8 ks + 4 h + j with ks = mu / 2 (+2 for imaginary),
// h = mu % 2, j = 2 (p % 2) + s'
__device__ __forceinline__ void pack_mu(const float (&th)[2][12], int c, bf4& re, bf4& im) {
#pragma unroll
  for (int d = 0; d < 2; ++d)
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      re[2 * d + s] = (short)f2bf(th[d][2 * (3 * s + c)]);
      im[2 * d + s] = (short)f2bf(th[d][2 * (3 * s + c) + 1]);
    }
}

// T tile layout (the B operand of v_mfma_f32_32x32x8bf16_1k for 32 sites x one colour):
//   T[((blk * 3 + c) * 4 + ks) * 256 + (h * 32 + n) * 4 + j]  (bf16)
// holds real-embedding row k = 8 ks + 4 h + j of column n (site blk*32+n, colour c); rows 0..15 are the
// real parts of the half-spinor components p*2+s' (path p, half-spin s'), rows 16..31 their imaginary parts.
__device__ __forceinline__ size_t tile_off(int blk, int c, int ks, int h, int n) {
  return ((size_t)((blk * 3 + c) * 4 + ks) * 64 + h * 32 + n) * 4;
}

template <int P, int RECON, bool ADJ>
__global__ void __launch_bounds__(256) k_ptc_gather(PtcLaunch L) {
  const int yi = blockIdx.x * blockDim.x + threadIdx.x;
  const Geom& g = L.g;
  if (yi >= g.Vh) return;
  int x[4];
  ptc_coords(g, 0, yi, x);
  const int blk = yi >> 5, n = yi & 31;
  unsigned short* T = L.T;
  auto one = [&](auto MUc) {
    constexpr int MU = decltype(MUc)::value;
    float th[2][12];
    gather_mu<P, RECON, MU, ADJ>(L, x, yi, th, L.in, L.in_nrm);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      bf4 re, im;
      pack_mu(th, c, re, im);
      *reinterpret_cast<bf4*>(T + tile_off(blk, c, MU >> 1, MU & 1, n)) = re;
      *reinterpret_cast<bf4*>(T + tile_off(blk, c, 2 + (MU >> 1), MU & 1, n)) = im;
    }
  };
  one(std::integral_constant<int, 0>());
  one(std::integral_constant<int, 1>());
  one(std::integral_constant<int, 2>());
  one(std::integral_constant<int, 3>());
}

// gh rows of one 32-site x colour MFMA result written to G: lane l (site l % 32, h = l / 32) holds rows
// 8 g + 4 h + i (g = 0, 1 real; 2, 3 imaginary), i.e. the half-spinor components of the paths
// q = 4 g + 2 h + m (m = 0, 1), both half-spins.
// G layout: G[((q * 3 + c) * Vh + y) * 4 + (2 s' + ri)]  (bf16), one 8-byte group per (q, c, y).
__device__ __forceinline__ void store_gh(unsigned short* G, int Vh, int y, int c, int h, const f16v& acc) {
#pragma unroll
  for (int gq = 0; gq < 2; ++gq)
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int q = 4 * gq + 2 * h + m;
      ushort4 w;
      w.x = f2bf(acc[4 * gq + 2 * m + 0]);
      w.y = f2bf(acc[8 + 4 * gq + 2 * m + 0]);
      w.z = f2bf(acc[4 * gq + 2 * m + 1]);
      w.w = f2bf(acc[8 + 4 * gq + 2 * m + 1]);
      *reinterpret_cast<ushort4*>(G + ((size_t)(q * 3 + c) * Vh + y) * 4) = w;
    }
}

// gh = Wh th: one wave per (32-site block, colour) tile, grid-stride over tiles.  The A operand (the
// real embedding [[Wr, -Wi], [Wi, Wr]] of Wh, 32x32) is read once per wave into registers.
__global__ void __launch_bounds__(256) k_ptc_mix(PtcLaunch L) {
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const int nblk = (L.g.Vh + 31) >> 5;
  const int ntiles = nblk * 3;
  bf4 a[4];
  const bf4* A = reinterpret_cast<const bf4*>(L.Afr);
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) a[ks] = A[ks * 64 + lane];
  for (int tile = wave; tile < ntiles; tile += nwaves) {
    const int blk = tile / 3, c = tile - 3 * blk;
    const bf4* B = reinterpret_cast<const bf4*>(L.T + tile_off(blk, c, 0, 0, 0));
    f16v acc = {0};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[ks], B[ks * 64 + lane], acc, 0, 0, 0);
    const int y = blk * 32 + (lane & 31);
    if (y < L.g.Vh) store_gh(L.G, L.g.Vh, y, c, lane >> 5, acc);
  }
}

// Fused gather + mix for inference, no LDS.  A wave owns 32 even sites; lane l takes site l % 32 and
// the directions mu = h, 2 + h (h = l / 32), both paths of each: in the B-operand layout these are
// exactly the rows lane l holds, so each lane packs its four transports straight into its fragments.
template <int P, int RECON, int LCP>
__device__ __forceinline__ void gather_mix_body(const PtcLaunch& L, int wg, const void* in, const float* in_nrm,
                                                unsigned short* G) {
  const Geom& g = L.g;
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5;
  const int y = (int)((wg * blockDim.x + threadIdx.x) >> 6) * 32 + (lane & 31);
  const bool live = y < g.Vh;
  bf4 b[3][4];
  if (live) {
    int x[4];
    ptc_coords(g, 0, y, x);
    float th[2][12];
    auto put = [&](int ksr) {
#pragma unroll
      for (int c = 0; c < 3; ++c) pack_mu(th, c, b[c][ksr], b[c][2 + ksr]);
    };
    if (h == 0) {
      gather_mu<P, RECON, 0, false, LCP>(L, x, y, th, in, in_nrm);
      put(0);
      gather_mu<P, RECON, 2, false, LCP>(L, x, y, th, in, in_nrm);
      put(1);
    } else {
      gather_mu<P, RECON, 1, false, LCP>(L, x, y, th, in, in_nrm);
      put(0);
      gather_mu<P, RECON, 3, false, LCP>(L, x, y, th, in, in_nrm);
      put(1);
    }
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) b[c][ks] = (bf4){0, 0, 0, 0};
  }
  const bf4* A = reinterpret_cast<const bf4*>(L.Afr);
  bf4 a[4];
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) a[ks] = A[ks * 64 + lane];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    f16v acc = {0};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[ks], b[c][ks], acc, 0, 0, 0);
    if (live) store_gh(G, g.Vh, y, c, h, acc);
  }
}

template <int P, int RECON, int LCP>
__global__ void __launch_bounds__(256) k_ptc_gather_mix(PtcLaunch L) {
  gather_mix_body<P, RECON, LCP>(L, blockIdx.x, L.in, L.in_nrm, L.G);
}

// Fused gather + mix for inference: a work-group of 128 threads transports its 128 even sites into an
// LDS copy of their four T tiles (24 KiB), then its two waves run the MFMA mix of those 12 (block,
// colour) tiles and write G.  T never goes to HBM.
template <int P, int RECON, int LCP>
__device__ __forceinline__ void gather_mix_lds_body(const PtcLaunch& L, int wg, const void* in, const float* in_nrm,
                                                    unsigned short* G, bf4* lds) {
  const Geom& g = L.g;
  const int yi = wg * 128 + threadIdx.x;
  const int lb = threadIdx.x >> 5, n = threadIdx.x & 31;  // local block, site within the block
  auto lds_at = [&](int c, int ks, int h) -> bf4& { return lds[((lb * 3 + c) * 4 + ks) * 64 + h * 32 + n]; };
  if (yi < g.Vh) {
    int x[4];
    ptc_coords(g, 0, yi, x);
    auto one = [&](auto MUc) {
      constexpr int MU = decltype(MUc)::value;
      float th[2][12];
      gather_mu<P, RECON, MU, false, LCP>(L, x, yi, th, in, in_nrm);
#pragma unroll
      for (int c = 0; c < 3; ++c) pack_mu(th, c, lds_at(c, MU >> 1, MU & 1), lds_at(c, 2 + (MU >> 1), MU & 1));
    };
    one(std::integral_constant<int, 0>());
    one(std::integral_constant<int, 1>());
    one(std::integral_constant<int, 2>());
    one(std::integral_constant<int, 3>());
  } else {  // sites past the end of the lattice contribute zero columns
    const bf4 z = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
#pragma unroll
        for (int h = 0; h < 2; ++h) lds_at(c, ks, h) = z;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  bf4 a[4];
  const bf4* A = reinterpret_cast<const bf4*>(L.Afr);
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) a[ks] = A[ks * 64 + lane];
  for (int tile = wave; tile < 12; tile += 2) {  // (local block, colour)
    const int tb = tile / 3, c = tile - 3 * tb;
    f16v acc = {0};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[ks], lds[((tb * 3 + c) * 4 + ks) * 64 + lane], acc, 0, 0, 0);
    const int y = wg * 128 + tb * 32 + (lane & 31);
    if (y < g.Vh) store_gh(G, g.Vh, y, c, lane >> 5, acc);
  }
}

template <int P, int RECON, int LCP>
__global__ void __launch_bounds__(128) k_ptc_gather_mix_lds(PtcLaunch L) {
  __shared__ bf4 lds[4 * 3 * 4 * 64];
  gather_mix_lds_body<P, RECON, LCP>(L, blockIdx.x, L.in, L.in_nrm, L.G, lds);
}

// P(v)(x) = W0 v(x) + sum_q U_q(x) R_q gh_q(x+q) at odd sites x
template <int P, int RECON, int MU, bool ADJ, int LCP = 0>
__device__ __forceinline__ void scatter_mu(const PtcLaunch& L, const int (&x)[4], int xi, float (&acc)[24],
                                           const unsigned short* G) {
  using R = typename SpinorIO<P>::R;
  constexpr int LCH = link_chunks(P, RECON), LCB = link_chunk_bytes(P, RECON);
  const Geom& g = L.g;
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const bool fwd = d == 0;
    bool last;
    const int yn = ptc_neighbour(g, x, MU, fwd, last);
    // ADJ: the adjoint of the gather, G^dag(t)(x) = sum_p Phat_p^dag U_p(x-p)^dag t_p(x-p): the path of the
    // opposite direction (t_{-q} at x + q) reconstructed with R_{-q} = 2 Phat_{-q}^dag (the factor 1/2 is in
    // the weights); the transports are the scatter's
    const int q = 2 * MU + (ADJ ? (d ^ 1) : d);
    float gh[12];  // gh_q(y) as [s'][colour][re,im]
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const ushort4 w = *reinterpret_cast<const ushort4*>(G + ((size_t)(q * 3 + c) * g.Vh + yn) * 4);
      gh[2 * c] = bf2f(w.x);
      gh[2 * c + 1] = bf2f(w.y);
      gh[2 * (3 + c)] = bf2f(w.z);
      gh[2 * (3 + c) + 1] = bf2f(w.w);
    }
    R u[18];
    float o[12];
    if (fwd) {  // U_mu(x), odd-parity links
      const char* gb = static_cast<const char*>(L.gauge[1]) + (size_t)MU * LCH * LCB * g.Vh;
      LinkIO<P, RECON>::template load<LCP>(gb, g.Vh, xi, u, R(L.ph[MU][last][0]), R(L.ph[MU][last][1]));
      transport2<false>(u, gh, o);
    } else {    // U_mu(x-mu)^dag, even-parity links
      const char* gb = static_cast<const char*>(L.gauge[0]) + (size_t)MU * LCH * LCB * g.Vh;
      LinkIO<P, RECON>::template load<LCP>(gb, g.Vh, yn, u, R(L.ph[MU][last][0]), R(L.ph[MU][last][1]));
      transport2<true>(u, gh, o);
    }
    reconstruct_add<MU>(o, (fwd ? -1.f : 1.f) * (ADJ ? -1.f : 1.f), acc);
  }
}

template <int P, int RECON, bool ADJ, int LCP>
__device__ __forceinline__ void scatter_body(const PtcLaunch& L, int wg, const void* in, const float* in_nrm,
                                             void* out, float* out_nrm, const unsigned short* G) {
  const int xi = wg * blockDim.x + threadIdx.x;
  const Geom& g = L.g;
  if (xi >= g.Vh) return;
  using R = typename SpinorIO<P>::R;
  int x[4];
  ptc_coords(g, 1, xi, x);
  float acc[24];
  {  // W0 v(x)
    R v[24];
    SpinorIO<P>::load(in, in_nrm, L.in_stride, xi, v);
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
  scatter_mu<P, RECON, 0, ADJ, LCP>(L, x, xi, acc, G);
  scatter_mu<P, RECON, 1, ADJ, LCP>(L, x, xi, acc, G);
  scatter_mu<P, RECON, 2, ADJ, LCP>(L, x, xi, acc, G);
  scatter_mu<P, RECON, 3, ADJ, LCP>(L, x, xi, acc, G);
  typename SpinorIO<P>::R o24[24];
#pragma unroll
  for (int k = 0; k < 24; ++k) o24[k] = (typename SpinorIO<P>::R)acc[k];
  SpinorIO<P>::store(out, out_nrm, L.out_stride, xi, o24);
}

template <int P, int RECON, bool ADJ, int LCP>
__global__ void __launch_bounds__(256) k_ptc_scatter(PtcLaunch L) {
  scatter_body<P, RECON, ADJ, LCP>(L, blockIdx.x, L.in, L.in_nrm, L.out, L.out_nrm, L.G);
}

// ---- several sources at once (the batched flexible solver) -----------------------------------------
// Work-group g runs on XCD g % 8 (round-robin dispatch); its k-th work-group there, k = g / 8, takes
// source k % ns of block 8 (k / ns) + g % 8.  A block's work-groups for all sources follow each other on
// one XCD, so its links (and transported neighbours' lines) are read from that XCD's L2 after the first.
__device__ __forceinline__ bool batch_map(int g, int nblk, int ns, int& blk, int& src) {
  const int k = g >> 3;
  src = k % ns;
  blk = 8 * (k / ns) + (g & 7);
  return blk < nblk;
}

template <int P, int RECON, int LCP>
__global__ void __launch_bounds__(256) k_ptc_gather_mix_b(PtcLaunch L, PtcBatch B) {
  int blk, src;
  if (!batch_map(blockIdx.x, (L.g.Vh + 127) / 128, B.ns, blk, src)) return;
  gather_mix_body<P, RECON, LCP>(L, blk, B.in[src], B.in_nrm[src], B.G[src]);
}

template <int P, int RECON, int LCP>
__global__ void __launch_bounds__(128) k_ptc_gather_mix_lds_b(PtcLaunch L, PtcBatch B) {
  __shared__ bf4 lds[4 * 3 * 4 * 64];
  int blk, src;
  if (!batch_map(blockIdx.x, (L.g.Vh + 127) / 128, B.ns, blk, src)) return;
  gather_mix_lds_body<P, RECON, LCP>(L, blk, B.in[src], B.in_nrm[src], B.G[src], lds);
}

template <int P, int RECON, int LCP>
__global__ void __launch_bounds__(256) k_ptc_scatter_b(PtcLaunch L, PtcBatch B) {
  int blk, src;
  if (!batch_map(blockIdx.x, (L.g.Vh + 255) / 256, B.ns, blk, src)) return;
  scatter_body<P, RECON, false, LCP>(L, blk, B.in[src], B.in_nrm[src], B.out[src], B.out_nrm[src], B.G[src]);
}

// Sh_real (32x32) = sum over tiles of dG_tile (32 x 32 sites) T_tile^T (32 sites x 32), per wave into
// partials[wave][64][16] (the MFMA accumulator layout).
__global__ void __launch_bounds__(256) k_ptc_grad(PtcLaunch L) {
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const int nblk = (L.g.Vh + 31) >> 5;
  const int ntiles = nblk * 3;
  f16v acc = {0};
  const int row = lane & 31, half = lane >> 5;
  // real-embedding row -> (ks, h, j) = (row / 8, (row % 8) / 4, row % 4)
  const size_t rbase = ((size_t)(row >> 3) * 64 + ((row >> 2) & 1) * 32) * 4 + (row & 3);
  for (int tile = wave; tile < ntiles; tile += nwaves) {
    const int blk = tile / 3, c = tile - 3 * blk;
    const unsigned short* dG = L.dT + tile_off(blk, c, 0, 0, 0);
    const unsigned short* T = L.T + tile_off(blk, c, 0, 0, 0);
#pragma unroll
    for (int kq = 0; kq < 4; ++kq) {  // 8 sites per step: lanes 0-31 sites 8kq..+3, lanes 32-63 8kq+4..+7
      const int s0 = 8 * kq + 4 * half;
      bf4 a, b;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a[j] = (short)dG[rbase + (size_t)(s0 + j) * 4];
        b[j] = (short)T[rbase + (size_t)(s0 + j) * 4];
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
  }
  float* out = L.gpart + (size_t)wave * 64 * 16;
#pragma unroll
  for (int r = 0; r < 16; ++r) out[lane * 16 + r] = acc[r];
}

// out[e] = sum over waves of partials[wave][e], e < 64*16 (FP64 accumulation, fixed order)
__global__ void __launch_bounds__(256) k_ptc_grad_sum(const float* part, int nwaves, double* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 64 * 16) return;
  double s = 0;
  for (int w = 0; w < nwaves; ++w) s += part[(size_t)w * 64 * 16 + e];
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
  if (prec == P64 && recon == 12) f(std::integral_constant<int, P64>(), std::integral_constant<int, 12>());
  else if (prec == P64) f(std::integral_constant<int, P64>(), std::integral_constant<int, 18>());
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

// variants 0 (operands in registers) and 1 (transports in LDS) load the links like the spinors; 3 and 4
// are the same with the links streamed (non-temporal), so that they displace less of the neighbours' lines
void launch_ptc_gather_mix(const PtcLaunch& L, int variant, hipStream_t s) {
  const bool nt = variant >= 3;
  const int v = nt ? variant - 3 : variant;
  dispatch_pr(L.prec, L.recon, [&](auto P, auto RC) {
    constexpr int PP = decltype(P)::value, RR = decltype(RC)::value;
    const dim3 grid((L.g.Vh + 127) / 128);
    if (v == 1) {
      if (nt) hipLaunchKernelGGL((k_ptc_gather_mix_lds<PP, RR, CMX_NT>), grid, dim3(128), 0, s, L);
      else hipLaunchKernelGGL((k_ptc_gather_mix_lds<PP, RR, 0>), grid, dim3(128), 0, s, L);
    } else {
      if (nt) hipLaunchKernelGGL((k_ptc_gather_mix<PP, RR, CMX_NT>), grid, dim3(256), 0, s, L);
      else hipLaunchKernelGGL((k_ptc_gather_mix<PP, RR, 0>), grid, dim3(256), 0, s, L);
    }
  });
}

static int batch_grid(int nblk, int ns) { return 8 * ns * ((nblk + 7) / 8); }

void launch_ptc_gather_mix_batch(const PtcLaunch& L, const PtcBatch& B, int variant, hipStream_t s) {
  const int nblk = (L.g.Vh + 127) / 128;
  const bool nt = variant >= 3;
  const int v = nt ? variant - 3 : variant;
  dispatch_pr(L.prec, L.recon, [&](auto P, auto RC) {
    constexpr int PP = decltype(P)::value, RR = decltype(RC)::value;
    const dim3 grid(batch_grid(nblk, B.ns));
    if (v == 1) {
      if (nt) hipLaunchKernelGGL((k_ptc_gather_mix_lds_b<PP, RR, CMX_NT>), grid, dim3(128), 0, s, L, B);
      else hipLaunchKernelGGL((k_ptc_gather_mix_lds_b<PP, RR, 0>), grid, dim3(128), 0, s, L, B);
    } else {
      if (nt) hipLaunchKernelGGL((k_ptc_gather_mix_b<PP, RR, CMX_NT>), grid, dim3(256), 0, s, L, B);
      else hipLaunchKernelGGL((k_ptc_gather_mix_b<PP, RR, 0>), grid, dim3(256), 0, s, L, B);
    }
  });
}

void launch_ptc_scatter_batch(const PtcLaunch& L, const PtcBatch& B, bool nt, hipStream_t s) {
  const int nblk = (L.g.Vh + 255) / 256;
  dispatch_pr(L.prec, L.recon, [&](auto P, auto RC) {
    constexpr int PP = decltype(P)::value, RR = decltype(RC)::value;
    const dim3 grid(batch_grid(nblk, B.ns));
    if (nt) hipLaunchKernelGGL((k_ptc_scatter_b<PP, RR, CMX_NT>), grid, dim3(256), 0, s, L, B);
    else hipLaunchKernelGGL((k_ptc_scatter_b<PP, RR, 0>), grid, dim3(256), 0, s, L, B);
  });
}

void launch_ptc_mix(const PtcLaunch& L, hipStream_t s) {
  const int ntiles = ((L.g.Vh + 31) / 32) * 3;
  const int waves = std::min(ntiles, ptc_waves());
  hipLaunchKernelGGL(k_ptc_mix, dim3((waves + 3) / 4), dim3(256), 0, s, L);
}

void launch_ptc_scatter(const PtcLaunch& L, bool adj, hipStream_t s, bool nt) {
  const dim3 grid((L.g.Vh + 255) / 256), block(256);
  dispatch_pr(L.prec, L.recon, [&](auto P, auto RC) {
    constexpr int PP = decltype(P)::value, RR = decltype(RC)::value;
    if (adj)
      hipLaunchKernelGGL((k_ptc_scatter<PP, RR, true, 0>), grid, block, 0, s, L);
    else if (nt)
      hipLaunchKernelGGL((k_ptc_scatter<PP, RR, false, CMX_NT>), grid, block, 0, s, L);
    else
      hipLaunchKernelGGL((k_ptc_scatter<PP, RR, false, 0>), grid, block, 0, s, L);
  });
}

void launch_ptc_grad(const PtcLaunch& L, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_ptc_grad, dim3(ptc_waves() / 4), dim3(256), 0, s, L);
  hipLaunchKernelGGL(k_ptc_grad_sum, dim3(64 * 16 / 256), dim3(256), 0, s, L.gpart, ptc_waves(), out);
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
