// Wilson hopping term on one checkerboard, one output site per lane, with fused epilogues.
//
//   chi(x) = sum_mu U_mu(x) (1 - isign gamma_mu) psi(x+mu) + U_mu^dag(x-mu) (1 + isign gamma_mu) psi(x-mu)
// (QDPWilsonDslashT::apply, lwldslash_w.h:252-295, DeGrand-Rossi gammas).
//
// Spin projection: every DR gamma_mu is off-diagonal in 2x2 spin blocks, gamma_mu = [[0, B],[B^dag, 0]]
// with B unitary and entries in {0, +-1, +-i}.  Hence (1 + s gamma_mu) psi = (h, s B^dag h) with the
// half spinor h = psi_up + s B psi_down.  Only h (2 spins x 3 colours) is multiplied by the link,
// and the lower two spins are rebuilt from it:
//   mu=0: B = [[0,i],[i,0]]      h0 = p0 + s i p3, h1 = p1 + s i p2;  lo2 = -s i h1, lo3 = -s i h0
//   mu=1: B = [[0,-1],[1,0]]     h0 = p0 - s p3,   h1 = p1 + s p2;    lo2 =  s h1,   lo3 = -s h0
//   mu=2: B = [[i,0],[0,-i]]     h0 = p0 + s i p2, h1 = p1 - s i p3;  lo2 = -s i h0, lo3 =  s i h1
//   mu=3: B = [[1,0],[0,1]]      h0 = p0 + s p2,   h1 = p1 + s p3;    lo2 =  s h0,   lo3 =  s h1
// Forward terms use s = -isign, backward terms s = +isign.
#pragma once
#include "cmx_internal.h"

namespace cmx {

// h[2 spin][3 colour][re,im] = 12 reals; psi index 2*(spin*3+c)+ri
template <int MU, int S, typename R>
__device__ __forceinline__ void project(const R (&p)[24], R (&h)[12]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const R p0r = p[2 * c], p0i = p[2 * c + 1];
    const R p1r = p[2 * (3 + c)], p1i = p[2 * (3 + c) + 1];
    const R p2r = p[2 * (6 + c)], p2i = p[2 * (6 + c) + 1];
    const R p3r = p[2 * (9 + c)], p3i = p[2 * (9 + c) + 1];
    R h0r, h0i, h1r, h1i;
    if constexpr (MU == 0) {  // i*(a+ib) = -b + ia
      h0r = p0r - S * p3i; h0i = p0i + S * p3r;
      h1r = p1r - S * p2i; h1i = p1i + S * p2r;
    } else if constexpr (MU == 1) {
      h0r = p0r - S * p3r; h0i = p0i - S * p3i;
      h1r = p1r + S * p2r; h1i = p1i + S * p2i;
    } else if constexpr (MU == 2) {
      h0r = p0r - S * p2i; h0i = p0i + S * p2r;
      h1r = p1r + S * p3i; h1i = p1i - S * p3r;
    } else {
      h0r = p0r + S * p2r; h0i = p0i + S * p2i;
      h1r = p1r + S * p3r; h1i = p1i + S * p3i;
    }
    h[2 * c] = h0r; h[2 * c + 1] = h0i;
    h[6 + 2 * c] = h1r; h[6 + 2 * c + 1] = h1i;
  }
}

template <int MU, int S, typename R>
__device__ __forceinline__ void accumulate(R (&acc)[24], const R (&h)[12]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const R h0r = h[2 * c], h0i = h[2 * c + 1], h1r = h[6 + 2 * c], h1i = h[6 + 2 * c + 1];
    acc[2 * c] += h0r; acc[2 * c + 1] += h0i;
    acc[2 * (3 + c)] += h1r; acc[2 * (3 + c) + 1] += h1i;
    R l2r, l2i, l3r, l3i;
    if constexpr (MU == 0) {  // -s i h: -s*(-b + ia) = s b - s i a
      l2r = S * h1i; l2i = -S * h1r;
      l3r = S * h0i; l3i = -S * h0r;
    } else if constexpr (MU == 1) {
      l2r = S * h1r; l2i = S * h1i;
      l3r = -S * h0r; l3i = -S * h0i;
    } else if constexpr (MU == 2) {
      l2r = S * h0i; l2i = -S * h0r;   // -s i h0
      l3r = -S * h1i; l3i = S * h1r;   //  s i h1
    } else {
      l2r = S * h0r; l2i = S * h0i;
      l3r = S * h1r; l3i = S * h1i;
    }
    acc[2 * (6 + c)] += l2r; acc[2 * (6 + c) + 1] += l2i;
    acc[2 * (9 + c)] += l3r; acc[2 * (9 + c) + 1] += l3i;
  }
}

// out = U h (DAG=false) or U^dag h (DAG=true), for both spins of the half spinor
template <bool DAG, typename R>
__device__ __forceinline__ void link_mul(const R (&u)[18], const R (&h)[12], R (&o)[12]) {
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      R re = 0, im = 0;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const R hr = h[6 * s + 2 * j], hi = h[6 * s + 2 * j + 1];
        if constexpr (!DAG) {
          const R ur = u[2 * (3 * i + j)], ui = u[2 * (3 * i + j) + 1];
          re += ur * hr - ui * hi;
          im += ur * hi + ui * hr;
        } else {
          const R ur = u[2 * (3 * j + i)], ui = u[2 * (3 * j + i) + 1];
          re += ur * hr + ui * hi;
          im += ur * hi - ui * hr;
        }
      }
      o[6 * s + 2 * i] = re;
      o[6 * s + 2 * i + 1] = im;
    }
}

template <int P, int RECON> struct HopTypes {
  using R = typename SpinorIO<P>::R;
  static constexpr int LCH = link_chunks(P, RECON);
  static constexpr int LCB = link_chunk_bytes(P, RECON);
};

template <int P, int RECON, int MU, int ISIGN, int CP = 0>
__device__ __forceinline__ void hop_dir(const HopLaunch& h, typename SpinorIO<P>::R (&acc)[24], int i, const int (&x)[4],
                                        int rowbase_unused) {
  using R = typename SpinorIO<P>::R;
  constexpr int LCH = HopTypes<P, RECON>::LCH, LCB = HopTypes<P, RECON>::LCB;
  const int Vh = h.g.Vh, Lxh = h.g.Lxh;
  const int L = h.g.L[MU];
  // neighbour indices on the other checkerboard
  int xf[4] = {x[0], x[1], x[2], x[3]}, xb[4] = {x[0], x[1], x[2], x[3]};
  xf[MU] = (x[MU] + 1 == L) ? 0 : x[MU] + 1;
  xb[MU] = (x[MU] == 0) ? L - 1 : x[MU] - 1;
  const int jf = ((xf[3] * h.g.L[2] + xf[2]) * h.g.L[1] + xf[1]) * Lxh + (xf[0] >> 1);
  const int jb = ((xb[3] * h.g.L[2] + xb[2]) * h.g.L[1] + xb[1]) * Lxh + (xb[0] >> 1);

  const char* gout = reinterpret_cast<const char*>(h.gauge[h.cb]) + (size_t)MU * LCH * LCB * Vh;
  const char* gin = reinterpret_cast<const char*>(h.gauge[1 - h.cb]) + (size_t)MU * LCH * LCB * Vh;

  R p[24], hs[12], o[12], u[18];
  // forward: U_mu(x) (1 - isign gamma_mu) psi(x+mu)
  {
    SpinorIO<P>::load(h.in, h.in_nrm, h.in_stride, jf, p);
    const int last = (x[MU] == L - 1);
    LinkIO<P, RECON>::template load<CP>(gout, Vh, i, u, R(h.ph[MU][last][0]), R(h.ph[MU][last][1]));
    project<MU, -ISIGN>(p, hs);
    link_mul<false>(u, hs, o);
    accumulate<MU, -ISIGN>(acc, o);
  }
  // backward: U_mu(x-mu)^dag (1 + isign gamma_mu) psi(x-mu)
  {
    SpinorIO<P>::load(h.in, h.in_nrm, h.in_stride, jb, p);
    const int last = (x[MU] == 0);
    LinkIO<P, RECON>::template load<CP>(gin, Vh, jb, u, R(h.ph[MU][last][0]), R(h.ph[MU][last][1]));
    project<MU, ISIGN>(p, hs);
    link_mul<true>(u, hs, o);
    accumulate<MU, ISIGN>(acc, o);
  }
}

template <int P> struct CloverBlockIO;
template <> struct CloverBlockIO<P64> {
  template <int CP = 0>
  __device__ __forceinline__ static void load(const void* base, long Vh, long i, int b, double (&c)[36]) {
    const double2* q = reinterpret_cast<const double2*>(base) + (size_t)18 * b * Vh;
#pragma unroll
    for (int k = 0; k < 18; ++k) {
      double2 v = q[k * Vh + i];
      c[2 * k] = v.x;
      c[2 * k + 1] = v.y;
    }
  }
};
template <> struct CloverBlockIO<P32> {
  template <int CP = 0>
  __device__ __forceinline__ static void load(const void* base, long Vh, long i, int b, float (&c)[36]) {
    const float4* q = reinterpret_cast<const float4*>(base) + (size_t)9 * b * Vh;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      float4 v = ld_cp<CP>(&q[k * Vh + i]);
      c[4 * k] = v.x;
      c[4 * k + 1] = v.y;
      c[4 * k + 2] = v.z;
      c[4 * k + 3] = v.w;
    }
  }
};
template <> struct CloverBlockIO<P16> {
  // per chiral block b: four 16 B chunks (reals 0-31, chunk 4b+k of every site), then one 8 B chunk
  // (reals 32-35) after the eight 16 B chunks; then one float norm per site
  template <int CP = 0>
  __device__ __forceinline__ static void load(const void* base, long Vh, long i, int b, float (&c)[36]) {
    const uint2* q = reinterpret_cast<const uint2*>(base);
    const uint4* q16 = reinterpret_cast<const uint4*>(base);
    const float n = ld_cp<CP>(&reinterpret_cast<const float*>(q + 18 * Vh)[i]);
    H4 v[9];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint4 w = ld_cp<CP>(&q16[(size_t)(4 * b + k) * Vh + i]);
      v[2 * k].u = make_uint2(w.x, w.y);
      v[2 * k + 1].u = make_uint2(w.z, w.w);
    }
    v[8].u = ld_cp<CP>(&q[(size_t)(16 + b) * Vh + i]);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      float2 f0 = __half22float2(v[k].h[0]), f1 = __half22float2(v[k].h[1]);
      c[4 * k] = f0.x * n;
      c[4 * k + 1] = f0.y * n;
      c[4 * k + 2] = f1.x * n;
      c[4 * k + 3] = f1.y * n;
    }
  }
};

// out(6 components of block b) = (L + D + L^dag) in(block b); c = {diag[6], offd[15] complex}
template <typename R>
__device__ __forceinline__ void clover_block_mv(const R (&c)[36], const R (&in)[24], R (&out)[24], int b) {
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    R re = c[i] * in[2 * (6 * b + i)];
    R im = c[i] * in[2 * (6 * b + i) + 1];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      if (j == i) continue;
      const R pr = in[2 * (6 * b + j)], pi = in[2 * (6 * b + j) + 1];
      if (j < i) {
        const int e = 6 + 2 * (i * (i - 1) / 2 + j);
        re += c[e] * pr - c[e + 1] * pi;
        im += c[e] * pi + c[e + 1] * pr;
      } else {
        const int e = 6 + 2 * (j * (j - 1) / 2 + i);
        re += c[e] * pr + c[e + 1] * pi;
        im += c[e] * pi - c[e + 1] * pr;
      }
    }
    out[2 * (6 * b + i)] = re;
    out[2 * (6 * b + i) + 1] = im;
  }
}

template <int P, int CP = 0>
__device__ __forceinline__ void clover_apply_site(const void* clov, long Vh, long i, const typename SpinorIO<P>::R (&in)[24],
                                                  typename SpinorIO<P>::R (&out)[24]) {
  using R = typename SpinorIO<P>::R;
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    R c[36];
    CloverBlockIO<P>::template load<CP>(clov, Vh, i, b, c);
    clover_block_mv<R>(c, in, out, b);
  }
}

// ------------------------------------------------------------------------------------------------
// Raw (undecoded) loads, so that the next direction's memory requests can be issued before the
// current direction is decoded and multiplied (double-buffered across the 8 hopping directions).
template <int P> struct RawSpinor;
template <> struct RawSpinor<P32> {
  float4 c[6];
  __device__ __forceinline__ void load(const void* base, const float*, int stride, int j) {
    const float4* b = reinterpret_cast<const float4*>(base);
#pragma unroll
    for (int k = 0; k < 6; ++k) c[k] = b[(size_t)k * stride + j];
  }
  __device__ __forceinline__ void decode(float (&s)[24]) const {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      s[4 * k] = c[k].x;
      s[4 * k + 1] = c[k].y;
      s[4 * k + 2] = c[k].z;
      s[4 * k + 3] = c[k].w;
    }
  }
};
template <> struct RawSpinor<P16> {
  uint4 c[3];
  float n;
  __device__ __forceinline__ void load(const void* base, const float* nrm, int stride, int j) {
    const uint4* b = reinterpret_cast<const uint4*>(base);
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] = b[(size_t)k * stride + j];
    n = nrm[j];
  }
  __device__ __forceinline__ void decode(float (&s)[24]) const {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      H8 v;
      v.u = c[k];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float2 f = __half22float2(v.h[q]);
        s[8 * k + 2 * q] = f.x * n;
        s[8 * k + 2 * q + 1] = f.y * n;
      }
    }
  }
};
template <int P, int RECON> struct RawLink;
template <> struct RawLink<P32, 12> {
  float4 c[3];
  __device__ __forceinline__ void load(const void* base, int Vh, int j) {
    const float4* b = reinterpret_cast<const float4*>(base);
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] = b[(size_t)k * Vh + j];
  }
  __device__ __forceinline__ void decode(float (&u)[18], float dre, float dim) const {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      u[4 * k] = c[k].x;
      u[4 * k + 1] = c[k].y;
      u[4 * k + 2] = c[k].z;
      u[4 * k + 3] = c[k].w;
    }
    recon_row2<float>(u, dre, dim);
  }
};
template <> struct RawLink<P32, 18> {
  float2 c[9];
  __device__ __forceinline__ void load(const void* base, int Vh, int j) {
    const float2* b = reinterpret_cast<const float2*>(base);
#pragma unroll
    for (int k = 0; k < 9; ++k) c[k] = b[(size_t)k * Vh + j];
  }
  __device__ __forceinline__ void decode(float (&u)[18], float, float) const {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      u[2 * k] = c[k].x;
      u[2 * k + 1] = c[k].y;
    }
  }
};
template <> struct RawLink<P16, 12> {
  uint2 c[3];
  __device__ __forceinline__ void load(const void* base, int Vh, int j) {
    // the 16 B chunk (reals 0-7) of every site, then the 8 B chunk (reals 8-11)
    const uint4 w = reinterpret_cast<const uint4*>(base)[j];
    c[0] = make_uint2(w.x, w.y);
    c[1] = make_uint2(w.z, w.w);
    c[2] = reinterpret_cast<const uint2*>(base)[(size_t)2 * Vh + j];
  }
  __device__ __forceinline__ void decode(float (&u)[18], float dre, float dim) const {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      H4 v;
      v.u = c[k];
      const float2 f0 = __half22float2(v.h[0]), f1 = __half22float2(v.h[1]);
      u[4 * k] = f0.x;
      u[4 * k + 1] = f0.y;
      u[4 * k + 2] = f1.x;
      u[4 * k + 3] = f1.y;
    }
    recon_row2<float>(u, dre, dim);
  }
};

// Per-site context shared by the 8 directions
struct HopSite {
  int i;          // output site (checkerboard index)
  int nb[8];      // neighbour index for step 2*mu (forward) and 2*mu+1 (backward)
  int last[8];    // the link of this step sits on the last slice of its direction
};

template <int P, int RECON> struct RawStep {
  RawSpinor<P> sp;
  RawLink<P, RECON> ln;
};

template <int P, int RECON, int STEP>
__device__ __forceinline__ void raw_step_load(const HopLaunch& h, const HopSite& st, RawStep<P, RECON>& r) {
  constexpr int MU = STEP / 2;
  constexpr bool FWD = (STEP % 2) == 0;
  constexpr int LCH = link_chunks(P, RECON), LCB = link_chunk_bytes(P, RECON);
  const int j = st.nb[STEP];
  r.sp.load(h.in, h.in_nrm, h.in_stride, j);
  const char* gb = reinterpret_cast<const char*>(h.gauge[FWD ? h.cb : 1 - h.cb]) + (size_t)MU * LCH * LCB * h.g.Vh;
  r.ln.load(gb, h.g.Vh, FWD ? st.i : j);
}

template <int P, int RECON, int STEP, int ISIGN>
__device__ __forceinline__ void raw_step_apply(const HopLaunch& h, const HopSite& st, const RawStep<P, RECON>& r,
                                               float (&acc)[24]) {
  constexpr int MU = STEP / 2;
  constexpr bool FWD = (STEP % 2) == 0;
  float p[24], u[18], hs[12], o[12];
  r.sp.decode(p);
  const int last = st.last[STEP];
  r.ln.decode(u, (float)h.ph[MU][last][0], (float)h.ph[MU][last][1]);
  if constexpr (FWD) {  // U_mu(x) (1 - isign gamma_mu) psi(x+mu)
    project<MU, -ISIGN>(p, hs);
    link_mul<false>(u, hs, o);
    accumulate<MU, -ISIGN>(acc, o);
  } else {  // U_mu(x-mu)^dag (1 + isign gamma_mu) psi(x-mu)
    project<MU, ISIGN>(p, hs);
    link_mul<true>(u, hs, o);
    accumulate<MU, ISIGN>(acc, o);
  }
}

// Double-buffered pass over the 8 steps: the loads of step s+1 are in flight while step s computes
template <int P, int RECON, int ISIGN>
__device__ __forceinline__ void hop_prefetched(const HopLaunch& h, const HopSite& st, float (&acc)[24]) {
  RawStep<P, RECON> a, b;
  raw_step_load<P, RECON, 0>(h, st, a);
  raw_step_load<P, RECON, 1>(h, st, b);
  raw_step_apply<P, RECON, 0, ISIGN>(h, st, a, acc);
  raw_step_load<P, RECON, 2>(h, st, a);
  raw_step_apply<P, RECON, 1, ISIGN>(h, st, b, acc);
  raw_step_load<P, RECON, 3>(h, st, b);
  raw_step_apply<P, RECON, 2, ISIGN>(h, st, a, acc);
  raw_step_load<P, RECON, 4>(h, st, a);
  raw_step_apply<P, RECON, 3, ISIGN>(h, st, b, acc);
  raw_step_load<P, RECON, 5>(h, st, b);
  raw_step_apply<P, RECON, 4, ISIGN>(h, st, a, acc);
  raw_step_load<P, RECON, 6>(h, st, a);
  raw_step_apply<P, RECON, 5, ISIGN>(h, st, b, acc);
  raw_step_load<P, RECON, 7>(h, st, b);
  raw_step_apply<P, RECON, 6, ISIGN>(h, st, a, acc);
  raw_step_apply<P, RECON, 7, ISIGN>(h, st, b, acc);
}

// Output site of this lane.  Tiled order: work-group b runs on XCD b mod 8; XCD k owns the t-slab
// [k*Lt/8, (k+1)*Lt/8) and walks it as (xy tile fastest, z within a block of tile_zb planes, t,
// z block), so the t +- 1 and z +- 1 neighbours of a site are read again within a few work-groups
// on the same XCD and hit its L2.  Linear order: contiguous sites with the XCD remap of xcd_tile.
__device__ __forceinline__ int hop_site_index(const HopLaunch& h, int blk, int nblk) {
  if (h.tiled) {
    const int b = blk, xcd = b & 7;
    int q = b >> 3;
    const int wp = q % h.tile_w;
    q /= h.tile_w;
    const int zi = q % h.tile_zb;
    q /= h.tile_zb;
    const int tl = q % h.tile_t8;
    const int zblk = q / h.tile_t8;
    const int z = zblk * h.tile_zb + zi, t = xcd * h.tile_t8 + tl;
    return (t * h.g.L[2] + z) * (h.g.Lxh * h.g.L[1]) + wp * (int)blockDim.x + (int)threadIdx.x;
  }
  return xcd_tile(blk, nblk) * (int)blockDim.x + (int)threadIdx.x;
}

// work-group blk of nblk of the launch h
// LCP: cache policy of the loads of data read once per launch (links, clover, xpay input); the
// double-buffered (PF) path loads its links with the default policy
// <out, x> of one site (conj(out) x, FP64 accumulation) into (dre, dim)
template <typename R>
__device__ __forceinline__ void out_dot_x(const R (&out)[24], const R (&xs)[24], double& dre, double& dim) {
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const double orr = out[2 * k], oi = out[2 * k + 1], xr = xs[2 * k], xi = xs[2 * k + 1];
    dre += orr * xr + oi * xi;
    dim += orr * xi - oi * xr;
  }
}
// the work-group's reductions: |out|^2 (h.partials[blk]), or with h.ndot |out|^2, Re <out, x>, Im <out, x>
// (h.partials[3 blk + 0..2], the layout the finalize kernel reads for three reductions)
__device__ __forceinline__ void write_norm_partials(const HopLaunch& h, int blk, double nsq, double dre, double dim) {
  if (h.ndot) {
    double v[3] = {nsq, dre, dim};
    block_sum<3>(v);
    if (threadIdx.x == 0) {
      h.partials[3 * blk] = v[0];
      h.partials[3 * blk + 1] = v[1];
      h.partials[3 * blk + 2] = v[2];
    }
  } else {
    double v[1] = {nsq};
    block_sum<1>(v);
    if (threadIdx.x == 0) h.partials[blk] = v[0];
  }
}

template <int P, int RECON, int ISIGN, int EPI, bool NORM, bool PF, int LCP>
__device__ __forceinline__ void hop_body(const HopLaunch& h, int blk, int nblk) {
  using R = typename SpinorIO<P>::R;
  if (h.skip && *h.skip) return;
  if (h.run && !*h.run) return;
  const int i = hop_site_index(h, blk, nblk);
  double nsq = 0.0, dre = 0.0, dim = 0.0;
  if (i < h.g.Vh) {
    int x[4];
    {
      int r = i;
      const int xh = r % h.g.Lxh;
      r /= h.g.Lxh;
      x[1] = r % h.g.L[1];
      r /= h.g.L[1];
      x[2] = r % h.g.L[2];
      x[3] = r / h.g.L[2];
      x[0] = 2 * xh + ((h.cb + x[1] + x[2] + x[3]) & 1);
    }
    R acc[24];
#pragma unroll
    for (int k = 0; k < 24; ++k) acc[k] = 0;
    if constexpr (PF) {
      HopSite st;
      st.i = i;
#pragma unroll
      for (int mu = 0; mu < 4; ++mu) {
        const int L = h.g.L[mu];
        int xf[4] = {x[0], x[1], x[2], x[3]}, xb[4] = {x[0], x[1], x[2], x[3]};
        xf[mu] = (x[mu] + 1 == L) ? 0 : x[mu] + 1;
        xb[mu] = (x[mu] == 0) ? L - 1 : x[mu] - 1;
        st.nb[2 * mu] = ((xf[3] * h.g.L[2] + xf[2]) * h.g.L[1] + xf[1]) * h.g.Lxh + (xf[0] >> 1);
        st.nb[2 * mu + 1] = ((xb[3] * h.g.L[2] + xb[2]) * h.g.L[1] + xb[1]) * h.g.Lxh + (xb[0] >> 1);
        st.last[2 * mu] = (x[mu] == L - 1);
        st.last[2 * mu + 1] = (x[mu] == 0);
      }
      hop_prefetched<P, RECON, ISIGN>(h, st, acc);
    } else {
      hop_dir<P, RECON, 0, ISIGN, LCP>(h, acc, i, x, 0);
      hop_dir<P, RECON, 1, ISIGN, LCP>(h, acc, i, x, 0);
      hop_dir<P, RECON, 2, ISIGN, LCP>(h, acc, i, x, 0);
      hop_dir<P, RECON, 3, ISIGN, LCP>(h, acc, i, x, 0);
    }

    R out[24];
    const R b = R(h.b);
    if constexpr (EPI == EPI_NONE) {
#pragma unroll
      for (int k = 0; k < 24; ++k) out[k] = b * acc[k];
    } else if constexpr (EPI == EPI_CLOV) {
#pragma unroll
      for (int k = 0; k < 24; ++k) acc[k] *= b;
      clover_apply_site<P, LCP>(h.clov, h.g.Vh, i, acc, out);
    } else if constexpr (EPI == EPI_XPAY) {
      R xs[24];
      SpinorIO<P>::template load<LCP>(h.x, h.x_nrm, h.x_stride, i, xs);
      const R a = R(h.a);
#pragma unroll
      for (int k = 0; k < 24; ++k) out[k] = a * xs[k] + b * acc[k];
      if constexpr (NORM)
        if (h.ndot) out_dot_x(out, xs, dre, dim);
    } else {
      R xs[24];
      SpinorIO<P>::template load<LCP>(h.x, h.x_nrm, h.x_stride, i, xs);
      clover_apply_site<P, LCP>(h.clov, h.g.Vh, i, xs, out);
#pragma unroll
      for (int k = 0; k < 24; ++k) out[k] += b * acc[k];
      if constexpr (NORM)
        if (h.ndot) out_dot_x(out, xs, dre, dim);
    }
    SpinorIO<P>::store(h.out, h.out_nrm, h.out_stride, i, out);
    if constexpr (NORM) {
      // |out|^2 of the values computed (below the FP16 storage rounding; FP64 accumulation)
#pragma unroll
      for (int k = 0; k < 24; ++k) nsq += (double)out[k] * (double)out[k];
    }
  }
  if constexpr (NORM) write_norm_partials(h, blk, nsq, dre, dim);
}

template <int P, int RECON, int ISIGN, int EPI, bool NORM, bool PF, int LCP>
__global__ void __launch_bounds__(256) k_hop(HopLaunch h) {
  hop_body<P, RECON, ISIGN, EPI, NORM, PF, LCP>(h, blockIdx.x, gridDim.x);
}
template <int P, int RECON, int ISIGN, int EPI, bool NORM, bool PF, int LCP>
__global__ void __launch_bounds__(256) k_hop_b(const HopLaunch* __restrict__ tab, const int4* __restrict__ map) {
  const int4 m = map[blockIdx.x];
  if (m.x < 0) return;  // alignment padding of the batched grid
  hop_body<P, RECON, ISIGN, EPI, NORM, PF, LCP>(tab[m.x], m.y, m.z);
}

template <int P, int RECON, bool PF> void launch_hop_t(const HopLaunch& h, hipStream_t s, const BatchGrid* bg) {
  const int nb = (h.g.Vh + h.block - 1) / h.block;
  const bool norm = h.partials != nullptr;
  const bool lnt = (h.fast & HOP_FAST_STREAM) != 0 && !PF;
#define CMX_HOP_GO2(IS, EP, NM, CP)                                                                       \
  if (bg)                                                                                                 \
    hipLaunchKernelGGL((k_hop_b<P, RECON, IS, EP, NM, PF, CP>), dim3(bg->nblocks), dim3(h.block), 0, s,  \
                       static_cast<const HopLaunch*>(bg->tab), bg->map);                                  \
  else                                                                                                    \
    hipLaunchKernelGGL((k_hop<P, RECON, IS, EP, NM, PF, CP>), dim3(nb), dim3(h.block), 0, s, h);
#define CMX_HOP_GO(IS, EP, NM)                                                                            \
  if (lnt) {                                                                                              \
    CMX_HOP_GO2(IS, EP, NM, CMX_NT)                                                                       \
  } else {                                                                                                \
    CMX_HOP_GO2(IS, EP, NM, 0)                                                                            \
  }
#define CMX_HOP_CASE(IS, EP)                                                                              \
  if (h.isign == IS && h.epi == EP) {                                                                     \
    if (norm) {                                                                                           \
      CMX_HOP_GO(IS, EP, true)                                                                            \
    } else {                                                                                              \
      CMX_HOP_GO(IS, EP, false)                                                                           \
    }                                                                                                     \
    return;                                                                                               \
  }
  CMX_HOP_CASE(1, EPI_NONE)
  CMX_HOP_CASE(1, EPI_CLOV)
  CMX_HOP_CASE(1, EPI_XPAY)
  CMX_HOP_CASE(1, EPI_CLOV_XPAY)
  CMX_HOP_CASE(-1, EPI_NONE)
  CMX_HOP_CASE(-1, EPI_CLOV)
  CMX_HOP_CASE(-1, EPI_XPAY)
  CMX_HOP_CASE(-1, EPI_CLOV_XPAY)
#undef CMX_HOP_CASE
#undef CMX_HOP_GO
#undef CMX_HOP_GO2
  throw Error(CMX_ERR_ARG, "launch_hop: bad isign/epilogue");
}

}  // namespace cmx
