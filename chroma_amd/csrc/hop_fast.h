// Hopping term, fast path for FP16 / FP32 storage with 12-real links and real boundary phases
// (periodic / antiperiodic; SimpleFermBC, simple_fermbc.h:87-104).  Same arithmetic contract as the
// generic kernel of hop_kernel.h, arranged for fewer vector instructions per site:
//  * 2-D grid: blockIdx.y is the (z,t) plane, so z, t are wave-uniform; x, y follow from one
//    multiply by a float reciprocal (exact for plane indices < 2^16) -- no per-lane integer division;
//  * buffer loads with 32-bit lane offsets and the chunk offset in the scalar offset operand;
//  * FP16 values enter the arithmetic without separate conversions: the projection and the link
//    multiply are written as float(half) products, which the compiler emits as mixed-precision FMAs;
//  * the per-site FP16 norm of the neighbour and the +-1 boundary sign of the rebuilt third link row
//    are folded into the factor of the accumulating FMAs.
#pragma once
#include "hop_kernel.h"

namespace cmx {

typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes) {
  // raw buffer, stride 0; word 3 = 0x00020000 for the gfx9 family (loads past `bytes` return 0)
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
// CP: cache-policy operand of the buffer loads; 0 = default, CMX_NT = non-temporal
template <int CP = 0>
__device__ __forceinline__ uint4 bload128(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const u32x4_t v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, CP);
  return make_uint4(v.x, v.y, v.z, v.w);
}
template <int CP = 0>
__device__ __forceinline__ uint2 bload64(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const u32x2_t v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, CP);
  return make_uint2(v.x, v.y);
}
__device__ __forceinline__ float bload32f(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// 12 complex numbers of a fermion (index spin*3+colour) or 6 of a link (rows 0,1), undecoded
template <int P> struct FastSpinor;
template <> struct FastSpinor<P16> {
  __half2 c[12];
  float n;
  __device__ __forceinline__ float re(int k) const { return __low2float(c[k]); }
  __device__ __forceinline__ float im(int k) const { return __high2float(c[k]); }
};
template <> struct FastSpinor<P32> {
  float2 c[12];
  static constexpr float n = 1.0f;
  __device__ __forceinline__ float re(int k) const { return c[k].x; }
  __device__ __forceinline__ float im(int k) const { return c[k].y; }
};
template <> struct FastSpinor<P64> {
  double2 c[12];
  static constexpr double n = 1.0;
  __device__ __forceinline__ double re(int k) const { return c[k].x; }
  __device__ __forceinline__ double im(int k) const { return c[k].y; }
};
template <int P> struct FastLink;
template <> struct FastLink<P16> {
  __half2 c[6];
  __device__ __forceinline__ float re(int k) const { return __low2float(c[k]); }
  __device__ __forceinline__ float im(int k) const { return __high2float(c[k]); }
};
template <> struct FastLink<P32> {
  float2 c[6];
  __device__ __forceinline__ float re(int k) const { return c[k].x; }
  __device__ __forceinline__ float im(int k) const { return c[k].y; }
};
template <> struct FastLink<P64> {
  double2 c[6];
  __device__ __forceinline__ double re(int k) const { return c[k].x; }
  __device__ __forceinline__ double im(int k) const { return c[k].y; }
};
// arithmetic type of the plane kernel: FP32 for FP16/FP32 storage, FP64 for FP64 storage
template <int P> struct FastReal { typedef float T; };
template <> struct FastReal<P64> { typedef double T; };
__device__ __forceinline__ float fmar(float a, float b, float c) { return __fmaf_rn(a, b, c); }
__device__ __forceinline__ double fmar(double a, double b, double c) { return __fma_rn(a, b, c); }
__device__ __forceinline__ double2 as_double2(uint4 v) {
  double2 d;
  __builtin_memcpy(&d, &v, 16);
  return d;
}

union H8x {
  uint4 u;
  __half2 h[4];
};
union H4x {
  uint2 u;
  __half2 h[2];
};

template <int P>
__device__ __forceinline__ void fast_load_spinor(__amdgpu_buffer_rsrc_t rs, __amdgpu_buffer_rsrc_t rn, int stride, int j,
                                                 FastSpinor<P>& s) {
  if constexpr (P == P16) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      H8x v;
      v.u = bload128(rs, j * 16, k * stride * 16);
#pragma unroll
      for (int q = 0; q < 4; ++q) s.c[4 * k + q] = v.h[q];
    }
    s.n = bload32f(rn, j * 4, 0);
  } else if constexpr (P == P32) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const uint4 v = bload128(rs, j * 16, k * stride * 16);
      s.c[2 * k] = make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
      s.c[2 * k + 1] = make_float2(__uint_as_float(v.z), __uint_as_float(v.w));
    }
  } else {  // 12 chunks: the chunk offset (up to 11 x stride x 16 B) is formed in 32-bit unsigned
#pragma unroll
    for (int k = 0; k < 12; ++k) s.c[k] = as_double2(bload128(rs, j * 16, (int)((unsigned)k * (unsigned)stride * 16u)));
  }
}

template <int P, int CP = 0>
__device__ __forceinline__ void fast_load_link(__amdgpu_buffer_rsrc_t rl, int Vh, int j, FastLink<P>& u) {
  if constexpr (P == P16) {
    H8x w;
    w.u = bload128<CP>(rl, j * 16, 0);
    u.c[0] = w.h[0];
    u.c[1] = w.h[1];
    u.c[2] = w.h[2];
    u.c[3] = w.h[3];
    H4x v;
    v.u = bload64<CP>(rl, j * 8, Vh * 16);
    u.c[4] = v.h[0];
    u.c[5] = v.h[1];
  } else if constexpr (P == P32) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const uint4 v = bload128<CP>(rl, j * 16, (int)((unsigned)k * (unsigned)Vh * 16u));
      u.c[2 * k] = make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
      u.c[2 * k + 1] = make_float2(__uint_as_float(v.z), __uint_as_float(v.w));
    }
  } else {
#pragma unroll
    for (int k = 0; k < 6; ++k) u.c[k] = as_double2(bload128<CP>(rl, j * 16, (int)((unsigned)k * (unsigned)Vh * 16u)));
  }
}

// FP16 step with the projection in packed half arithmetic: h is formed as half2 (re, im) pairs,
// one rounding to FP16 like the storage itself, and both the link entries and h enter the link
// multiply as FP16 operands of mixed-precision FMAs with FP32 accumulation.
typedef _Float16 hf2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ hf2 as_hf2(__half2 v) {
  hf2 r;
  __builtin_memcpy(&r, &v, 4);
  return r;
}
template <int SS> __device__ __forceinline__ hf2 proj_add(hf2 a, hf2 b) { return SS > 0 ? a + b : a - b; }
template <int SS> __device__ __forceinline__ hf2 proj_addi(hf2 a, hf2 b) {  // a + SS i b
  const hf2 ib = {-b.y, b.x};
  return SS > 0 ? a + ib : a - ib;
}

// third link row c = conj(u0 x u1) in FP32 from the FP16 rows (the boundary sign is applied by the caller)
__device__ __forceinline__ void fast_row2_h(const FastLink<P16>& lk, float (&cr)[3], float (&ci)[3]) {
  hf2 u[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) u[k] = as_hf2(lk.c[k]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int a = (k + 1) % 3, b = (k + 2) % 3;
    // chains start from fma(., ., +0): with signed zeros honoured this is not folded into a multiply,
    // so FP16 operands stay in mixed-precision FMAs instead of being converted first
    float xr = __fmaf_rn((float)u[a].x, (float)u[3 + b].x, 0.0f);
    xr = __fmaf_rn(-(float)u[a].y, (float)u[3 + b].y, xr);
    xr = __fmaf_rn(-(float)u[b].x, (float)u[3 + a].x, xr);
    xr = __fmaf_rn((float)u[b].y, (float)u[3 + a].y, xr);
    float xi = __fmaf_rn((float)u[a].x, (float)u[3 + b].y, 0.0f);
    xi = __fmaf_rn((float)u[a].y, (float)u[3 + b].x, xi);
    xi = __fmaf_rn(-(float)u[b].x, (float)u[3 + a].y, xi);
    xi = __fmaf_rn(-(float)u[b].y, (float)u[3 + a].x, xi);
    cr[k] = xr;
    ci[k] = -xi;
  }
}

template <int MU, bool FWD, int ISIGN>
__device__ __forceinline__ void fast_step_h_body(const FastSpinor<P16>& ps, const FastLink<P16>& lk,
                                                 const float (&cr)[3], const float (&ci)[3], float sg,
                                                 float (&acc)[24]);

template <int MU, bool FWD, int ISIGN>
__device__ __forceinline__ void fast_step_h(const FastSpinor<P16>& ps, const FastLink<P16>& lk, float sg,
                                            float (&acc)[24]) {
  float cr[3], ci[3];
  fast_row2_h(lk, cr, ci);
  fast_step_h_body<MU, FWD, ISIGN>(ps, lk, cr, ci, sg, acc);
}

template <int MU, bool FWD, int ISIGN>
__device__ __forceinline__ void fast_step_h_body(const FastSpinor<P16>& ps, const FastLink<P16>& lk,
                                                 const float (&cr)[3], const float (&ci)[3], float sg,
                                                 float (&acc)[24]) {
  constexpr int S = FWD ? -ISIGN : ISIGN;  // (1 + S gamma_mu)
  hf2 h[2][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const hf2 p0 = as_hf2(ps.c[c]), p1 = as_hf2(ps.c[3 + c]), p2 = as_hf2(ps.c[6 + c]), p3 = as_hf2(ps.c[9 + c]);
    if constexpr (MU == 0) {
      h[0][c] = proj_addi<S>(p0, p3);
      h[1][c] = proj_addi<S>(p1, p2);
    } else if constexpr (MU == 1) {
      h[0][c] = proj_add<-S>(p0, p3);
      h[1][c] = proj_add<S>(p1, p2);
    } else if constexpr (MU == 2) {
      h[0][c] = proj_addi<S>(p0, p2);
      h[1][c] = proj_addi<-S>(p1, p3);
    } else {
      h[0][c] = proj_add<S>(p0, p2);
      h[1][c] = proj_add<S>(p1, p3);
    }
  }
  hf2 u[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) u[k] = as_hf2(lk.c[k]);
  const float n = ps.n;
  float orr[2][3], oi[2][3];
#pragma unroll
  for (int sp = 0; sp < 2; ++sp) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      float re, im;
      if constexpr (FWD) {
        if (i < 2) {
          re = __fmaf_rn((float)u[3 * i].x, (float)h[sp][0].x, 0.0f);
          im = __fmaf_rn((float)u[3 * i].x, (float)h[sp][0].y, 0.0f);
          re = __fmaf_rn(-(float)u[3 * i].y, (float)h[sp][0].y, re);
          im = __fmaf_rn((float)u[3 * i].y, (float)h[sp][0].x, im);
#pragma unroll
          for (int j = 1; j < 3; ++j) {
            re = __fmaf_rn((float)u[3 * i + j].x, (float)h[sp][j].x, re);
            re = __fmaf_rn(-(float)u[3 * i + j].y, (float)h[sp][j].y, re);
            im = __fmaf_rn((float)u[3 * i + j].x, (float)h[sp][j].y, im);
            im = __fmaf_rn((float)u[3 * i + j].y, (float)h[sp][j].x, im);
          }
        } else {
          re = __fmaf_rn(cr[0], (float)h[sp][0].x, 0.0f);
          im = __fmaf_rn(cr[0], (float)h[sp][0].y, 0.0f);
          re = __fmaf_rn(-ci[0], (float)h[sp][0].y, re);
          im = __fmaf_rn(ci[0], (float)h[sp][0].x, im);
#pragma unroll
          for (int j = 1; j < 3; ++j) {
            re = __fmaf_rn(cr[j], (float)h[sp][j].x, re);
            re = __fmaf_rn(-ci[j], (float)h[sp][j].y, re);
            im = __fmaf_rn(cr[j], (float)h[sp][j].y, im);
            im = __fmaf_rn(ci[j], (float)h[sp][j].x, im);
          }
          re *= sg;
          im *= sg;
        }
      } else {  // o_i = sum_j conj(U_ji) h_j with U_2i = sg c_i
        re = __fmaf_rn((float)u[i].x, (float)h[sp][0].x, 0.0f);
        im = __fmaf_rn((float)u[i].x, (float)h[sp][0].y, 0.0f);
        re = __fmaf_rn((float)u[i].y, (float)h[sp][0].y, re);
        im = __fmaf_rn(-(float)u[i].y, (float)h[sp][0].x, im);
        re = __fmaf_rn((float)u[3 + i].x, (float)h[sp][1].x, re);
        re = __fmaf_rn((float)u[3 + i].y, (float)h[sp][1].y, re);
        im = __fmaf_rn((float)u[3 + i].x, (float)h[sp][1].y, im);
        im = __fmaf_rn(-(float)u[3 + i].y, (float)h[sp][1].x, im);
        const float sc = sg * cr[i], sd = sg * ci[i];
        re = __fmaf_rn(sc, (float)h[sp][2].x, re);
        re = __fmaf_rn(sd, (float)h[sp][2].y, re);
        im = __fmaf_rn(sc, (float)h[sp][2].y, im);
        im = __fmaf_rn(-sd, (float)h[sp][2].x, im);
      }
      orr[sp][i] = re;
      oi[sp][i] = im;
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float o0r = orr[0][c], o0i = oi[0][c], o1r = orr[1][c], o1i = oi[1][c];
    acc[2 * c] = __fmaf_rn(n, o0r, acc[2 * c]);
    acc[2 * c + 1] = __fmaf_rn(n, o0i, acc[2 * c + 1]);
    acc[2 * (3 + c)] = __fmaf_rn(n, o1r, acc[2 * (3 + c)]);
    acc[2 * (3 + c) + 1] = __fmaf_rn(n, o1i, acc[2 * (3 + c) + 1]);
    const float ns = n * (float)S;
    float* a2 = &acc[2 * (6 + c)];
    float* a3 = &acc[2 * (9 + c)];
    if constexpr (MU == 0) {
      a2[0] = __fmaf_rn(ns, o1i, a2[0]);
      a2[1] = __fmaf_rn(-ns, o1r, a2[1]);
      a3[0] = __fmaf_rn(ns, o0i, a3[0]);
      a3[1] = __fmaf_rn(-ns, o0r, a3[1]);
    } else if constexpr (MU == 1) {
      a2[0] = __fmaf_rn(ns, o1r, a2[0]);
      a2[1] = __fmaf_rn(ns, o1i, a2[1]);
      a3[0] = __fmaf_rn(-ns, o0r, a3[0]);
      a3[1] = __fmaf_rn(-ns, o0i, a3[1]);
    } else if constexpr (MU == 2) {
      a2[0] = __fmaf_rn(ns, o0i, a2[0]);
      a2[1] = __fmaf_rn(-ns, o0r, a2[1]);
      a3[0] = __fmaf_rn(-ns, o1i, a3[0]);
      a3[1] = __fmaf_rn(ns, o1r, a3[1]);
    } else {
      a2[0] = __fmaf_rn(ns, o0r, a2[0]);
      a2[1] = __fmaf_rn(ns, o0i, a2[1]);
      a3[0] = __fmaf_rn(ns, o1r, a3[0]);
      a3[1] = __fmaf_rn(ns, o1i, a3[1]);
    }
  }
}

// One hopping step (direction MU, forward or backward) accumulated into acc with factor `nf`
// (neighbour norm) and boundary sign `sg` of the rebuilt third link row.
template <int P, int MU, bool FWD, int ISIGN>
__device__ __forceinline__ void fast_step(const FastSpinor<P>& ps, const FastLink<P>& lk, typename FastReal<P>::T sg,
                                          typename FastReal<P>::T (&acc)[24]) {
  using R = typename FastReal<P>::T;
  constexpr int S = FWD ? -ISIGN : ISIGN;  // (1 + S gamma_mu)
  // half spinor h[2][3] (complex), unscaled by the neighbour norm
  R hr[2][3], hi[2][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int k0 = c, k1 = 3 + c, k2 = 6 + c, k3 = 9 + c;
    if constexpr (MU == 0) {  // h0 = p0 + S i p3, h1 = p1 + S i p2
      hr[0][c] = fmar(ps.im(k3), R(-S), ps.re(k0));
      hi[0][c] = fmar(ps.re(k3), R(S), ps.im(k0));
      hr[1][c] = fmar(ps.im(k2), R(-S), ps.re(k1));
      hi[1][c] = fmar(ps.re(k2), R(S), ps.im(k1));
    } else if constexpr (MU == 1) {  // h0 = p0 - S p3, h1 = p1 + S p2
      hr[0][c] = fmar(ps.re(k3), R(-S), ps.re(k0));
      hi[0][c] = fmar(ps.im(k3), R(-S), ps.im(k0));
      hr[1][c] = fmar(ps.re(k2), R(S), ps.re(k1));
      hi[1][c] = fmar(ps.im(k2), R(S), ps.im(k1));
    } else if constexpr (MU == 2) {  // h0 = p0 + S i p2, h1 = p1 - S i p3
      hr[0][c] = fmar(ps.im(k2), R(-S), ps.re(k0));
      hi[0][c] = fmar(ps.re(k2), R(S), ps.im(k0));
      hr[1][c] = fmar(ps.im(k3), R(S), ps.re(k1));
      hi[1][c] = fmar(ps.re(k3), R(-S), ps.im(k1));
    } else {  // h0 = p0 + S p2, h1 = p1 + S p3
      hr[0][c] = fmar(ps.re(k2), R(S), ps.re(k0));
      hi[0][c] = fmar(ps.im(k2), R(S), ps.im(k0));
      hr[1][c] = fmar(ps.re(k3), R(S), ps.re(k1));
      hi[1][c] = fmar(ps.im(k3), R(S), ps.im(k1));
    }
  }
  // third row of the link: c = conj(u0 x u1) (the boundary sign is applied through sg)
  R cr[3], ci[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int a = (k + 1) % 3, b = (k + 2) % 3;
    // u0_a u1_b - u0_b u1_a
    R xr = lk.re(a) * lk.re(3 + b);
    xr = fmar(-lk.im(a), lk.im(3 + b), xr);
    xr = fmar(-lk.re(b), lk.re(3 + a), xr);
    xr = fmar(lk.im(b), lk.im(3 + a), xr);
    R xi = lk.re(a) * lk.im(3 + b);
    xi = fmar(lk.im(a), lk.re(3 + b), xi);
    xi = fmar(-lk.re(b), lk.im(3 + a), xi);
    xi = fmar(-lk.im(b), lk.re(3 + a), xi);
    cr[k] = xr;
    ci[k] = -xi;
  }
  // o = U h (forward) or U^dag h (backward); row/column 2 carries the boundary sign
  R orr[2][3], oi[2][3];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      R re, im;
      if constexpr (FWD) {
        if (i < 2) {  // rows 0,1 from storage
          re = lk.re(3 * i) * hr[s][0];
          re = fmar(-lk.im(3 * i), hi[s][0], re);
          im = lk.re(3 * i) * hi[s][0];
          im = fmar(lk.im(3 * i), hr[s][0], im);
#pragma unroll
          for (int j = 1; j < 3; ++j) {
            re = fmar(lk.re(3 * i + j), hr[s][j], re);
            re = fmar(-lk.im(3 * i + j), hi[s][j], re);
            im = fmar(lk.re(3 * i + j), hi[s][j], im);
            im = fmar(lk.im(3 * i + j), hr[s][j], im);
          }
        } else {  // rebuilt row: sg * c . h
          re = cr[0] * hr[s][0];
          re = fmar(-ci[0], hi[s][0], re);
          im = cr[0] * hi[s][0];
          im = fmar(ci[0], hr[s][0], im);
#pragma unroll
          for (int j = 1; j < 3; ++j) {
            re = fmar(cr[j], hr[s][j], re);
            re = fmar(-ci[j], hi[s][j], re);
            im = fmar(cr[j], hi[s][j], im);
            im = fmar(ci[j], hr[s][j], im);
          }
          re *= sg;
          im *= sg;
        }
      } else {  // o_i = sum_j conj(U_ji) h_j, U_2i = sg * c_i
        re = lk.re(i) * hr[s][0];
        re = fmar(lk.im(i), hi[s][0], re);
        im = lk.re(i) * hi[s][0];
        im = fmar(-lk.im(i), hr[s][0], im);
        re = fmar(lk.re(3 + i), hr[s][1], re);
        re = fmar(lk.im(3 + i), hi[s][1], re);
        im = fmar(lk.re(3 + i), hi[s][1], im);
        im = fmar(-lk.im(3 + i), hr[s][1], im);
        const R h2r = sg * hr[s][2], h2i = sg * hi[s][2];
        re = fmar(cr[i], h2r, re);
        re = fmar(ci[i], h2i, re);
        im = fmar(cr[i], h2i, im);
        im = fmar(-ci[i], h2r, im);
      }
      orr[s][i] = re;
      oi[s][i] = im;
    }
  }
  // accumulate n * o into the upper spins and n * S B^dag o into the lower spins
  const R n = ps.n;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const R o0r = orr[0][c], o0i = oi[0][c], o1r = orr[1][c], o1i = oi[1][c];
    acc[2 * c] = fmar(n, o0r, acc[2 * c]);
    acc[2 * c + 1] = fmar(n, o0i, acc[2 * c + 1]);
    acc[2 * (3 + c)] = fmar(n, o1r, acc[2 * (3 + c)]);
    acc[2 * (3 + c) + 1] = fmar(n, o1i, acc[2 * (3 + c) + 1]);
    const R ns = n * R(S);
    R* a2 = &acc[2 * (6 + c)];
    R* a3 = &acc[2 * (9 + c)];
    if constexpr (MU == 0) {  // lo2 = -S i o1, lo3 = -S i o0
      a2[0] = fmar(ns, o1i, a2[0]);
      a2[1] = fmar(-ns, o1r, a2[1]);
      a3[0] = fmar(ns, o0i, a3[0]);
      a3[1] = fmar(-ns, o0r, a3[1]);
    } else if constexpr (MU == 1) {  // lo2 = S o1, lo3 = -S o0
      a2[0] = fmar(ns, o1r, a2[0]);
      a2[1] = fmar(ns, o1i, a2[1]);
      a3[0] = fmar(-ns, o0r, a3[0]);
      a3[1] = fmar(-ns, o0i, a3[1]);
    } else if constexpr (MU == 2) {  // lo2 = -S i o0, lo3 = S i o1
      a2[0] = fmar(ns, o0i, a2[0]);
      a2[1] = fmar(-ns, o0r, a2[1]);
      a3[0] = fmar(-ns, o1i, a3[0]);
      a3[1] = fmar(ns, o1r, a3[1]);
    } else {  // lo2 = S o0, lo3 = S o1
      a2[0] = fmar(ns, o0r, a2[0]);
      a2[1] = fmar(ns, o0i, a2[1]);
      a3[0] = fmar(ns, o1r, a3[0]);
      a3[1] = fmar(ns, o1i, a3[1]);
    }
  }
}

// Materialises the accumulator after each step: without it the compiler defers the products of
// earlier steps to the end of the kernel and the register count doubles.
template <typename T> __device__ __forceinline__ void pin_acc(T (&acc)[24]) {
#pragma unroll
  for (int k = 0; k < 24; ++k) asm volatile("" : "+v"(acc[k]));
}

// work-group blk of the launch h
template <int P, int ISIGN, int EPI, bool NORM, int LCP>
__device__ __forceinline__ void hop_fast_body(const HopLaunch& h, int blk) {
  if (h.skip && *h.skip) return;
  if (h.run && !*h.run) return;
  const int Lxh = h.g.Lxh, Ly = h.g.L[1], Lz = h.g.L[2], Lt = h.g.L[3];
  const int Pl = Lxh * Ly;  // sites per (z,t) plane
  // work-group -> (xy tile, plane).  Work-group b runs on XCD b mod 8; with the XCD-aware order each
  // XCD owns a contiguous slab of Lt/8 time slices and walks it (xy tile, z within a block of
  // tile_zb planes, t, z block), so the z +- 1 and t +- 1 neighbours are re-read from its own L2.
  const int gx = (Pl + (int)blockDim.x - 1) / (int)blockDim.x;
  int bx, z, t;
  {
    const int b = blk;
    if (h.tiled) {
      const int xcd = b & 7;
      int q = b >> 3;
      bx = q % gx;
      q /= gx;
      const int zi = q % h.tile_zb;
      q /= h.tile_zb;
      const int tl = q % h.tile_t8;
      z = (q / h.tile_t8) * h.tile_zb + zi;
      t = xcd * h.tile_t8 + tl;
    } else {
      bx = b % gx;
      const int plane_ = b / gx;
      z = plane_ % Lz;
      t = plane_ / Lz;
    }
  }
  const int plane = t * Lz + z;
  const int r = bx * blockDim.x + threadIdx.x;
  double nsq = 0.0, dre = 0.0, dim = 0.0;
  if (r < Pl) {
    int y = (int)((float)r * h.inv_lxh);
    int xh = r - y * Lxh;
    if (xh < 0) { --y; xh += Lxh; }
    if (xh >= Lxh) { ++y; xh -= Lxh; }
    const int x = 2 * xh + ((h.cb + y + z + t) & 1);
    const int base = plane * Pl;
    const int i = base + r;
    const int Vh = h.g.Vh;
    // neighbour indices on the other checkerboard
    const int xp = (x + 1 == 2 * Lxh) ? 0 : x + 1, xm = (x == 0) ? 2 * Lxh - 1 : x - 1;
    const int yp = (y + 1 == Ly) ? 0 : y + 1, ym = (y == 0) ? Ly - 1 : y - 1;
    const int zp_off = ((z + 1 == Lz) ? -z : 1) * Pl, zm_off = ((z == 0) ? Lz - 1 : -1) * Pl;
    const int tp_off = ((t + 1 == Lt) ? -t : 1) * Lz * Pl, tm_off = ((t == 0) ? Lt - 1 : -1) * Lz * Pl;
    const int nb[8] = {base + y * Lxh + (xp >> 1), base + y * Lxh + (xm >> 1),
                       base + yp * Lxh + xh,       base + ym * Lxh + xh,
                       i + zp_off,                 i + zm_off,
                       i + tp_off,                 i + tm_off};
    const bool lastf[4] = {x == 2 * Lxh - 1, y == Ly - 1, z == Lz - 1, t == Lt - 1};
    const bool lastb[4] = {x == 0, y == 0, z == 0, t == 0};
    using R = typename FastReal<P>::T;
    constexpr int SB = P == P16 ? 48 : (P == P32 ? 96 : 192);  // bytes per fermion site (chunks)
    const __amdgpu_buffer_rsrc_t rin = make_rsrc(h.in, (unsigned)(SB * h.in_stride));
    const __amdgpu_buffer_rsrc_t rnin = make_rsrc(h.in_nrm, (unsigned)(4 * h.in_stride));
    constexpr int LB = P == P16 ? 24 : (P == P32 ? 48 : 96);  // bytes per 12-real link
    R acc[24];
#pragma unroll
    for (int k = 0; k < 24; ++k) acc[k] = 0.f;
    // 8 steps (mu forward / backward) with a one-step lookahead of the raw loads, so that two steps
    // of raw data are live at a time.
    FastSpinor<P> sp[2];
    FastLink<P> lk[2];
    auto load_step = [&](int step, int slot) {
      const int mu = step >> 1;
      const bool fwd = (step & 1) == 0;
      const __amdgpu_buffer_rsrc_t rl = make_rsrc(
          static_cast<const char*>(h.gauge[fwd ? h.cb : 1 - h.cb]) + (size_t)mu * LB * Vh, (unsigned)(LB * Vh));
      fast_load_spinor<P>(rin, rnin, h.in_stride, nb[step], sp[slot]);
      fast_load_link<P, LCP>(rl, Vh, fwd ? i : nb[step], lk[slot]);
    };
#define CMX_FAST_STEP(STEP)                                                                            \
    {                                                                                                  \
      if ((STEP) < 7) load_step((STEP) + 1, ((STEP) + 1) & 1);                                         \
      constexpr int MU_ = (STEP) >> 1;                                                                 \
      constexpr bool FWD_ = ((STEP) & 1) == 0;                                                         \
      const bool last_ = FWD_ ? lastf[MU_] : lastb[MU_];                                               \
      const float sg_ = last_ ? h.sgn[MU_][1] : h.sgn[MU_][0];                                         \
      if constexpr (P == P16)                                                                          \
        fast_step_h<MU_, FWD_, ISIGN>(sp[(STEP) & 1], lk[(STEP) & 1], sg_, acc);                       \
      else                                                                                             \
        fast_step<P, MU_, FWD_, ISIGN>(sp[(STEP) & 1], lk[(STEP) & 1], sg_, acc);                      \
      pin_acc(acc);                                                                                    \
    }
    load_step(0, 0);
    CMX_FAST_STEP(0)
    CMX_FAST_STEP(1)
    CMX_FAST_STEP(2)
    CMX_FAST_STEP(3)
    CMX_FAST_STEP(4)
    CMX_FAST_STEP(5)
    CMX_FAST_STEP(6)
    CMX_FAST_STEP(7)
#undef CMX_FAST_STEP
    R out[24];
    const R b = (R)h.b;
    if constexpr (EPI == EPI_NONE) {
#pragma unroll
      for (int k = 0; k < 24; ++k) out[k] = b * acc[k];
    } else if constexpr (EPI == EPI_CLOV) {
#pragma unroll
      for (int k = 0; k < 24; ++k) acc[k] *= b;
      clover_apply_site<P, LCP>(h.clov, Vh, i, acc, out);
    } else if constexpr (EPI == EPI_XPAY) {
      R xs[24];
      SpinorIO<P>::template load<LCP>(h.x, h.x_nrm, h.x_stride, i, xs);
      const R a = (R)h.a;
#pragma unroll
      for (int k = 0; k < 24; ++k) out[k] = fmar(a, xs[k], b * acc[k]);
      if constexpr (NORM)
        if (h.ndot) out_dot_x(out, xs, dre, dim);
    } else {
      R xs[24];
      SpinorIO<P>::template load<LCP>(h.x, h.x_nrm, h.x_stride, i, xs);
      clover_apply_site<P, LCP>(h.clov, Vh, i, xs, out);
#pragma unroll
      for (int k = 0; k < 24; ++k) out[k] = fmar(b, acc[k], out[k]);
      if constexpr (NORM)
        if (h.ndot) out_dot_x(out, xs, dre, dim);
    }
    SpinorIO<P>::store(h.out, h.out_nrm, h.out_stride, i, out);
    if constexpr (NORM) {
#pragma unroll
      for (int k = 0; k < 24; ++k) nsq += (double)out[k] * (double)out[k];
    }
  }
  if constexpr (NORM) write_norm_partials(h, blk, nsq, dre, dim);
}

template <int P, int ISIGN, int EPI, bool NORM, int LCP>
__global__ void __launch_bounds__(256) k_hop_fast(HopLaunch h) {
  hop_fast_body<P, ISIGN, EPI, NORM, LCP>(h, blockIdx.x);
}
template <int P, int ISIGN, int EPI, bool NORM, int LCP>
__global__ void __launch_bounds__(256) k_hop_fast_b(const HopLaunch* __restrict__ tab, const int4* __restrict__ map) {
  const int4 m = map[blockIdx.x];
  if (m.x < 0) return;  // alignment padding of the batched grid
  hop_fast_body<P, ISIGN, EPI, NORM, LCP>(tab[m.x], m.y);
}

template <int P> void launch_hop_fast_t(const HopLaunch& h, hipStream_t s, const BatchGrid* bg) {
  constexpr size_t SB = P == P16 ? 48 : (P == P32 ? 96 : 192);
  if (!bg && SB * std::max((size_t)h.in_stride, (size_t)h.g.Vh) >= ((size_t)1 << 31))
    throw Error(CMX_ERR_UNSUPPORTED, "plane-grid hopping kernel: field above 2 GiB (32-bit buffer offsets)");
  const int Pl = h.g.Lxh * h.g.L[1];
  const dim3 grid(((Pl + h.block - 1) / h.block) * h.g.L[2] * h.g.L[3]);
  const bool norm = h.partials != nullptr;
  const bool lnt = (h.fast & HOP_FAST_STREAM) != 0;
#define CMX_HOPF_GO(IS, EP, NM, CP)                                                             \
  if (bg)                                                                                       \
    hipLaunchKernelGGL((k_hop_fast_b<P, IS, EP, NM, CP>), dim3(bg->nblocks), dim3(h.block), 0, \
                       s, static_cast<const HopLaunch*>(bg->tab), bg->map);                     \
  else                                                                                          \
    hipLaunchKernelGGL((k_hop_fast<P, IS, EP, NM, CP>), grid, dim3(h.block), 0, s, h);
#define CMX_HOPF_NT(IS, EP, NM)                                                                 \
  if (lnt) {                                                                                    \
    CMX_HOPF_GO(IS, EP, NM, CMX_NT)                                                             \
  } else {                                                                                      \
    CMX_HOPF_GO(IS, EP, NM, 0)                                                                  \
  }
#define CMX_HOPF_CASE(IS, EP)                                                                   \
  if (h.isign == IS && h.epi == EP) {                                                           \
    if (norm) {                                                                                 \
      CMX_HOPF_NT(IS, EP, true)                                                                 \
    } else {                                                                                    \
      CMX_HOPF_NT(IS, EP, false)                                                                \
    }                                                                                           \
    return;                                                                                     \
  }
  CMX_HOPF_CASE(1, EPI_NONE)
  CMX_HOPF_CASE(1, EPI_CLOV)
  CMX_HOPF_CASE(1, EPI_XPAY)
  CMX_HOPF_CASE(1, EPI_CLOV_XPAY)
  CMX_HOPF_CASE(-1, EPI_NONE)
  CMX_HOPF_CASE(-1, EPI_CLOV)
  CMX_HOPF_CASE(-1, EPI_XPAY)
  CMX_HOPF_CASE(-1, EPI_CLOV_XPAY)
#undef CMX_HOPF_CASE
#undef CMX_HOPF_NT
#undef CMX_HOPF_GO
  throw Error(CMX_ERR_ARG, "launch_hop_fast: bad isign/epilogue");
}

}  // namespace cmx
