// Device helpers shared by the FP64 set-up and force kernels: checkerboard coordinates and indices,
// 3x3 complex matrices in registers, and FP64 links read from the per-parity layout
// g64[parity][mu][9 x double2][Vh].
#pragma once
#include "cmx_internal.h"

namespace cmx {

// ---- coordinates ----------------------------------------------------------------------------------
__device__ __forceinline__ void cb_coords(const Geom& g, int cb, int i, int (&x)[4]) {
  int r = i;
  const int xh = r % g.Lxh;
  r /= g.Lxh;
  x[1] = r % g.L[1];
  r /= g.L[1];
  x[2] = r % g.L[2];
  x[3] = r / g.L[2];
  x[0] = 2 * xh + ((cb + x[1] + x[2] + x[3]) & 1);
}
__device__ __forceinline__ int cb_index(const Geom& g, const int (&x)[4]) {
  return ((x[3] * g.L[2] + x[2]) * g.L[1] + x[1]) * g.Lxh + (x[0] >> 1);
}
__device__ __forceinline__ int parity(const int (&x)[4]) { return (x[0] + x[1] + x[2] + x[3]) & 1; }

// ---- 3x3 complex FP64 matrices in registers -----------------------------------------------------------
struct M3 {
  double a[18];
};
__device__ __forceinline__ M3 mul(const M3& A, const M3& B) {
  M3 Cm;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double re = 0, im = 0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double ar = A.a[2 * (3 * i + k)], ai = A.a[2 * (3 * i + k) + 1];
        const double br = B.a[2 * (3 * k + j)], bi = B.a[2 * (3 * k + j) + 1];
        re += ar * br - ai * bi;
        im += ar * bi + ai * br;
      }
      Cm.a[2 * (3 * i + j)] = re;
      Cm.a[2 * (3 * i + j) + 1] = im;
    }
  return Cm;
}
__device__ __forceinline__ M3 dag(const M3& A) {
  M3 Cm;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      Cm.a[2 * (3 * i + j)] = A.a[2 * (3 * j + i)];
      Cm.a[2 * (3 * i + j) + 1] = -A.a[2 * (3 * j + i) + 1];
    }
  return Cm;
}

// Link U_mu at full coordinates x (periodic wrap applied by the caller)
__device__ __forceinline__ M3 get_link(const Geom& g, const double2* const (&g64)[2], int mu, const int (&x)[4]) {
  const int p = parity(x), i = cb_index(g, x);
  const double2* b = g64[p] + (size_t)mu * 9 * g.Vh;
  M3 u;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    double2 v = b[(size_t)k * g.Vh + i];
    u.a[2 * k] = v.x;
    u.a[2 * k + 1] = v.y;
  }
  return u;
}
__device__ __forceinline__ void shiftc(const Geom& g, const int (&x)[4], int mu, int d, int (&y)[4]) {
  y[0] = x[0], y[1] = x[1], y[2] = x[2], y[3] = x[3];
  y[mu] = (x[mu] + d + g.L[mu]) % g.L[mu];
}


}  // namespace cmx
