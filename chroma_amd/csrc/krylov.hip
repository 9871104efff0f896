// Multi-vector kernels of the Krylov-basis solvers (FGMRES-DR, syssolver_linop_fgmres_dr.cc:78-183):
// all inner products of one vector with a basis in one pass, and a linear combination of a basis
// added to one vector in one pass.  Reductions are two-stage and deterministic: each work-group
// writes its partial sums, and one work-group per column adds them in work-group order.
#include "blas.h"

namespace cmx {

namespace {

template <int P>
__device__ __forceinline__ void kload(const VRef& r, int i, float (&o)[24]) {
  SpinorIO<P>::load(r.v, r.nrm, r.stride, i, o);
}
template <int P>
__device__ __forceinline__ void kload(const VRef& r, int i, double (&o)[24]) {
  typename SpinorIO<P>::R t[24];
  SpinorIO<P>::load(r.v, r.nrm, r.stride, i, t);
#pragma unroll
  for (int k = 0; k < 24; ++k) o[k] = (double)t[k];
}

// partials[blockIdx][2i, 2i+1] = sum over this work-group's sites of <v_i, w>.  The basis is taken
// in groups of 8 vectors, so w is read once per group.
template <int PV, int PW>
__global__ void __launch_bounds__(256) k_mdot(MultiLaunch L) {
  constexpr int G = 8;
  __shared__ double sh[2 * KRY_MAXV][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  using R = typename std::conditional<PV == P64 || PW == P64, double, float>::type;
  for (int q0 = 0; q0 < L.m; q0 += G) {
    double d[G][2];
#pragma unroll
    for (int g = 0; g < G; ++g) d[g][0] = d[g][1] = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < L.n; i += gridDim.x * 256) {
      R w[24];
      kload<PW>(L.w, i, w);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        if (q0 + g < L.m) {
          R v[24];
          kload<PV>(L.v[q0 + g], i, v);
          double dr = 0, di = 0;
#pragma unroll
          for (int k = 0; k < 12; ++k) {
            const double ar = v[2 * k], ai = v[2 * k + 1], br = w[2 * k], bi = w[2 * k + 1];
            dr += ar * br + ai * bi;
            di += ar * bi - ai * br;
          }
          d[g][0] += dr;
          d[g][1] += di;
        }
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (q0 + g < L.m) {
        const double dr = wave_sum(d[g][0]), di = wave_sum(d[g][1]);
        if (lane == 0) {
          sh[2 * (q0 + g)][wid] = dr;
          sh[2 * (q0 + g) + 1][wid] = di;
        }
      }
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < 2 * L.m; r += 256)
    L.partials[(size_t)blockIdx.x * 2 * L.m + r] = sh[r][0] + sh[r][1] + sh[r][2] + sh[r][3];
}

// w (+)= sgn * sum_q c_q v_q (complex c_q = S[coef + 2q], S[coef + 2q + 1]); optional |w|^2 partials
template <int PV, int PW, bool NORM>
__global__ void __launch_bounds__(256) k_maxpy(MultiLaunch L) {
  __shared__ double cs[2 * KRY_MAXV];
  for (int r = threadIdx.x; r < 2 * L.m; r += 256) cs[r] = L.sgn * L.S[L.coef + r];
  __syncthreads();
  using R = typename std::conditional<PV == P64 || PW == P64, double, float>::type;
  double nsq = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < L.n; i += gridDim.x * 256) {
    R w[24];
    if (L.overwrite) {
#pragma unroll
      for (int k = 0; k < 24; ++k) w[k] = 0;
    } else {
      kload<PW>(L.w, i, w);
    }
    for (int q = 0; q < L.m; ++q) {
      R v[24];
      kload<PV>(L.v[q], i, v);
      const R cr = (R)cs[2 * q], ci = (R)cs[2 * q + 1];
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        const R xr = v[2 * k], xi = v[2 * k + 1];
        w[2 * k] += cr * xr - ci * xi;
        w[2 * k + 1] += cr * xi + ci * xr;
      }
    }
    typename SpinorIO<PW>::R o[24];
#pragma unroll
    for (int k = 0; k < 24; ++k) o[k] = (typename SpinorIO<PW>::R)w[k];
    SpinorIO<PW>::store(L.w.v, L.w.nrm, L.w.stride, i, o);
    if constexpr (NORM) {
#pragma unroll
      for (int k = 0; k < 24; ++k) nsq += (double)o[k] * (double)o[k];
    }
  }
  if constexpr (NORM) {
    double v[1] = {nsq};
    block_sum<1>(v);
    if (threadIdx.x == 0) L.partials[blockIdx.x] = v[0];
  }
}

__global__ void __launch_bounds__(256) k_colsum(const double* partials, int nb, int nred, double* out, int mode) {
  const int r = blockIdx.x;
  double acc = 0;
  for (int b = threadIdx.x; b < nb; b += 256) acc += partials[(size_t)b * nred + r];
  double v[1] = {acc};
  block_sum<1>(v);
  if (threadIdx.x == 0) {
    out[r] = v[0];
    if (mode == 1 && r == 0) out[nred] = v[0] > 0 ? 1.0 / sqrt(v[0]) : 0.0;
  }
}

int kry_grid(int n) { return std::min(KRY_BLOCKS, (n + 255) / 256); }

}  // namespace

void launch_mdot(const MultiLaunch& L, int pv, int pw, hipStream_t s) {
  CMX_REQUIRE(L.m >= 1 && L.m <= KRY_MAXV, CMX_ERR_ARG, "mdot: basis size");
  const dim3 grid(kry_grid(L.n)), block(256);
#define CMX_K(A, B)                                                            \
  if (pv == A && pw == B) {                                                    \
    hipLaunchKernelGGL((k_mdot<A, B>), grid, block, 0, s, L);                  \
    return;                                                                    \
  }
  CMX_K(P64, P64) CMX_K(P32, P32) CMX_K(P16, P16) CMX_K(P32, P64) CMX_K(P16, P64) CMX_K(P16, P32)
#undef CMX_K
  throw Error(CMX_ERR_UNSUPPORTED, "mdot: precision combination not instantiated");
}

void launch_maxpy(const MultiLaunch& L, int pv, int pw, bool norm, hipStream_t s) {
  CMX_REQUIRE(L.m >= 1 && L.m <= KRY_MAXV, CMX_ERR_ARG, "maxpy: basis size");
  const dim3 grid(kry_grid(L.n)), block(256);
#define CMX_K(A, B)                                                            \
  if (pv == A && pw == B) {                                                    \
    if (norm)                                                                  \
      hipLaunchKernelGGL((k_maxpy<A, B, true>), grid, block, 0, s, L);         \
    else                                                                       \
      hipLaunchKernelGGL((k_maxpy<A, B, false>), grid, block, 0, s, L);        \
    return;                                                                    \
  }
  CMX_K(P64, P64) CMX_K(P32, P32) CMX_K(P16, P16) CMX_K(P32, P64) CMX_K(P16, P64) CMX_K(P16, P32)
#undef CMX_K
  throw Error(CMX_ERR_UNSUPPORTED, "maxpy: precision combination not instantiated");
}

void launch_colsum(const double* partials, int nb, int nred, double* out, int mode, hipStream_t s) {
  hipLaunchKernelGGL(k_colsum, dim3(nred), dim3(256), 0, s, partials, nb, nred, out, mode);
}

int kry_blocks(int n) { return kry_grid(n); }

}  // namespace cmx
