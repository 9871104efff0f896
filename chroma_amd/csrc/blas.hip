// Fused BLAS-1 kernels and the finalize/scalar-program kernel (see blas.h).
// The reductions follow the fused kernels of bicgstab_kernels_scalarsite.h:29-742 (one pass that
// updates a vector and returns its norm) with FP64 accumulation and a deterministic two-stage sum.
#include <cmath>

#include "blas.h"

namespace cmx {

namespace {

template <int P, typename R>
__device__ __forceinline__ void vload(const VRef& r, int i, R (&o)[24]) {
  typename SpinorIO<P>::R t[24];
  SpinorIO<P>::load(r.v, r.nrm, r.stride, i, t);
#pragma unroll
  for (int k = 0; k < 24; ++k) o[k] = (R)t[k];
}
template <int P, typename R>
__device__ __forceinline__ void vstore(const VRef& r, int i, const R (&o)[24]) {
  typename SpinorIO<P>::R t[24];
#pragma unroll
  for (int k = 0; k < 24; ++k) t[k] = (typename SpinorIO<P>::R)o[k];
  SpinorIO<P>::store(r.v, r.nrm, r.stride, i, t);
}

template <typename R> __device__ __forceinline__ double nrm2(const R (&v)[24]) {
  double s = 0;
#pragma unroll
  for (int k = 0; k < 24; ++k) s += (double)v[k] * (double)v[k];
  return s;
}

#define CMX_F(NAME, NR, LD, ST)                                                                      \
  struct NAME {                                                                                      \
    static constexpr int NRED = NR, LOAD = LD, STORE = ST;                                          \
    template <typename R>                                                                            \
    __device__ static void f(R(&v0)[24], R(&v1)[24], R(&v2)[24], R(&v3)[24], const double* S, int ia, \
                             int ib, double(&red)[4])

CMX_F(FCopy, 0, 2, 1) {
#pragma unroll
  for (int k = 0; k < 24; ++k) v0[k] = v1[k];
}
};
CMX_F(FZero, 0, 0, 1) {
#pragma unroll
  for (int k = 0; k < 24; ++k) v0[k] = 0;
}
};
CMX_F(FNorm2, 1, 1, 0) { red[0] += nrm2(v0); }
};
CMX_F(FXmyNorm, 1, 6, 1) {
#pragma unroll
  for (int k = 0; k < 24; ++k) v0[k] = v1[k] - v2[k];
  red[0] += nrm2(v0);
}
};
CMX_F(FXmyNormCopy, 1, 6, 9) {
#pragma unroll
  for (int k = 0; k < 24; ++k) {
    v0[k] = v1[k] - v2[k];
    v3[k] = v0[k];
  }
  red[0] += nrm2(v0);
}
};
CMX_F(FAxpyNorm, 1, 3, 1) {
  const R a = (R)S[ia];
#pragma unroll
  for (int k = 0; k < 24; ++k) v0[k] += a * v1[k];
  red[0] += nrm2(v0);
}
};
CMX_F(FAxpy, 0, 3, 1) {
  const R a = (R)S[ia];
#pragma unroll
  for (int k = 0; k < 24; ++k) v0[k] += a * v1[k];
}
};
CMX_F(FXpay, 0, 3, 1) {
  const R a = (R)S[ia];
#pragma unroll
  for (int k = 0; k < 24; ++k) v0[k] = v1[k] + a * v0[k];
}
};
CMX_F(FAxpby, 0, 3, 1) {
  const R a = (R)S[ia], b = (R)S[ib];
#pragma unroll
  for (int k = 0; k < 24; ++k) v0[k] = a * v1[k] + b * v0[k];
}
};
CMX_F(FCgUpdate, 0, 7, 3) {
  const R a = (R)S[ia], b = (R)S[ib];
#pragma unroll
  for (int k = 0; k < 24; ++k) {
    v0[k] += a * v1[k];
    v1[k] = v2[k] + b * v1[k];
  }
}
};
CMX_F(FAxpy2Norm, 1, 15, 3) {
  const R a = (R)S[ia];
#pragma unroll
  for (int k = 0; k < 24; ++k) {
    v0[k] += a * v2[k];
    v1[k] -= a * v3[k];
  }
  red[0] += nrm2(v1);
}
};
CMX_F(FAddZeroCopy, 0, 11, 7) {
#pragma unroll
  for (int k = 0; k < 24; ++k) {
    v0[k] += v1[k];
    v1[k] = 0;
    v2[k] = v3[k];
  }
}
};
CMX_F(FAdd, 0, 3, 1) {
#pragma unroll
  for (int k = 0; k < 24; ++k) v0[k] += v1[k];
}
};
CMX_F(FCdot, 2, 3, 0) {
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const double ar = v0[2 * k], ai = v0[2 * k + 1], br = v1[2 * k], bi = v1[2 * k + 1];
    red[0] += ar * br + ai * bi;
    red[1] += ar * bi - ai * br;
  }
}
};
CMX_F(FBicgP, 0, 7, 1) {
  const R wr = (R)S[ia], wi = (R)S[ia + 1], br = (R)S[ib], bi = (R)S[ib + 1];
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const R tr = v0[2 * k] - (wr * v2[2 * k] - wi * v2[2 * k + 1]);
    const R ti = v0[2 * k + 1] - (wr * v2[2 * k + 1] + wi * v2[2 * k]);
    v0[2 * k] = v1[2 * k] + (br * tr - bi * ti);
    v0[2 * k + 1] = v1[2 * k + 1] + (br * ti + bi * tr);
  }
}
};
CMX_F(FCaxpy, 0, 3, 1) {
  const R ar = (R)S[ia], ai = (R)S[ia + 1];
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const R xr = v1[2 * k], xi = v1[2 * k + 1];
    v0[2 * k] += ar * xr - ai * xi;
    v0[2 * k + 1] += ar * xi + ai * xr;
  }
}
};
CMX_F(FTnormTdot, 3, 3, 0) {
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const double ar = v0[2 * k], ai = v0[2 * k + 1], br = v1[2 * k], bi = v1[2 * k + 1];
    red[0] += ar * ar + ai * ai;
    red[1] += ar * br + ai * bi;
    red[2] += ar * bi - ai * br;
  }
}
};
CMX_F(FBicgUpdate, 1, 15, 3) {
  const R wr = (R)S[ia], wi = (R)S[ia + 1], ar = (R)S[ib], ai = (R)S[ib + 1];
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const R rr = v1[2 * k], ri = v1[2 * k + 1], pr = v2[2 * k], pi = v2[2 * k + 1];
    v0[2 * k] += (wr * rr - wi * ri) + (ar * pr - ai * pi);
    v0[2 * k + 1] += (wr * ri + wi * rr) + (ar * pi + ai * pr);
    const R tr = v3[2 * k], ti = v3[2 * k + 1];
    v1[2 * k] = rr - (wr * tr - wi * ti);
    v1[2 * k + 1] = ri - (wr * ti + wi * tr);
  }
  red[0] += nrm2(v1);
}
};
CMX_F(FCaxpy2, 0, 15, 5) {
  const R ar = (R)S[ia], ai = (R)S[ia + 1], br = (R)S[ib], bi = (R)S[ib + 1];
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const R xr = v1[2 * k], xi = v1[2 * k + 1], yr = v3[2 * k], yi = v3[2 * k + 1];
    v0[2 * k] += ar * xr - ai * xi;
    v0[2 * k + 1] += ar * xi + ai * xr;
    v2[2 * k] += br * yr - bi * yi;
    v2[2 * k + 1] += br * yi + bi * yr;
  }
}
};
CMX_F(FRbicgUpdate, 3, 15, 3) {
  const R wr = (R)S[ia], wi = (R)S[ia + 1];
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const R sr = v1[2 * k], si = v1[2 * k + 1], tr = v2[2 * k], ti = v2[2 * k + 1];
    v0[2 * k] += wr * sr - wi * si;
    v0[2 * k + 1] += wr * si + wi * sr;
    v1[2 * k] = sr - (wr * tr - wi * ti);
    v1[2 * k + 1] = si - (wr * ti + wi * tr);
    const double rr = v1[2 * k], ri = v1[2 * k + 1], zr = v3[2 * k], zi = v3[2 * k + 1];
    red[0] += rr * rr + ri * ri;
    red[1] += zr * rr + zi * ri;
    red[2] += zr * ri - zi * rr;
  }
}
};
CMX_F(FScale, 0, 2, 1) {
  const R a = (R)S[ia];
#pragma unroll
  for (int k = 0; k < 24; ++k) v0[k] = a * v1[k];
}
};
CMX_F(FAxpyCNorm, 1, 3, 1) {
  const R ar = (R)S[ia], ai = (R)S[ia + 1];
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const R xr = v1[2 * k], xi = v1[2 * k + 1];
    v0[2 * k] += ar * xr - ai * xi;
    v0[2 * k + 1] += ar * xi + ai * xr;
  }
  red[0] += nrm2(v0);
}
};
// the two updates of an MR step in one pass: the same arithmetic as FCaxpy then FAxpyCNorm (v0 is
// updated from v1 before v1 changes)
CMX_F(FMrUpdate, 1, 7, 3) {
  const R ar = (R)S[ia], ai = (R)S[ia + 1], br = (R)S[ib], bi = (R)S[ib + 1];
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const R xr = v1[2 * k], xi = v1[2 * k + 1], yr = v2[2 * k], yi = v2[2 * k + 1];
    v0[2 * k] += ar * xr - ai * xi;
    v0[2 * k + 1] += ar * xi + ai * xr;
    v1[2 * k] += br * yr - bi * yi;
    v1[2 * k + 1] += br * yi + bi * yr;
  }
  red[0] += nrm2(v1);
}
};
#undef CMX_F

// work-group blk of the launch b
template <class F, int P0, int P1, int P2, int P3>
__device__ __forceinline__ void blas_body(const BlasLaunch& b, int blk) {
  if (b.skip_flag >= 0 && b.F[b.skip_flag]) return;
  if (b.run_flag >= 0 && !b.F[b.run_flag]) return;
  constexpr int U = F::LOAD | F::STORE;
  constexpr bool D = ((U & 1) && P0 == P64) || ((U & 2) && P1 == P64) || ((U & 4) && P2 == P64) ||
                     ((U & 8) && P3 == P64);
  using R = typename std::conditional<D, double, float>::type;
  double red[4] = {0, 0, 0, 0};
  const int i = blk * blockDim.x + threadIdx.x;
  if (i < b.n) {
    R v0[24], v1[24], v2[24], v3[24];
    if constexpr (F::LOAD & 1) vload<P0>(b.v[0], i, v0);
    if constexpr (F::LOAD & 2) vload<P1>(b.v[1], i, v1);
    if constexpr (F::LOAD & 4) vload<P2>(b.v[2], i, v2);
    if constexpr (F::LOAD & 8) vload<P3>(b.v[3], i, v3);
    F::f(v0, v1, v2, v3, b.S, b.ia, b.ib, red);
    if constexpr (F::STORE & 1) vstore<P0>(b.v[0], i, v0);
    if constexpr (F::STORE & 2) vstore<P1>(b.v[1], i, v1);
    if constexpr (F::STORE & 4) vstore<P2>(b.v[2], i, v2);
    if constexpr (F::STORE & 8) vstore<P3>(b.v[3], i, v3);
  }
  if constexpr (F::NRED > 0) {
    double v[F::NRED];
#pragma unroll
    for (int r = 0; r < F::NRED; ++r) v[r] = red[r];
    block_sum<F::NRED>(v);
    if (threadIdx.x == 0)
#pragma unroll
      for (int r = 0; r < F::NRED; ++r) b.partials[(size_t)blk * F::NRED + r] = v[r];
  }
}

template <class F, int P0, int P1, int P2, int P3>
__global__ void __launch_bounds__(256) k_blas(BlasLaunch b) {
  blas_body<F, P0, P1, P2, P3>(b, blockIdx.x);
}
template <class F, int P0, int P1, int P2, int P3>
__global__ void __launch_bounds__(256) k_blas_b(const BlasLaunch* __restrict__ tab, const int4* __restrict__ map) {
  const int4 m = map[blockIdx.x];
  blas_body<F, P0, P1, P2, P3>(tab[m.x], m.y);
}

template <class F>
void dispatch(const BlasLaunch& b, const int p[4], hipStream_t s, const BatchGrid* bg) {
  const dim3 grid(blas_blocks(b.n)), block(256);
  const int key = p[0] + 3 * p[1] + 9 * p[2] + 27 * p[3];
#define CMX_T(A, B, C, D)                                                                  \
  case A + 3 * B + 9 * C + 27 * D:                                                         \
    if (bg)                                                                                \
      hipLaunchKernelGGL((k_blas_b<F, A, B, C, D>), dim3(bg->nblocks), block, 0, s,       \
                         static_cast<const BlasLaunch*>(bg->tab), bg->map);               \
    else                                                                                   \
      hipLaunchKernelGGL((k_blas<F, A, B, C, D>), grid, block, 0, s, b);                  \
    return;
  switch (key) {
    CMX_T(P64, P64, P64, P64)
    CMX_T(P32, P32, P32, P32)
    CMX_T(P16, P16, P16, P16)
    // precision conversions and mixed-precision solver steps
    CMX_T(P64, P32, P64, P64)
    CMX_T(P64, P16, P64, P64)
    CMX_T(P32, P64, P32, P32)
    CMX_T(P16, P64, P16, P16)
    CMX_T(P32, P16, P32, P32)
    CMX_T(P16, P32, P16, P16)
    CMX_T(P64, P64, P64, P32)
    CMX_T(P64, P64, P64, P16)
    CMX_T(P32, P32, P16, P16)
    CMX_T(P16, P32, P32, P32)
    CMX_T(P16, P16, P32, P16)
    CMX_T(P32, P16, P16, P16)
    default:
      break;
  }
#undef CMX_T
  throw Error(CMX_ERR_UNSUPPORTED, "blas: precision combination not instantiated");
}

// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void finalize_body(const FinLaunch& L) {
  const double* partials = L.partials;
  const int nb = L.nb, nred = L.nred, prog = L.prog, slot = L.slot;
  double* S = L.S;
  int* F = L.F;
  double red[3] = {0, 0, 0};
  for (int r = 0; r < nred; ++r) {
    // each thread sums partials threadIdx.x, + blockDim.x, ... in that order; the loads of eight terms
    // are issued before their additions, so the sum waits for one load latency per eight terms (the
    // partials of a 48^3x96 reduction are 20736 per entry) and the order, hence the result, is unchanged
    double acc = 0;
    const int step = blockDim.x;
    int k = threadIdx.x;
    for (; k + 7 * step < nb; k += 8 * step) {
      double t[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) t[j] = partials[(size_t)(k + j * step) * nred + r];
#pragma unroll
      for (int j = 0; j < 8; ++j) acc += t[j];
    }
    for (; k < nb; k += step) acc += partials[(size_t)k * nred + r];
    red[r] = acc;
  }
  block_sum<3>(red);
  if (threadIdx.x != 0) return;
  if (slot >= 0)
    for (int r = 0; r < nred; ++r) S[slot + r] = red[r];
  switch (prog) {
    case PG_STORE:
      break;
    case PG_CG_INIT: {  // invcg2.cc:109-132
      S[S_CP] = red[0];
      S[S_C] = red[0];
      F[F_K] = 0;
      if (S[S_CP] <= S[S_RSDSQ]) {
        F[F_CONV] = 1;
        F[F_NCOUNT] = 0;
      }
      break;
    }
    case PG_CG_D: {  // invcg2.cc:160-175
      F[F_DONE] = F[F_CONV];
      if (F[F_DONE]) break;
      S[S_D] = red[0];
      if (!(S[S_D] > 0.0) || !isfinite(S[S_D])) {
        F[F_ERR] = CMX_ERR_BREAKDOWN;
        F[F_CONV] = 1;
        F[F_DONE] = 1;
        F[F_NCOUNT] = F[F_K];
        break;
      }
      S[S_A] = S[S_C] / S[S_D];
      S[S_NEGA] = -S[S_A];
      break;
    }
    case PG_CG_R: {  // invcg2.cc:178-219
      if (F[F_DONE]) break;
      S[S_CP] = red[0];
      const int k = ++F[F_K];
      S[S_B] = S[S_CP] / S[S_C];
      S[S_C] = S[S_CP];
      if (S[S_CP] <= S[S_RSDSQ]) {
        F[F_CONV] = 1;
        F[F_NCOUNT] = k;
      } else if (k >= F[F_MAXIT] || !isfinite(S[S_CP])) {
        F[F_CONV] = 1;
        F[F_NCOUNT] = k;
        F[F_ERR] = CMX_ERR_NOCONV;
      }
      break;
    }
    case PG_RCG_INIT: {  // reliable_cg.cc:60-70
      S[S_RSQ] = red[0];
      S[S_RNORM] = sqrt(red[0]);
      S[S_R0NORM] = S[S_RNORM];
      S[S_MAXRX] = S[S_RNORM];
      S[S_MAXRR] = S[S_RNORM];
      S[S_BETA] = 0.0;
      F[F_K] = 0;
      F[F_NUPD] = 0;
      break;
    }
    case PG_RCG_D: {  // reliable_cg.cc:84-96: c = r_sq; d = |mp|^2; a = c/d
      F[F_DONE] = F[F_CONV];
      if (F[F_DONE]) break;
      S[S_C] = S[S_RSQ];
      S[S_D] = red[0];
      if (!(S[S_D] > 0.0) || !isfinite(S[S_D])) {
        F[F_ERR] = CMX_ERR_BREAKDOWN;
        F[F_CONV] = 1;
        F[F_DONE] = 1;
        F[F_NCOUNT] = F[F_K];
        break;
      }
      S[S_A] = S[S_C] / S[S_D];
      S[S_NEGA] = -S[S_A];
      break;
    }
    case PG_RCG_R: {  // reliable_cg.cc:98-118
      if (F[F_DONE]) break;
      S[S_RSQ] = red[0];
      const double rn = sqrt(red[0]);
      S[S_RNORM] = rn;
      if (rn > S[S_MAXRX]) S[S_MAXRX] = rn;
      if (rn > S[S_MAXRR]) S[S_MAXRR] = rn;
      const int updX = (rn < S[S_DELTA] * S[S_R0NORM] && S[S_R0NORM] <= S[S_MAXRX]);
      const int updR = (rn < S[S_DELTA] * S[S_MAXRR] && S[S_R0NORM] <= S[S_MAXRR]) || updX;
      F[F_UPDX] = updX;
      F[F_UPDR] = updR;
      if (updR) F[F_NUPD]++;
      break;
    }
    case PG_RCG_UPD: {  // reliable_cg.cc:120-170 (update part, then the convergence test)
      if (F[F_DONE]) break;
      if (F[F_UPDR]) {
        S[S_RSQ] = red[0];
        S[S_RNORM] = sqrt(red[0]);
        S[S_MAXRR] = S[S_RNORM];
        if (F[F_UPDX]) {
          S[S_R0NORM] = S[S_RNORM];
          S[S_MAXRX] = S[S_RNORM];
        }
      }
      const int k = F[F_K];
      F[F_K] = k + 1;
      S[S_BETA] = S[S_RSQ] / S[S_C];
      if (S[S_RSQ] < S[S_RSDSQ]) {
        F[F_CONV] = 1;
        F[F_NCOUNT] = k;
      } else if (k + 1 >= F[F_MAXIT] || !isfinite(S[S_RSQ])) {
        F[F_CONV] = 1;
        F[F_NCOUNT] = k + 1;
        F[F_ERR] = CMX_ERR_NOCONV;
      }
      break;
    }
    case PG_BICG_RHO: {  // invbicgstab.cc:72-95
      F[F_DONE] = F[F_CONV];
      if (F[F_DONE]) break;
      const double rr = red[0], ri = red[1];
      S[S_RHO] = rr;
      S[S_RHO + 1] = ri;
      if (rr == 0.0 && ri == 0.0) {
        F[F_ERR] = CMX_ERR_BREAKDOWN;
        F[F_CONV] = F[F_DONE] = 1;
        F[F_NCOUNT] = F[F_K];
        break;
      }
      // beta = (rho / rho_prev) * (alpha / omega)
      auto cdiv = [](double ar, double ai, double br, double bi, double& cr, double& ci) {
        const double d = br * br + bi * bi;
        cr = (ar * br + ai * bi) / d;
        ci = (ai * br - ar * bi) / d;
      };
      double q1r, q1i, q2r, q2i;
      cdiv(rr, ri, S[S_RHOP], S[S_RHOP + 1], q1r, q1i);
      cdiv(S[S_ALPHA], S[S_ALPHA + 1], S[S_OMEGA], S[S_OMEGA + 1], q2r, q2i);
      S[S_BETAC] = q1r * q2r - q1i * q2i;
      S[S_BETAC + 1] = q1r * q2i + q1i * q2r;
      break;
    }
    case PG_BICG_ALPHA: {  // invbicgstab.cc:100-118
      if (F[F_DONE]) break;
      const double cr = red[0], ci = red[1];
      if (cr == 0.0 && ci == 0.0) {
        F[F_ERR] = CMX_ERR_BREAKDOWN;
        F[F_CONV] = F[F_DONE] = 1;
        F[F_NCOUNT] = F[F_K];
        break;
      }
      const double d = cr * cr + ci * ci;
      const double ar = (S[S_RHO] * cr + S[S_RHO + 1] * ci) / d;
      const double ai = (S[S_RHO + 1] * cr - S[S_RHO] * ci) / d;
      S[S_ALPHA] = ar;
      S[S_ALPHA + 1] = ai;
      S[S_NALPHA] = -ar;
      S[S_NALPHA + 1] = -ai;
      S[S_RHOP] = S[S_RHO];
      S[S_RHOP + 1] = S[S_RHO + 1];
      break;
    }
    case PG_BICG_OMEGA: {  // invbicgstab.cc:122-150
      if (F[F_DONE]) break;
      const double tn = red[0];
      if (tn == 0.0) {
        F[F_ERR] = CMX_ERR_BREAKDOWN;
        F[F_CONV] = F[F_DONE] = 1;
        F[F_NCOUNT] = F[F_K];
        break;
      }
      S[S_OMEGA] = red[1] / tn;
      S[S_OMEGA + 1] = red[2] / tn;
      S[S_NOMEGA] = -S[S_OMEGA];
      S[S_NOMEGA + 1] = -S[S_OMEGA + 1];
      break;
    }
    case PG_BICG_R: {  // invbicgstab.cc:160-172
      if (F[F_DONE]) break;
      const int k = ++F[F_K];
      S[S_CP] = red[0];
      if (red[0] < S[S_RSDSQ]) {
        F[F_CONV] = 1;
        F[F_NCOUNT] = k;
      } else if (k >= F[F_MAXIT] || !isfinite(red[0])) {
        F[F_CONV] = 1;
        F[F_NCOUNT] = k;
        F[F_ERR] = CMX_ERR_NOCONV;
      }
      break;
    }
    case PG_RBICG_INIT: {  // reliable_bicgstab.cc:61-104
      S[S_RSQ] = red[0];
      S[S_RNORM] = sqrt(red[0]);
      S[S_R0NORM] = S[S_MAXRX] = S[S_MAXRR] = S[S_RNORM];
      S[S_RHO] = red[0];  // <r0|r> = |r|^2 at the start (:109-113)
      S[S_RHO + 1] = 0.0;
      S[S_RHOP] = S[S_ALPHA] = S[S_OMEGA] = 1.0;
      S[S_RHOP + 1] = S[S_ALPHA + 1] = S[S_OMEGA + 1] = 0.0;
      S[S_BETAC] = S[S_BETAC + 1] = 0.0;  // first p = r (p = v = 0 on entry)
      F[F_K] = 0;
      F[F_NUPD] = 0;
      break;
    }
    case PG_RBICG_ALPHA: {  // reliable_bicgstab.cc:135-150
      F[F_DONE] = F[F_CONV];
      if (F[F_DONE]) break;
      const double cr = red[0], ci = red[1];
      if (cr == 0.0 && ci == 0.0) {
        F[F_ERR] = CMX_ERR_BREAKDOWN;
        F[F_CONV] = F[F_DONE] = 1;
        F[F_NCOUNT] = F[F_K];
        break;
      }
      const double d = cr * cr + ci * ci;
      S[S_ALPHA] = (S[S_RHO] * cr + S[S_RHO + 1] * ci) / d;
      S[S_ALPHA + 1] = (S[S_RHO + 1] * cr - S[S_RHO] * ci) / d;
      S[S_NALPHA] = -S[S_ALPHA];
      S[S_NALPHA + 1] = -S[S_ALPHA + 1];
      S[S_RHOP] = S[S_RHO];
      S[S_RHOP + 1] = S[S_RHO + 1];
      break;
    }
    case PG_RBICG_R:      // reliable_bicgstab.cc:199-214
    case PG_RBICG_UPD: {  // reliable_bicgstab.cc:216-271
      if (F[F_DONE]) break;
      if (prog == PG_RBICG_R) {
        S[S_RSQ] = red[0];
        S[S_RHO] = red[1];
        S[S_RHO + 1] = red[2];
        const double rn = sqrt(red[0]);
        S[S_RNORM] = rn;
        if (rn > S[S_MAXRX]) S[S_MAXRX] = rn;
        if (rn > S[S_MAXRR]) S[S_MAXRR] = rn;
        const int updX = (rn < S[S_DELTA] * S[S_R0NORM] && S[S_R0NORM] <= S[S_MAXRX]);
        const int updR = (rn < S[S_DELTA] * S[S_MAXRR] && S[S_R0NORM] <= S[S_MAXRR]) || updX;
        F[F_UPDX] = updX;
        F[F_UPDR] = updR;
        if (updR) {  // the host launches the FP64 update, then PG_RBICG_UPD finishes the iteration
          F[F_NUPD]++;
          break;
        }
      } else {  // r = b - A x in FP64 (rho keeps the value from the sloppy r, as the reference does)
        S[S_RSQ] = red[0];
        S[S_RNORM] = sqrt(red[0]);
        S[S_MAXRR] = S[S_RNORM];
        if (F[F_UPDX]) {
          S[S_R0NORM] = S[S_RNORM];
          S[S_MAXRX] = S[S_RNORM];
        }
      }
      const int k = F[F_K];
      F[F_K] = k + 1;
      if (S[S_RSQ] < S[S_RSDSQ]) {
        F[F_CONV] = 1;
        F[F_NCOUNT] = k;
        break;
      }
      if (k + 1 >= F[F_MAXIT] || !isfinite(S[S_RSQ])) {
        F[F_CONV] = 1;
        F[F_NCOUNT] = k + 1;
        F[F_ERR] = CMX_ERR_NOCONV;
        break;
      }
      // beta = (rho / rho_prev) (alpha / omega) for the next p (:115-117)
      auto cdiv = [](double ar, double ai, double br, double bi, double& cr, double& ci) {
        const double d = br * br + bi * bi;
        cr = (ar * br + ai * bi) / d;
        ci = (ai * br - ar * bi) / d;
      };
      double q1r, q1i, q2r, q2i;
      cdiv(S[S_RHO], S[S_RHO + 1], S[S_RHOP], S[S_RHOP + 1], q1r, q1i);
      cdiv(S[S_ALPHA], S[S_ALPHA + 1], S[S_OMEGA], S[S_OMEGA + 1], q2r, q2i);
      S[S_BETAC] = q1r * q2r - q1i * q2i;
      S[S_BETAC + 1] = q1r * q2i + q1i * q2r;
      break;
    }
    case PG_MR_INIT: {
      S[S_MR + 8] = red[0];
      const double rsd = S[S_MR + 10];
      S[S_MR + 11] = rsd * rsd * S[S_MR + 7];
      F[F_MRDONE] = red[0] <= S[S_MR + 11] ? 1 : 0;
      break;
    }
    case PG_MR_A: {
      if (F[F_MRDONE]) break;
      const double d = red[0], cr = red[1], ci = red[2], w = S[S_MR + 9];
      if (!(d > 0.0)) {  // M r = 0: nothing left to reduce
        F[F_MRDONE] = 1;
        break;
      }
      S[S_MR + 3] = w * cr / d;
      S[S_MR + 4] = w * ci / d;
      S[S_MR + 5] = -S[S_MR + 3];
      S[S_MR + 6] = -S[S_MR + 4];
      break;
    }
    case PG_MR_R: {
      if (F[F_MRDONE]) break;
      S[S_MR + 8] = red[0];
      if (red[0] <= S[S_MR + 11]) F[F_MRDONE] = 1;
      break;
    }
    default:
      break;
  }
}

__global__ void __launch_bounds__(256) k_finalize(FinLaunch L) { finalize_body(L); }
// one work-group per problem
__global__ void __launch_bounds__(256) k_finalize_b(const FinLaunch* __restrict__ tab) { finalize_body(tab[blockIdx.x]); }

}  // namespace

int blas_nred(int op) {
  switch (op) {
    case B_NORM2: case B_XMY_NORM: case B_XMY_NORM_COPY: case B_AXPY_NORM: case B_AXPY2_NORM: case B_BICG_UPDATE:
    case B_AXPY_CNORM: case B_MR_UPDATE:
      return 1;
    case B_CDOT:
      return 2;
    case B_TNORM_TDOT: case B_RBICG_UPDATE:
      return 3;
    default:
      return 0;
  }
}

int blas_blocks(int n) { return (n + 255) / 256; }

void launch_blas(int op, const BlasLaunch& b, const int prec[4], hipStream_t s, const BatchGrid* bg) {
  switch (op) {
    case B_COPY: dispatch<FCopy>(b, prec, s, bg); break;
    case B_ZERO: dispatch<FZero>(b, prec, s, bg); break;
    case B_NORM2: dispatch<FNorm2>(b, prec, s, bg); break;
    case B_XMY_NORM: dispatch<FXmyNorm>(b, prec, s, bg); break;
    case B_XMY_NORM_COPY: dispatch<FXmyNormCopy>(b, prec, s, bg); break;
    case B_AXPY_NORM: dispatch<FAxpyNorm>(b, prec, s, bg); break;
    case B_AXPY: dispatch<FAxpy>(b, prec, s, bg); break;
    case B_XPAY: dispatch<FXpay>(b, prec, s, bg); break;
    case B_AXPBY: dispatch<FAxpby>(b, prec, s, bg); break;
    case B_CG_UPDATE: dispatch<FCgUpdate>(b, prec, s, bg); break;
    case B_AXPY2_NORM: dispatch<FAxpy2Norm>(b, prec, s, bg); break;
    case B_ADD_ZERO_COPY: dispatch<FAddZeroCopy>(b, prec, s, bg); break;
    case B_ADD: dispatch<FAdd>(b, prec, s, bg); break;
    case B_CDOT: dispatch<FCdot>(b, prec, s, bg); break;
    case B_BICG_P: dispatch<FBicgP>(b, prec, s, bg); break;
    case B_CAXPY: dispatch<FCaxpy>(b, prec, s, bg); break;
    case B_TNORM_TDOT: dispatch<FTnormTdot>(b, prec, s, bg); break;
    case B_BICG_UPDATE: dispatch<FBicgUpdate>(b, prec, s, bg); break;
    case B_CAXPY2: dispatch<FCaxpy2>(b, prec, s, bg); break;
    case B_RBICG_UPDATE: dispatch<FRbicgUpdate>(b, prec, s, bg); break;
    case B_SCALE: dispatch<FScale>(b, prec, s, bg); break;
    case B_AXPY_CNORM: dispatch<FAxpyCNorm>(b, prec, s, bg); break;
    case B_MR_UPDATE: dispatch<FMrUpdate>(b, prec, s, bg); break;
    default: throw Error(CMX_ERR_ARG, "blas: unknown op");
  }
}

void launch_finalize(const FinLaunch& L, hipStream_t s) { hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, L); }
void launch_finalize_batch(const FinLaunch* tab, int nproblems, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize_b, dim3(nproblems), dim3(256), 0, s, tab);
}

}  // namespace cmx
