// Multigrid kernels: null-vector block orthonormalisation, restriction, prolongation, the probes that
// build the coarse operator, the coarse operator itself and the BLAS of the coarse GCR.
//
// Fine fields are the library's spinor layout (SpinorIO); the null vectors P_k are full-lattice FP32
// vectors (even sites then odd, stride 2 Vh), so chunks 0-2 of a site hold the spins 0,1 (chirality +)
// and chunks 3-5 the spins 2,3 (chirality -).  Coarse vectors are complex FP32 [block][c * nvec + k].
// Blocks have even extents, so a block's origin has even parity and its odd sites are enumerated as
// in the odd checkerboard.
#include "blas.h"
#include "mg.h"
#include "su3_dev.h"

namespace cmx {

namespace {

__device__ __forceinline__ void block_coords(const MGGeom& m, int b, int (&bc)[4]) {
  int r = b;
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    bc[mu] = r % m.Lc[mu];
    r /= m.Lc[mu];
  }
}

__device__ __forceinline__ int lex_of(const Geom& g, const int (&x)[4]) {
  return x[0] + g.L[0] * (x[1] + g.L[1] * (x[2] + g.L[2] * x[3]));
}

// full-lattice index (parity * Vh + lex / 2) of local site l of block b
__device__ __forceinline__ int full_site(const MGGeom& m, int b, int l) {
  int bc[4], x[4];
  block_coords(m, b, bc);
  int r = l;
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    x[mu] = bc[mu] * m.B[mu] + r % m.B[mu];
    r /= m.B[mu];
  }
  const int par = (x[0] + x[1] + x[2] + x[3]) & 1;
  return par * m.g.Vh + lex_of(m.g, x) / 2;
}

// odd-checkerboard index of local odd site j of block b
__device__ __forceinline__ int odd_site(const MGGeom& m, int b, int j) {
  int bc[4], x[4];
  block_coords(m, b, bc);
  const int bxh = m.B[0] >> 1;
  const int lxh = j % bxh;
  int r = j / bxh;
  const int ly = r % m.B[1];
  r /= m.B[1];
  const int lz = r % m.B[2];
  const int lt = r / m.B[2];
  x[0] = bc[0] * m.B[0] + 2 * lxh + ((1 + ly + lz + lt) & 1);
  x[1] = bc[1] * m.B[1] + ly;
  x[2] = bc[2] * m.B[2] + lz;
  x[3] = bc[3] * m.B[3] + lt;
  return lex_of(m.g, x) / 2;
}

__device__ __forceinline__ int block_of(const MGGeom& m, const int (&x)[4]) {
  return x[0] / m.B[0] + m.Lc[0] * (x[1] / m.B[1] + m.Lc[1] * (x[2] / m.B[2] + m.Lc[2] * (x[3] / m.B[3])));
}

// block of the full-lattice site hs
__device__ __forceinline__ int block_of_full(const MGGeom& m, int hs) {
  const int par = hs >= m.g.Vh ? 1 : 0;
  int x[4];
  cb_coords(m.g, par, hs - par * m.g.Vh, x);
  return block_of(m, x);
}

// Sums over the work-group (256 threads, 4 waves) of N complex values per thread: tot[j] (LDS) receives
// the sum of entry j; sh is LDS scratch of 4 N entries
template <int N, typename T>
__device__ __forceinline__ void group_reduce(const T (&v)[N], T* sh, T* tot) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    auto a = v[j].x, b = v[j].y;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, 64);
      b += __shfl_xor(b, off, 64);
    }
    if (lane == 0) sh[j * 4 + wid].x = a, sh[j * 4 + wid].y = b;
  }
  __syncthreads();
  const int nw = blockDim.x >> 6;
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    T t = sh[j * 4];
    for (int w = 1; w < nw; ++w) t.x += sh[j * 4 + w].x, t.y += sh[j * 4 + w].y;
    tot[j] = t;
  }
  __syncthreads();
}

// ---- block orthonormalisation -------------------------------------------------------------------------
struct OrthArgs {
  MGGeom m;
  float4* P[MG_MAXVEC];
};

__device__ __forceinline__ void load_half(const float4* P, int stride, int hs, int c, float (&v)[12]) {
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const float4 t = P[(3 * c + q) * stride + hs];
    v[4 * q] = t.x, v[4 * q + 1] = t.y, v[4 * q + 2] = t.z, v[4 * q + 3] = t.w;
  }
}
__device__ __forceinline__ void store_half(float4* P, int stride, int hs, int c, const float (&v)[12]) {
#pragma unroll
  for (int q = 0; q < 3; ++q) P[(3 * c + q) * stride + hs] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
}
__device__ __forceinline__ double2 dot6(const float (&a)[12], const float (&b)[12]) {  // conj(a) . b
  double re = 0, im = 0;
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    re += (double)a[2 * s] * b[2 * s] + (double)a[2 * s + 1] * b[2 * s + 1];
    im += (double)a[2 * s] * b[2 * s + 1] - (double)a[2 * s + 1] * b[2 * s];
  }
  return make_double2(re, im);
}

// one work-group per (block, chirality): classical Gram-Schmidt applied twice, then normalisation
__global__ void __launch_bounds__(256) k_mg_orth(OrthArgs A) {
  __shared__ double2 sh[MG_MAXVEC * 4], tot[MG_MAXVEC];
  const MGGeom& m = A.m;
  const int b = blockIdx.x >> 1, c = blockIdx.x & 1;
  const int stride = 2 * m.g.Vh;
  for (int k = 0; k < m.nvec; ++k) {
    for (int pass = 0; pass < 3; ++pass) {  // passes 0, 1: projections; pass 2: the norm
      if (pass < 2 && k == 0) continue;
      double2 acc[MG_MAXVEC];
#pragma unroll
      for (int j = 0; j < MG_MAXVEC; ++j) acc[j] = make_double2(0, 0);
      for (int l = threadIdx.x; l < m.bvol; l += blockDim.x) {
        const int hs = full_site(m, b, l);
        float vk[12];
        load_half(A.P[k], stride, hs, c, vk);
        if (pass == 2) {
          acc[0].x += dot6(vk, vk).x;
        } else {
#pragma unroll
          for (int j = 0; j < MG_MAXVEC; ++j) {
            if (j < k) {
              float vj[12];
              load_half(A.P[j], stride, hs, c, vj);
              const double2 d = dot6(vj, vk);
              acc[j].x += d.x, acc[j].y += d.y;
            }
          }
        }
      }
      group_reduce<MG_MAXVEC>(acc, sh, tot);
      for (int l = threadIdx.x; l < m.bvol; l += blockDim.x) {
        const int hs = full_site(m, b, l);
        float vk[12];
        load_half(A.P[k], stride, hs, c, vk);
        if (pass == 2) {
          const double n2 = tot[0].x;
          const float inv = n2 > 0 ? (float)(1.0 / sqrt(n2)) : 0.f;
#pragma unroll
          for (int q = 0; q < 12; ++q) vk[q] *= inv;
        } else {
          for (int j = 0; j < k; ++j) {
            float vj[12];
            load_half(A.P[j], stride, hs, c, vj);
            const float dr = (float)tot[j].x, di = (float)tot[j].y;
#pragma unroll
            for (int s = 0; s < 6; ++s) {
              vk[2 * s] -= dr * vj[2 * s] - di * vj[2 * s + 1];
              vk[2 * s + 1] -= dr * vj[2 * s + 1] + di * vj[2 * s];
            }
          }
        }
        store_half(A.P[k], stride, hs, c, vk);
      }
      __syncthreads();
    }
  }
}

// ---- restriction ------------------------------------------------------------------------------------
struct RestrictArgs {
  MGGeom m;
  MGVecs P;
  VRef v;
  float2* out;
};

// the 24 reals of null vector k at full-lattice site hs: the FP32 vectors (PP = P32) or their FP16
// copies (PP = P16); reals 0-11 are chirality +, 12-23 chirality -
template <int PP>
__device__ __forceinline__ void load_pvec(const MGVecs& P, int k, int stride, int hs, float (&pk)[24]) {
  if constexpr (PP == P16) {
    SpinorIO<P16>::load(P.H[k], P.Hn[k], stride, hs, pk);
  } else {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      float h[12];
      load_half(P.P[k], stride, hs, c, h);
#pragma unroll
      for (int q = 0; q < 12; ++q) pk[12 * c + q] = h[q];
    }
  }
}

template <int PR, int PP, bool FULL> __global__ void __launch_bounds__(256) k_mg_restrict(RestrictArgs A) {
  __shared__ float2 sh[MG_MAXD * 4], tot[MG_MAXD];
  const MGGeom& m = A.m;
  // work-group g runs on XCD g mod 8: each XCD takes a contiguous range of blocks, so neighbouring
  // blocks along x, whose sites share the cache lines of the fine fields, follow each other on one L2
  const int per = (m.Nb + 7) / 8;
  const int b = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (b >= m.Nb) return;  // the whole work-group: the padding of the last XCD's range
  const int stride = 2 * m.g.Vh;
  const int nsite = FULL ? m.bvol : m.bvol / 2;
  float2 acc[MG_MAXD];
#pragma unroll
  for (int j = 0; j < MG_MAXD; ++j) acc[j] = make_float2(0.f, 0.f);
  for (int l = threadIdx.x; l < nsite; l += blockDim.x) {
    int hs, iv;
    if (FULL) {
      hs = full_site(m, b, l);
      iv = hs;
    } else {
      iv = odd_site(m, b, l);
      hs = m.g.Vh + iv;
    }
    typename SpinorIO<PR>::R s[24];
    SpinorIO<PR>::load(A.v.v, A.v.nrm, A.v.stride, iv, s);
    float v[24];
#pragma unroll
    for (int q = 0; q < 24; ++q) v[q] = (float)s[q];
#pragma unroll
    for (int k = 0; k < MG_MAXVEC; ++k) {
      if (k < m.nvec) {
        float pk[24];
        load_pvec<PP>(A.P, k, stride, hs, pk);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          float re = 0.f, im = 0.f;
#pragma unroll
          for (int q = 0; q < 6; ++q) {
            const float ar = pk[12 * c + 2 * q], ai = pk[12 * c + 2 * q + 1], br = v[12 * c + 2 * q],
                        bi = v[12 * c + 2 * q + 1];
            re += ar * br + ai * bi;
            im += ar * bi - ai * br;
          }
          acc[c * MG_MAXVEC + k].x += re;
          acc[c * MG_MAXVEC + k].y += im;
        }
      }
    }
  }
  group_reduce<MG_MAXD>(acc, sh, tot);
  if ((int)threadIdx.x < m.D) {
    const int c = threadIdx.x / m.nvec, k = threadIdx.x % m.nvec;
    A.out[(size_t)b * m.D + threadIdx.x] = tot[c * MG_MAXVEC + k];
  }
}

// ---- prolongation -----------------------------------------------------------------------------------
struct ProlongArgs {
  MGGeom m;
  MGVecs P;
  const float2* e;
  VRef out;
  int add;
};

template <int PR, int PP, bool FULL> __global__ void __launch_bounds__(256) k_mg_prolong(ProlongArgs A) {
  const MGGeom& m = A.m;
  const int n = FULL ? 2 * m.g.Vh : m.g.Vh;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int hs = FULL ? i : m.g.Vh + i;
  const int b = block_of_full(m, hs);
  const int stride = 2 * m.g.Vh;
  float v[24];
#pragma unroll
  for (int q = 0; q < 24; ++q) v[q] = 0.f;
  const float2* e = A.e + (size_t)b * m.D;
#pragma unroll
  for (int k = 0; k < MG_MAXVEC; ++k) {
    if (k < m.nvec) {
      float pk[24];
      load_pvec<PP>(A.P, k, stride, hs, pk);
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const float2 w = e[c * m.nvec + k];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          v[12 * c + 2 * q] += pk[12 * c + 2 * q] * w.x - pk[12 * c + 2 * q + 1] * w.y;
          v[12 * c + 2 * q + 1] += pk[12 * c + 2 * q] * w.y + pk[12 * c + 2 * q + 1] * w.x;
        }
      }
    }
  }
  typename SpinorIO<PR>::R s[24];
  if (A.add) {
    SpinorIO<PR>::load(A.out.v, A.out.nrm, A.out.stride, i, s);
#pragma unroll
    for (int q = 0; q < 24; ++q) s[q] += v[q];
  } else {
#pragma unroll
    for (int q = 0; q < 24; ++q) s[q] = v[q];
  }
  SpinorIO<PR>::store(A.out.v, A.out.nrm, A.out.stride, i, s);
}

// ---- random starts of the null vectors: Z2 noise from a counter-based hash (splitmix64) --------------
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) k_mg_noise(VRef out, int n, unsigned long long seed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long h = mix64(seed ^ mix64((unsigned long long)i));
  float s[24];
#pragma unroll
  for (int q = 0; q < 24; ++q) s[q] = ((h >> q) & 1ull) ? 0.70710678f : -0.70710678f;
  SpinorIO<P32>::store(out.v, nullptr, out.stride, i, s);
}

// ---- the coarse operator from face-restricted probes ---------------------------------------------
// With a colouring of period 2 in every coarse direction (period 1 where the extent is 1), a probe on
// the blocks of colour cl reaches a block b of another colour only through one direction mu, from
// b + mu across b's +mu face and from b - mu across its -mu face: restricting M probe over the whole
// block (colour cl: the block's own entry) or over each of the two faces separates the stencil entries.
// One work-group per block; the entries go straight into column dof of Ac[b][d] (column-major).
struct FaceArgs {
  MGGeom m;
  MGVecs P;
  const int* colour;
  const int* nbr;
  int col, dof;
  VRef v;  // M probe, full lattice FP64
  float2* Ac;
};

__device__ __forceinline__ int face_site(const MGGeom& m, int b, int mu, int side, int f) {
  int bc[4], x[4];
  block_coords(m, b, bc);
  int r = f;
#pragma unroll
  for (int nu = 0; nu < 4; ++nu) {
    int l;
    if (nu == mu) {
      l = side ? m.B[nu] - 1 : 0;
    } else {
      l = r % m.B[nu];
      r /= m.B[nu];
    }
    x[nu] = bc[nu] * m.B[nu] + l;
  }
  const int par = (x[0] + x[1] + x[2] + x[3]) & 1;
  return par * m.g.Vh + lex_of(m.g, x) / 2;
}

__global__ void __launch_bounds__(256) k_mg_face_restrict(FaceArgs A) {
  __shared__ float2 sh[MG_MAXD * 4], tot[MG_MAXD];
  const MGGeom& m = A.m;
  // each XCD takes a contiguous range of blocks (the order of k_mg_restrict)
  const int per = (m.Nb + 7) / 8;
  const int b = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (b >= m.Nb) return;  // the whole work-group: the padding of the last XCD's range
  const int stride = 2 * m.g.Vh;
  int mu = -1;
  if (A.colour[b] != A.col) {
    for (int nu = 0; nu < 4; ++nu)
      if (A.colour[A.nbr[b * MG_NST + 1 + 2 * nu]] == A.col) {
        mu = nu;
        break;
      }
    if (mu < 0) return;  // no stencil entry of this block has the probed colour
  }
  const int nregion = mu < 0 ? 1 : 2;
  for (int reg = 0; reg < nregion; ++reg) {
    const int nsite = mu < 0 ? m.bvol : m.bvol / m.B[mu];
    float2 acc[MG_MAXD];
#pragma unroll
    for (int j = 0; j < MG_MAXD; ++j) acc[j] = make_float2(0.f, 0.f);
    for (int l = threadIdx.x; l < nsite; l += blockDim.x) {
      const int hs = mu < 0 ? full_site(m, b, l) : face_site(m, b, mu, reg == 0 ? 1 : 0, l);
      double sd[24];
      SpinorIO<P64>::load(A.v.v, nullptr, A.v.stride, hs, sd);
      float v[24];
#pragma unroll
      for (int q = 0; q < 24; ++q) v[q] = (float)sd[q];
#pragma unroll
      for (int k = 0; k < MG_MAXVEC; ++k) {
        if (k < m.nvec) {
#pragma unroll
          for (int c = 0; c < 2; ++c) {
            float pk[12];
            load_half(A.P.P[k], stride, hs, c, pk);
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
              const float ar = pk[2 * q], ai = pk[2 * q + 1], br = v[12 * c + 2 * q], bi = v[12 * c + 2 * q + 1];
              re += ar * br + ai * bi;
              im += ar * bi - ai * br;
            }
            acc[c * MG_MAXVEC + k].x += re;
            acc[c * MG_MAXVEC + k].y += im;
          }
        }
      }
    }
    group_reduce<MG_MAXD>(acc, sh, tot);
    // stencil entry: the block itself, or +mu (region 0, the +mu face) / -mu (region 1)
    const int d = mu < 0 ? 0 : 1 + 2 * mu + reg;
    if ((int)threadIdx.x < m.D) {
      const int c = threadIdx.x / m.nvec, k = threadIdx.x % m.nvec;
      A.Ac[(((size_t)b * MG_NST + d) * m.D + A.dof) * m.D + threadIdx.x] = tot[c * MG_MAXVEC + k];
    }
    __syncthreads();
  }
}

// ---- probes and the coarse operator ---------------------------------------------------------------
struct ProbeArgs {
  MGGeom m;
  MGVecs P;
  const int* colour;
  int col, k, c;
  VRef out;
};

__global__ void __launch_bounds__(256) k_mg_probe(ProbeArgs A) {
  const MGGeom& m = A.m;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= 2 * m.g.Vh) return;
  double s[24];
#pragma unroll
  for (int q = 0; q < 24; ++q) s[q] = 0.0;
  if (A.colour[block_of_full(m, hs)] == A.col) {
    float pk[12];
    load_half(A.P.P[A.k], 2 * m.g.Vh, hs, A.c, pk);
#pragma unroll
    for (int q = 0; q < 12; ++q) s[12 * A.c + q] = pk[q];
  }
  SpinorIO<P64>::store(A.out.v, nullptr, A.out.stride, hs, s);
}

__global__ void __launch_bounds__(256) k_mg_scatter(MGGeom m, const int* nbr, const int* colour, int col, int dof,
                                                    const float2* y, float2* Ac) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m.Nb * m.D) return;
  const int b = t / m.D, r = t % m.D;
  for (int d = 0; d < MG_NST; ++d) {
    if (colour[nbr[b * MG_NST + d]] == col) {
      Ac[(((size_t)b * MG_NST + d) * m.D + dof) * m.D + r] = y[t];
      return;
    }
  }
}

__device__ __forceinline__ float2 ac_load(const float2* a, size_t i) { return a[i]; }
__device__ __forceinline__ float2 ac_load(const __half2* a, size_t i) { return __half22float2(a[i]); }

// AT: float2 (the set-up's FP32 entries) or __half2 (their FP16 copy, read by an FP16 V-cycle)
template <typename AT>
__global__ void __launch_bounds__(256) k_mg_coarse_apply(MGGeom m, const int* nbr, const AT* Ac, const float2* x,
                                                         float2* y, const int* skip) {
  if (skip && *skip) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m.Nb * m.D) return;
  const int b = t / m.D, r = t % m.D;
  float re = 0.f, im = 0.f;
  for (int d = 0; d < MG_NST; ++d) {
    // column-major entry: the D lanes of a block read one contiguous column per step
    const AT* a = Ac + ((size_t)b * MG_NST + d) * m.D * m.D + r;
    const float2* xv = x + (size_t)nbr[b * MG_NST + d] * m.D;
    for (int c = 0; c < m.D; ++c) {
      const float2 av = ac_load(a, (size_t)c * m.D), xx = xv[c];
      re += av.x * xx.x - av.y * xx.y;
      im += av.x * xx.y + av.y * xx.x;
    }
  }
  y[t] = make_float2(re, im);
}

__global__ void __launch_bounds__(256) k_mg_to_half(long n, const float2* a, __half2* h) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) h[i] = __float22half2_rn(a[i]);
}

// ---- coarse BLAS ------------------------------------------------------------------------------------
struct CdotArgs {
  const float2* v[MG_MAXKRY + 2];
  const float2* w;
  int n;
  double2* partials;
  const int* skip;
};

__global__ void __launch_bounds__(256) k_mg_cdot(CdotArgs A) {
  __shared__ double2 sh[4], tot[1];
  if (A.skip && *A.skip) return;
  const float2* v = A.v[blockIdx.y];
  double re = 0, im = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += gridDim.x * blockDim.x) {
    const float2 a = v[i], b = A.w[i];
    re += (double)a.x * b.x + (double)a.y * b.y;
    im += (double)a.x * b.y - (double)a.y * b.x;
  }
  const double2 acc[1] = {make_double2(re, im)};
  group_reduce<1>(acc, sh, tot);
  if (threadIdx.x == 0) A.partials[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = tot[0];
}

// one work-group: the m dot products from their partials, then the stage's scalar work in thread 0
__global__ void __launch_bounds__(256) k_mg_gcr_fin(const double2* part, int nblk, int m, int j, int stage,
                                                    double tol2, double* S, int* F) {
  __shared__ double2 sh[4], tot[1], d[MG_MAXKRY + 2];
  if (stage != MG_FIN_INIT && F[MGF_SKIP]) {
    if (threadIdx.x == 0) F[MGF_UPD] = 0;
    return;
  }
  for (int q = 0; q < m; ++q) {
    double re = 0, im = 0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) re += part[q * nblk + b].x, im += part[q * nblk + b].y;
    const double2 acc[1] = {make_double2(re, im)};
    group_reduce<1>(acc, sh, tot);
    if (threadIdx.x == 0) d[q] = tot[0];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  if (stage == MG_FIN_INIT) {
    const double b2 = d[0].x;
    S[MGS_B2] = b2;
    S[MGS_R2] = b2;
    S[MGS_STOP] = tol2 * b2;
    F[MGF_SKIP] = b2 > 0 ? 0 : 1;
    F[MGF_UPD] = 0;
  } else if (stage == MG_FIN_RESTART) {
    S[MGS_R2] = d[0].x;
    if (d[0].x <= S[MGS_STOP]) F[MGF_SKIP] = 1;
    F[MGF_UPD] = 0;
  } else {
    double qq = d[j].x;
    for (int i = 0; i < j; ++i) {
      qq -= d[i].x * d[i].x + d[i].y * d[i].y;
      S[MGS_C + 2 * i] = d[i].x;
      S[MGS_C + 2 * i + 1] = d[i].y;
    }
    if (!(qq > 1e-30 * S[MGS_B2])) {  // no new direction
      F[MGF_SKIP] = 1;
      F[MGF_UPD] = 0;
      return;
    }
    const double nq = sqrt(qq);
    // alpha = <q_j', r> / |q_j'| = conj(d[j + 1]) / nq   (r is orthogonal to the earlier q_i)
    const double ar = d[j + 1].x / nq, ai = -d[j + 1].y / nq;
    S[MGS_INV] = 1.0 / nq;
    S[MGS_AR] = ar;
    S[MGS_AI] = ai;
    S[MGS_R2] -= ar * ar + ai * ai;
    F[MGF_UPD] = 1;
    if (S[MGS_R2] <= S[MGS_STOP]) F[MGF_SKIP] = 1;
  }
}

__global__ void __launch_bounds__(256) k_mg_update(MGUpdate u) {
  if (!u.F[MGF_UPD]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= u.n) return;
  float2 q = u.q[u.j][i], p = u.p[u.j][i];
  for (int k = 0; k < u.j; ++k) {
    const float cr = (float)u.S[MGS_C + 2 * k], ci = (float)u.S[MGS_C + 2 * k + 1];
    const float2 qk = u.q[k][i], pk = u.p[k][i];
    q.x -= cr * qk.x - ci * qk.y;
    q.y -= cr * qk.y + ci * qk.x;
    p.x -= cr * pk.x - ci * pk.y;
    p.y -= cr * pk.y + ci * pk.x;
  }
  const float inv = (float)u.S[MGS_INV], ar = (float)u.S[MGS_AR], ai = (float)u.S[MGS_AI];
  q.x *= inv, q.y *= inv, p.x *= inv, p.y *= inv;
  u.q[u.j][i] = q;
  u.p[u.j][i] = p;
  float2 x = u.x[i], r = u.r[i];
  x.x += ar * p.x - ai * p.y;
  x.y += ar * p.y + ai * p.x;
  r.x -= ar * q.x - ai * q.y;
  r.y -= ar * q.y + ai * q.x;
  u.x[i] = x;
  u.r[i] = r;
}

int grid_of(long n) { return (int)((n + 255) / 256); }

}  // namespace

void launch_mg_orth(const MGGeom& m, float4* const* P, hipStream_t s) {
  OrthArgs A;
  A.m = m;
  for (int k = 0; k < MG_MAXVEC; ++k) A.P[k] = k < m.nvec ? P[k] : nullptr;
  hipLaunchKernelGGL(k_mg_orth, dim3(2 * m.Nb), dim3(256), 0, s, A);
}

void launch_mg_restrict(const MGGeom& m, const MGVecs& P, int prec, VRef v, bool full, float2* out, hipStream_t s) {
  RestrictArgs A{m, P, v, out};
  const int per = (m.Nb + 7) / 8;  // blocks per XCD (k_mg_restrict's work-group order)
  // one thread per site of the block where that fits a work-group of 64-256 threads (a 4^4 block's odd
  // sites: 128 threads, two waves, instead of four waves half idle)
  const int nsite = full ? m.bvol : m.bvol / 2;
  const int threads = std::min(256, std::max(64, (nsite + 63) / 64 * 64));
  auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(8 * per), dim3(threads), 0, s, A); };
  if (full) {
    if (prec == P64) go(k_mg_restrict<P64, P32, true>);
    else if (prec == P32) go(k_mg_restrict<P32, P32, true>);
    else go(k_mg_restrict<P16, P32, true>);
  } else {
    if (prec == P64) go(k_mg_restrict<P64, P32, false>);
    else if (prec == P32) go(k_mg_restrict<P32, P32, false>);
    else if (P.H[0]) go(k_mg_restrict<P16, P16, false>);  // FP16 V-cycle: the FP16 copies
    else go(k_mg_restrict<P16, P32, false>);
  }
}

void launch_mg_prolong(const MGGeom& m, const MGVecs& P, int prec, const float2* e, VRef out, bool full, bool add,
                       hipStream_t s) {
  ProlongArgs A{m, P, e, out, add ? 1 : 0};
  const int g = grid_of(full ? 2L * m.g.Vh : (long)m.g.Vh);
  auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(g), dim3(256), 0, s, A); };
  if (full) {
    if (prec == P64) go(k_mg_prolong<P64, P32, true>);
    else if (prec == P32) go(k_mg_prolong<P32, P32, true>);
    else go(k_mg_prolong<P16, P32, true>);
  } else {
    if (prec == P64) go(k_mg_prolong<P64, P32, false>);
    else if (prec == P32) go(k_mg_prolong<P32, P32, false>);
    else if (P.H[0]) go(k_mg_prolong<P16, P16, false>);  // FP16 V-cycle: the FP16 copies
    else go(k_mg_prolong<P16, P32, false>);
  }
}

void launch_mg_probe(const MGGeom& m, const MGVecs& P, const int* colour, int col, int k, int c, VRef out,
                     hipStream_t s) {
  ProbeArgs A{m, P, colour, col, k, c, out};
  hipLaunchKernelGGL(k_mg_probe, dim3(grid_of(2L * m.g.Vh)), dim3(256), 0, s, A);
}

void launch_mg_noise(VRef out, int n, unsigned long long seed, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_noise, dim3(grid_of(n)), dim3(256), 0, s, out, n, seed);
}

void launch_mg_face_restrict(const MGGeom& m, const MGVecs& P, const int* colour, const int* nbr, int col, int dof,
                             VRef v, float2* Ac, hipStream_t s) {
  FaceArgs A{m, P, colour, nbr, col, dof, v, Ac};
  const int per = (m.Nb + 7) / 8;  // blocks per XCD (k_mg_face_restrict's work-group order)
  hipLaunchKernelGGL(k_mg_face_restrict, dim3(8 * per), dim3(256), 0, s, A);
}

void launch_mg_scatter(const MGGeom& m, const int* nbr, const int* colour, int col, int dof, const float2* y,
                       float2* Ac, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_scatter, dim3(grid_of((long)m.Nb * m.D)), dim3(256), 0, s, m, nbr, colour, col, dof, y, Ac);
}

void launch_mg_coarse_apply(const MGGeom& m, const int* nbr, const float2* Ac, const float2* x, float2* y,
                            const int* skip, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_coarse_apply<float2>, dim3(grid_of((long)m.Nb * m.D)), dim3(256), 0, s, m, nbr, Ac, x, y,
                     skip);
}

void launch_mg_coarse_apply_h(const MGGeom& m, const int* nbr, const void* Ah, const float2* x, float2* y,
                              const int* skip, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_coarse_apply<__half2>, dim3(grid_of((long)m.Nb * m.D)), dim3(256), 0, s, m, nbr,
                     static_cast<const __half2*>(Ah), x, y, skip);
}

void launch_mg_to_half(long n, const float2* a, void* h, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_to_half, dim3(grid_of(n)), dim3(256), 0, s, n, a, static_cast<__half2*>(h));
}

int mg_cdot_blocks(int n) { return std::max(1, std::min(256, grid_of(n))); }

void launch_mg_cdot(const float2* const* v, int m, const float2* w, int n, double2* partials, const int* skip,
                    hipStream_t s) {
  CdotArgs A;
  for (int j = 0; j < MG_MAXKRY + 2; ++j) A.v[j] = j < m ? v[j] : nullptr;
  A.w = w;
  A.n = n;
  A.partials = partials;
  A.skip = skip;
  hipLaunchKernelGGL(k_mg_cdot, dim3(mg_cdot_blocks(n), m), dim3(256), 0, s, A);
}

void launch_mg_gcr_fin(const double2* partials, int nblk, int m, int j, int stage, double tol2, double* S, int* F,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_mg_gcr_fin, dim3(1), dim3(256), 0, s, partials, nblk, m, j, stage, tol2, S, F);
}

void launch_mg_update(const MGUpdate& u, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_update, dim3(grid_of(u.n)), dim3(256), 0, s, u);
}

}  // namespace cmx
