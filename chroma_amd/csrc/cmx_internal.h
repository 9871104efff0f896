// Internal declarations of the MI355X Wilson-clover library (HIP, gfx950).
//
// Device layout (one checkerboard, "SoA in 16-byte chunks"):
//   fermion  : chunk k of site i at base[k*Vh + i]; FP64 12 x double2 (1 complex each),
//              FP32 6 x float4 (2 complex), FP16 3 x 16 B (4 complex as half2) + float norm[Vh]
//              (value = half * norm, norm = max |component| of the site)
//   link     : per (parity, mu) block of nchunk*Vh chunks; FP64/18 9 x double2, FP64/12 6 x double2,
//              FP32/18 9 x float2, FP32/12 3 x float4 (rows 0,1), FP16/12 a 16 B chunk (8 halves)
//              of every site followed by an 8 B chunk (4 halves) of every site: two loads per link
//   clover   : per site 72 reals {diag[0][6], offd[0][15] c, diag[1][6], offd[1][15] c}, so each
//              chiral block is 36 consecutive reals; FP64 36 x double2, FP32 18 x float4,
//              FP16 per block four 16 B chunks (reals 0-31) and one 8 B chunk (reals 32-35), the
//              eight 16 B chunks first, + float norm[Vh] (max |entry| of the site)
// Consecutive lanes own consecutive sites, so every chunk load of a wavefront is one contiguous
// 64 x 16 B (or 64 x 8 B) segment.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "chroma_mi355x.h"

namespace cmx {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define CMX_HIP(x)                                                                                   \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess) throw ::cmx::Error(CMX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CMX_REQUIRE(cond, code, msg)                  \
  do {                                                \
    if (!(cond)) throw ::cmx::Error((code), (msg));   \
  } while (0)

enum { P64 = CMX_DOUBLE, P32 = CMX_SINGLE, P16 = CMX_HALF };

// ------------------------------------------------------------------------------------------------
// Geometry passed by value to kernels
struct Geom {
  int L[4];
  int Lxh;  // Lx/2
  int Vh;   // sites per checkerboard
};

// Bytes per fermion site (one checkerboard) at a precision, excluding the FP16 norm
__host__ __device__ constexpr int spinor_chunks(int p) { return p == P64 ? 12 : (p == P32 ? 6 : 3); }
inline size_t spinor_bytes(int p, long Vh) {
  return (size_t)spinor_chunks(p) * 16 * Vh + (p == P16 ? sizeof(float) * Vh : 0);
}
__host__ __device__ constexpr int link_chunks(int p, int recon) { return recon == 18 ? 9 : (p == P64 ? 6 : 3); }
__host__ __device__ constexpr int link_chunk_bytes(int p, int recon) {
  return p == P64 ? 16 : (p == P32 ? (recon == 18 ? 8 : 16) : 8);
}
inline size_t clover_bytes(int p, long Vh) {
  return p == P64 ? (size_t)36 * 16 * Vh : (p == P32 ? (size_t)18 * 16 * Vh : (size_t)18 * 8 * Vh + 4 * Vh);
}

// A device fermion on one checkerboard
struct Field {
  void* v = nullptr;
  float* nrm = nullptr;  // FP16 only: points into the same allocation
  int prec = P64;
  long Vh = 0;
};

// ------------------------------------------------------------------------------------------------
// Complex helpers in registers
template <typename R> struct C {
  R re, im;
};
template <typename R> __device__ __forceinline__ C<R> cmul(C<R> a, C<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> __device__ __forceinline__ C<R> cmulc(C<R> a, C<R> b) {  // conj(a) * b
  return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}

// ------------------------------------------------------------------------------------------------
// Fermion load/store (24 reals: index 2*(spin*3+colour) + re/im)
template <int P> struct SpinorIO;

template <> struct SpinorIO<P64> {
  using R = double;
  template <int CP = 0>
  __device__ __forceinline__ static void load(const void* base, const float*, long Vh, long i, double (&s)[24]) {
    const double2* b = reinterpret_cast<const double2*>(base);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      double2 v = b[k * Vh + i];
      s[2 * k] = v.x;
      s[2 * k + 1] = v.y;
    }
  }
  __device__ __forceinline__ static void store(void* base, float*, long Vh, long i, const double (&s)[24]) {
    double2* b = reinterpret_cast<double2*>(base);
#pragma unroll
    for (int k = 0; k < 12; ++k) b[k * Vh + i] = make_double2(s[2 * k], s[2 * k + 1]);
  }
};

// Cache policy of a load: 0 an ordinary load, CMX_NT non-temporal (streamed through L2 with early
// eviction: for data read once per launch, so that it displaces less of the data that is re-read)
#define CMX_NT 2
template <int CP, typename T> __device__ __forceinline__ T ld_cp(const T* p) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "ld_cp: 4, 8 or 16 bytes");
  if constexpr (CP == 0) {
    return *p;
  } else {
    typedef unsigned int W __attribute__((ext_vector_type(sizeof(T) / 4)));
    const W w = __builtin_nontemporal_load(reinterpret_cast<const W*>(p));
    T r;
    __builtin_memcpy(&r, &w, sizeof(T));
    return r;
  }
}

template <> struct SpinorIO<P32> {
  using R = float;
  template <int CP = 0>
  __device__ __forceinline__ static void load(const void* base, const float*, long Vh, long i, float (&s)[24]) {
    const float4* b = reinterpret_cast<const float4*>(base);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      float4 v = ld_cp<CP>(&b[k * Vh + i]);
      s[4 * k] = v.x;
      s[4 * k + 1] = v.y;
      s[4 * k + 2] = v.z;
      s[4 * k + 3] = v.w;
    }
  }
  __device__ __forceinline__ static void store(void* base, float*, long Vh, long i, const float (&s)[24]) {
    float4* b = reinterpret_cast<float4*>(base);
#pragma unroll
    for (int k = 0; k < 6; ++k) b[k * Vh + i] = make_float4(s[4 * k], s[4 * k + 1], s[4 * k + 2], s[4 * k + 3]);
  }
};

union H8 {
  uint4 u;
  __half2 h[4];
};

template <> struct SpinorIO<P16> {
  using R = float;
  template <int CP = 0>
  __device__ __forceinline__ static void load(const void* base, const float* nrm, long Vh, long i, float (&s)[24]) {
    const uint4* b = reinterpret_cast<const uint4*>(base);
    const float n = ld_cp<CP>(&nrm[i]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      H8 v;
      v.u = ld_cp<CP>(&b[k * Vh + i]);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        float2 f = __half22float2(v.h[j]);
        s[8 * k + 2 * j] = f.x * n;
        s[8 * k + 2 * j + 1] = f.y * n;
      }
    }
  }
  __device__ __forceinline__ static void store(void* base, float* nrm, long Vh, long i, const float (&s)[24]) {
    uint4* b = reinterpret_cast<uint4*>(base);
    float m = 0.f;
#pragma unroll
    for (int k = 0; k < 24; ++k) m = fmaxf(m, fabsf(s[k]));
    // a site whose largest component is below the FP32 normal range is stored as zero: 1/m would
    // overflow to infinity (the front of an exponentially decaying propagator reaches such values on
    // large lattices, and inf * 0 would turn the site into NaN)
    const bool keep = m >= __FLT_MIN__;
    const float inv = keep ? 1.0f / m : 0.f;
    nrm[i] = keep ? m : 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      H8 v;
#pragma unroll
      for (int j = 0; j < 4; ++j) v.h[j] = __floats2half2_rn(s[8 * k + 2 * j] * inv, s[8 * k + 2 * j + 1] * inv);
      b[k * Vh + i] = v.u;
    }
  }
};

// ------------------------------------------------------------------------------------------------
// Links.  load() returns the full 3x3 (18 reals, row-major re/im) for a link whose two stored rows
// are u0, u1; for 12-real storage u2 = det(U) * conj(u0 x u1), which holds for every U = z*W with
// W in SU(3) and |z| = 1 -- i.e. SU(3) links carrying a boundary phase.  det(U) is z^3 and is
// passed per direction (bulk value and the value on the last slice x_mu = L_mu - 1).
template <int P, int RECON> struct LinkIO;

template <typename R>
__device__ __forceinline__ void recon_row2(R (&u)[18], R dre, R dim) {
  // c = conj(u0 x u1); u2 = d * c
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int a = (k + 1) % 3, b = (k + 2) % 3;
    // (u0 x u1)_k = u0_a u1_b - u0_b u1_a
    R xr = (u[2 * a] * u[6 + 2 * b] - u[2 * a + 1] * u[6 + 2 * b + 1]) - (u[2 * b] * u[6 + 2 * a] - u[2 * b + 1] * u[6 + 2 * a + 1]);
    R xi = (u[2 * a] * u[6 + 2 * b + 1] + u[2 * a + 1] * u[6 + 2 * b]) - (u[2 * b] * u[6 + 2 * a + 1] + u[2 * b + 1] * u[6 + 2 * a]);
    // conj
    xi = -xi;
    u[12 + 2 * k] = dre * xr - dim * xi;
    u[12 + 2 * k + 1] = dre * xi + dim * xr;
  }
}

template <> struct LinkIO<P64, 18> {
  using R = double;
  template <int CP = 0>
  __device__ __forceinline__ static void load(const void* base, long Vh, long i, R (&u)[18], R, R) {
    const double2* b = reinterpret_cast<const double2*>(base);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      double2 v = ld_cp<CP>(&b[k * Vh + i]);
      u[2 * k] = v.x;
      u[2 * k + 1] = v.y;
    }
  }
};
template <> struct LinkIO<P64, 12> {
  using R = double;
  template <int CP = 0>
  __device__ __forceinline__ static void load(const void* base, long Vh, long i, R (&u)[18], R dre, R dim) {
    const double2* b = reinterpret_cast<const double2*>(base);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      double2 v = ld_cp<CP>(&b[k * Vh + i]);
      u[2 * k] = v.x;
      u[2 * k + 1] = v.y;
    }
    recon_row2<R>(u, dre, dim);
  }
};
template <> struct LinkIO<P32, 18> {
  using R = float;
  template <int CP = 0>
  __device__ __forceinline__ static void load(const void* base, long Vh, long i, R (&u)[18], R, R) {
    const float2* b = reinterpret_cast<const float2*>(base);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      float2 v = ld_cp<CP>(&b[k * Vh + i]);
      u[2 * k] = v.x;
      u[2 * k + 1] = v.y;
    }
  }
};
template <> struct LinkIO<P32, 12> {
  using R = float;
  template <int CP = 0>
  __device__ __forceinline__ static void load(const void* base, long Vh, long i, R (&u)[18], R dre, R dim) {
    const float4* b = reinterpret_cast<const float4*>(base);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float4 v = ld_cp<CP>(&b[k * Vh + i]);
      u[4 * k] = v.x;
      u[4 * k + 1] = v.y;
      u[4 * k + 2] = v.z;
      u[4 * k + 3] = v.w;
    }
    recon_row2<R>(u, dre, dim);
  }
};
union H4 {
  uint2 u;
  __half2 h[2];
};
template <> struct LinkIO<P16, 12> {
  using R = float;
  template <int CP = 0>
  __device__ __forceinline__ static void load(const void* base, long Vh, long i, R (&u)[18], R dre, R dim) {
    const uint2* b = reinterpret_cast<const uint2*>(base);
    const uint4 w = ld_cp<CP>(&reinterpret_cast<const uint4*>(base)[i]);
    H4 vv[3];
    vv[0].u = make_uint2(w.x, w.y);
    vv[1].u = make_uint2(w.z, w.w);
    vv[2].u = ld_cp<CP>(&b[2 * Vh + i]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const H4 v = vv[k];
      float2 f0 = __half22float2(v.h[0]), f1 = __half22float2(v.h[1]);
      u[4 * k] = f0.x;
      u[4 * k + 1] = f0.y;
      u[4 * k + 2] = f1.x;
      u[4 * k + 3] = f1.y;
    }
    recon_row2<R>(u, dre, dim);
  }
};

// ------------------------------------------------------------------------------------------------
// Reductions: wavefront (64 lanes) butterfly through cross-lane shuffles, then LDS across waves.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum NR values over a block of blockDim.x threads (multiple of 64, <= 1024); result valid in thread 0.
template <int NR> __device__ __forceinline__ void block_sum(double (&v)[NR]) {
  __shared__ double sh[NR][16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double w = wave_sum(v[r]);
    if (lane == 0) sh[r][wid] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      double s = 0;
      for (int w = 0; w < nw; ++w) s += sh[r][w];
      v[r] = s;
    }
  }
}

// Work-group -> lattice-tile mapping that keeps each XCD on one contiguous slab of sites.
// The hardware hands work-group b to XCD (b mod 8); remapping b -> (b mod 8)*(n/8) + b/8 gives XCD k
// the contiguous logical range [k*n/8, (k+1)*n/8), so the t +- 1 neighbours of a slab are mostly in
// that XCD's own L2.  Identity when n is not a multiple of 8.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  if ((n & 7) != 0) return b;
  const int per = n >> 3;
  return (b & 7) * per + (b >> 3);
}

// ------------------------------------------------------------------------------------------------
// Batched launches: one grid covers the work-groups of several problems (solves on different
// lattices, or several solves on one).  Work-group b runs work-group map[b].y (of map[b].z) of
// problem map[b].x, whose launch record is tab[map[b].x].
struct BatchGrid {
  const void* tab;   // device array of launch records (HopLaunch, BlasLaunch or FinLaunch)
  const int4* map;   // device array, one entry per work-group of the batched grid
  int nblocks;       // work-groups of the batched grid
};

// ------------------------------------------------------------------------------------------------
// Host-side launch interfaces (implemented in the .hip files)

// Hop kernel epilogues
enum Epi {
  EPI_NONE = 0,       // out = b * D in
  EPI_CLOV = 1,       // out = A * (b * D in)
  EPI_XPAY = 2,       // out = a * x + b * D in
  EPI_CLOV_XPAY = 3,  // out = A x + b * D in
};

struct HopLaunch {
  Geom g;
  int cb;          // output checkerboard
  int isign;       // +1 / -1
  int prec, recon;
  const void* gauge[2];  // per parity block [mu][chunk][site]
  double ph[4][2][2];    // det phases: [mu][bulk|last slice][re,im]
  const void* in;  // fermion views: chunk k of site i at base[k*stride + i]
  const float* in_nrm;
  int in_stride;
  void* out;
  float* out_nrm;
  int out_stride;
  int epi;
  const void* clov;  // clover at the output parity, precision prec, stride Vh
  const void* x;
  const float* x_nrm;
  int x_stride;
  double a, b;
  double* partials;  // |out|^2 per work-group (nullable)
  int ndot;          // with partials and an xpay epilogue: |out|^2, Re <out, x>, Im <out, x> per work-group
  const int* skip;   // device flag: return immediately when *skip != 0 (nullable)
  const int* run;    // device flag: run only when *run != 0 (nullable)
  // launch shape (chosen by the operator's tuner, see hop_config)
  int block;         // threads per work-group
  int tiled;         // 1: tiled work-group order below, 0: linear order with the XCD remap
  int tile_w;        // work-groups per (z,t) plane = Lx/2*Ly / block
  int tile_zb;       // z planes per tile
  int tile_t8;       // t slices per XCD slab = Lt/8
  int prefetch;      // 1: double-buffered raw loads across the 8 directions
  int fast;          // bit 0: 2-D grid kernel of hop_fast.h (FP16/FP32, 12-real links, real phases),
                     // else the site-linear kernel of hop_kernel.h; bit 1 (HOP_FAST_STREAM): the
                     // loads of data read once per launch (links, clover, the xpay input) carry the
                     // non-temporal hint
  float inv_lxh;     // 1 / (Lx/2), for the fast kernel's plane decomposition
  float sgn[4][2];   // real det(U) phase per direction: [bulk | last slice]
};

constexpr int HOP_FAST_STREAM = 2;
// Launch-shape choice for the hopping kernels of one lattice and precision
struct HopConfig {
  int block = 256, tiled = 0, zb = 1, prefetch = 0, fast = 1;
};
// Fills the shape fields of h from c, falling back to the linear order when the lattice does not
// admit the tiling (returns the number of work-groups)
int hop_apply_config(HopLaunch& h, const HopConfig& c);
// Number of work-groups (= reduction partials) of a hopping launch
int hop_num_blocks(const HopLaunch& h);
// bg != nullptr: the batched grid of bg, whose records all select the kernel that h selects
void launch_hop(const HopLaunch& h, hipStream_t s, const BatchGrid* bg = nullptr);
void launch_hop_p64(const HopLaunch& h, hipStream_t s, const BatchGrid* bg);
void launch_hop_p32(const HopLaunch& h, hipStream_t s, const BatchGrid* bg);
void launch_hop_p16(const HopLaunch& h, hipStream_t s, const BatchGrid* bg);

// Device solver scalars live in one double array; flags in one int array
struct DevState {
  double* S;  // scalars
  int* F;     // flags
};

}  // namespace cmx
