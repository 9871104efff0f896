// Hopping term on up to 12 right-hand sides with the link products on the matrix cores (FP16 storage,
// 12-real links, real boundary phases): the 12 spin-colour sources of a propagator (quarkProp4_a,
// quarkprop4_w.cc:70-117) share every link, so per site and direction the product
// U (3x3) x [h_1 ... h_24] (the two half spinors of 12 sources) is one small matrix product with a
// shared left operand.  v_mfma_f32_4x4x4_16b_f16 computes 16 independent 4x4x4 products per wave;
// here each block of 4 lanes is one site: lane q holds row q of U (row 3 zero), column q of the right
// operand is one half spinor (source 2p + q/2, half q%2, p = source pair), and the complex product is
// four real products (U_re h_re - U_im h_im, U_re h_im + U_im h_re) with FP32 accumulation.
//
// The spin projection (h = psi_up + s B psi_down) is formed in packed FP16 like the single-source
// kernel (hop_fast.h), the link's third row is rebuilt in FP32 and rounded to FP16 as an operand,
// and the reconstruction of the lower spins and the per-site FP16 norms stay on the vector units.
// A wave covers 16 sites, a 256-lane work-group 64 sites of one (z, t) plane.  The links of a site
// are loaded and turned into matrix operands once for all sources; the sources are taken in pairs,
// each pair accumulating its 8 directions before it is stored.
#include "hop_fast.h"
#include "hop_multi.h"

namespace cmx {

namespace {

typedef _Float16 h4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4_t mfma4(h4_t a, h4_t b, f4_t c) { return __builtin_amdgcn_mfma_f32_4x4x4f16(a, b, c, 0, 0, 0); }

// quad_perm(1, 0, 3, 2): each lane reads its neighbour within the pair of lanes of one source
__device__ __forceinline__ float swap_pair(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}

union H8m {
  uint4 u;
  hf2 h[4];
};

// row q of U (FWD) or of U^dag (backward link), q = 0..3 (row 3 zero), as (re, im) FP16 operands
template <bool FWD>
__device__ __forceinline__ void link_operand(const FastLink<P16>& lk, float sg, int q, h4_t& are, h4_t& aim, h4_t& aimn) {
  float cr[3], ci[3];
  fast_row2_h(lk, cr, ci);
  hf2 u[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) u[k] = as_hf2(lk.c[k]);
  hf2 r2[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) r2[k] = hf2{(_Float16)(sg * cr[k]), (_Float16)(sg * ci[k])};
  const hf2 zero = {(_Float16)0.0f, (_Float16)0.0f};
  hf2 a[3];
  if constexpr (FWD) {  // A[q][k] = U[q][k]
#pragma unroll
    for (int k = 0; k < 3; ++k) a[k] = q == 0 ? u[k] : q == 1 ? u[3 + k] : q == 2 ? r2[k] : zero;
  } else {  // A[q][k] = conj(U[k][q])
    const hf2 c0 = q == 0 ? u[0] : q == 1 ? u[1] : q == 2 ? u[2] : zero;
    const hf2 c1 = q == 0 ? u[3] : q == 1 ? u[4] : q == 2 ? u[5] : zero;
    const hf2 c2 = q == 0 ? r2[0] : q == 1 ? r2[1] : q == 2 ? r2[2] : zero;
    a[0] = hf2{c0.x, -c0.y};
    a[1] = hf2{c1.x, -c1.y};
    a[2] = hf2{c2.x, -c2.y};
  }
  are = h4_t{a[0].x, a[1].x, a[2].x, (_Float16)0.0f};
  aim = h4_t{a[0].y, a[1].y, a[2].y, (_Float16)0.0f};
  aimn = -aim;
}

// half spinor `hs` (0/1) of direction MU, sign S of (1 + S gamma_mu), from the FP16 spinor p[12]
template <int MU, int S>
__device__ __forceinline__ void project_half(const hf2 (&p)[12], int hs, hf2 (&h)[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const hf2 p0 = p[c], p1 = p[3 + c], p2 = p[6 + c], p3 = p[9 + c];
    hf2 h0, h1;
    if constexpr (MU == 0) {
      h0 = proj_addi<S>(p0, p3);
      h1 = proj_addi<S>(p1, p2);
    } else if constexpr (MU == 1) {
      h0 = proj_add<-S>(p0, p3);
      h1 = proj_add<S>(p1, p2);
    } else if constexpr (MU == 2) {
      h0 = proj_addi<S>(p0, p2);
      h1 = proj_addi<-S>(p1, p3);
    } else {
      h0 = proj_add<S>(p0, p2);
      h1 = proj_add<S>(p1, p3);
    }
    h[c] = hs ? h1 : h0;
  }
}

template <int ISIGN, int LCP>
__global__ void __launch_bounds__(256) k_hop_mfma(MultiHopLaunch M) {
  const HopLaunch& h = M.h;
  const int Lxh = h.g.Lxh, Ly = h.g.L[1], Lz = h.g.L[2], Lt = h.g.L[3];
  const int Pl = Lxh * Ly;
  const int gx = (Pl + 63) / 64;  // 64 sites per 256-lane work-group
  int bx, z, t;
  {
    const int b = blockIdx.x;
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
  const int lane = (int)threadIdx.x & 63;
  const int q = lane & 3;   // row of U / column of the half-spinor operand within the site's quad
  const int hs = q & 1;     // half spinor of this lane's column
  const int sp = q >> 1;    // source of this lane within the pair
  const int r_raw = bx * 64 + ((int)threadIdx.x >> 2);
  const bool valid = r_raw < Pl;
  const int r = valid ? r_raw : Pl - 1;  // the tail's spare quads compute a real site and store nothing
  const int plane = t * Lz + z;
  int y = (int)((float)r * h.inv_lxh);
  int xh = r - y * Lxh;
  if (xh < 0) { --y; xh += Lxh; }
  if (xh >= Lxh) { ++y; xh -= Lxh; }
  const int x = 2 * xh + ((h.cb + y + z + t) & 1);
  const int base = plane * Pl;
  const int i = base + r;
  const int Vh = h.g.Vh;
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

  // the link operands of the 8 directions, once for all sources
  h4_t are[8], aim[8], aimn[8];
#define CMX_MFMA_LINK(STEP)                                                                                \
  {                                                                                                        \
    constexpr int MU_ = (STEP) >> 1;                                                                       \
    constexpr bool FWD_ = ((STEP) & 1) == 0;                                                               \
    const __amdgpu_buffer_rsrc_t rl = make_rsrc(                                                           \
        static_cast<const char*>(h.gauge[FWD_ ? h.cb : 1 - h.cb]) + (size_t)MU_ * 24 * Vh, (unsigned)(24 * Vh)); \
    FastLink<P16> lk;                                                                                      \
    fast_load_link<P16, LCP>(rl, Vh, FWD_ ? i : nb[STEP], lk);                                             \
    const float sg_ = (FWD_ ? lastf[MU_] : lastb[MU_]) ? h.sgn[MU_][1] : h.sgn[MU_][0];                   \
    link_operand<FWD_>(lk, sg_, q, are[STEP], aim[STEP], aimn[STEP]);                                      \
  }
  CMX_MFMA_LINK(0)
  CMX_MFMA_LINK(1)
  CMX_MFMA_LINK(2)
  CMX_MFMA_LINK(3)
  CMX_MFMA_LINK(4)
  CMX_MFMA_LINK(5)
  CMX_MFMA_LINK(6)
  CMX_MFMA_LINK(7)
#undef CMX_MFMA_LINK
  const float sgn_h = hs ? -1.0f : 1.0f;
  const float b = (float)h.b;
  const f4_t fzero = {0.f, 0.f, 0.f, 0.f};
  const int npair = (M.nrhs + 1) >> 1;
  for (int pr = 0; pr < npair; ++pr) {
    const int s_raw = 2 * pr + sp;
    const bool s_ok = s_raw < M.nrhs;
    const int s = s_ok ? s_raw : 2 * pr;  // an odd last source: the spare lanes repeat the pair's first
    const VRef vin = M.in[s];
    const uint4* vb = static_cast<const uint4*>(vin.v);
    const size_t st = (size_t)vin.stride;
    // up: upper spin hs; lz: lower spin 2 + hs (z, t directions); lx: lower spin 3 - hs (x, y)
    float up[6], lz[6], lx[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) up[k] = lz[k] = lx[k] = 0.f;
#define CMX_MFMA_STEP(STEP)                                                                        \
    {                                                                                              \
      constexpr int MU_ = (STEP) >> 1;                                                             \
      constexpr bool FWD_ = ((STEP) & 1) == 0;                                                     \
      constexpr int S_ = FWD_ ? -ISIGN : ISIGN;                                                    \
      const int j_ = nb[STEP];                                                                     \
      hf2 p_[12];                                                                                  \
      _Pragma("unroll") for (int k = 0; k < 3; ++k) {                                              \
        H8m v_;                                                                                    \
        v_.u = vb[k * st + j_];                                                                    \
        _Pragma("unroll") for (int w = 0; w < 4; ++w) p_[4 * k + w] = v_.h[w];                     \
      }                                                                                            \
      const float n_ = vin.nrm[j_];                                                                \
      hf2 hh_[3];                                                                                  \
      project_half<MU_, S_>(p_, hs, hh_);                                                          \
      const h4_t bre_ = {hh_[0].x, hh_[1].x, hh_[2].x, (_Float16)0.0f};                           \
      const h4_t bim_ = {hh_[0].y, hh_[1].y, hh_[2].y, (_Float16)0.0f};                            \
      f4_t dre_ = mfma4(are[STEP], bre_, fzero);                                                   \
      dre_ = mfma4(aimn[STEP], bim_, dre_);                                                        \
      f4_t dim_ = mfma4(are[STEP], bim_, fzero);                                                   \
      dim_ = mfma4(aim[STEP], bre_, dim_);                                                         \
      _Pragma("unroll") for (int c = 0; c < 3; ++c) {                                              \
        up[2 * c] = __fmaf_rn(n_, dre_[c], up[2 * c]);                                             \
        up[2 * c + 1] = __fmaf_rn(n_, dim_[c], up[2 * c + 1]);                                     \
        const float ns_ = n_ * (float)S_;                                                          \
        if constexpr (MU_ == 0) {                                                                  \
          lx[2 * c] = __fmaf_rn(ns_, dim_[c], lx[2 * c]);                                          \
          lx[2 * c + 1] = __fmaf_rn(-ns_, dre_[c], lx[2 * c + 1]);                                 \
        } else if constexpr (MU_ == 1) {                                                           \
          lx[2 * c] = __fmaf_rn(-sgn_h * ns_, dre_[c], lx[2 * c]);                                 \
          lx[2 * c + 1] = __fmaf_rn(-sgn_h * ns_, dim_[c], lx[2 * c + 1]);                         \
        } else if constexpr (MU_ == 2) {                                                           \
          lz[2 * c] = __fmaf_rn(sgn_h * ns_, dim_[c], lz[2 * c]);                                  \
          lz[2 * c + 1] = __fmaf_rn(-sgn_h * ns_, dre_[c], lz[2 * c + 1]);                         \
        } else {                                                                                   \
          lz[2 * c] = __fmaf_rn(ns_, dre_[c], lz[2 * c]);                                          \
          lz[2 * c + 1] = __fmaf_rn(ns_, dim_[c], lz[2 * c + 1]);                                  \
        }                                                                                          \
      }                                                                                            \
    }
    CMX_MFMA_STEP(0)
    CMX_MFMA_STEP(1)
    CMX_MFMA_STEP(2)
    CMX_MFMA_STEP(3)
    CMX_MFMA_STEP(4)
    CMX_MFMA_STEP(5)
    CMX_MFMA_STEP(6)
    CMX_MFMA_STEP(7)
#undef CMX_MFMA_STEP
    // lower spin 2 + hs = own z/t part + the partner lane's x/y part (its lower spin 3 - (1 - hs))
    float o_up[6], o_lo[6];
    float m = 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      o_up[k] = b * up[k];
      o_lo[k] = b * (lz[k] + swap_pair(lx[k]));
      m = fmaxf(m, fmaxf(fabsf(o_up[k]), fabsf(o_lo[k])));
    }
    m = fmaxf(m, swap_pair(m));  // the largest component over the 24 reals of the site
    if (valid && s_ok) {
      const VRef vo = M.out[s];
      const bool keep = m >= __FLT_MIN__;
      const float inv = keep ? 1.0f / m : 0.f;
      if (hs == 0) vo.nrm[i] = keep ? m : 0.f;
      // complex k of the site lives in chunk k / 4 at position k % 4; this lane holds complex
      // 3 hs .. 3 hs + 2 (spin hs) and 6 + 3 hs .. 8 + 3 hs (spin 2 + hs)
      __half2 cu[3], cl[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        cu[c] = __floats2half2_rn(o_up[2 * c] * inv, o_up[2 * c + 1] * inv);
        cl[c] = __floats2half2_rn(o_lo[2 * c] * inv, o_lo[2 * c + 1] * inv);
      }
      __half2* ob = static_cast<__half2*>(vo.v);
      const size_t so = (size_t)vo.stride * 4;  // half2 per chunk row
      auto at = [&](int k) -> __half2& { return ob[(size_t)(k >> 2) * so + (size_t)i * 4 + (k & 3)]; };
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        at(3 * hs + c) = cu[c];
        at(6 + 3 * hs + c) = cl[c];
      }
    }
  }
}

}  // namespace

void launch_hop_mfma(const MultiHopLaunch& M, hipStream_t s) {
  const HopLaunch& h = M.h;
  CMX_REQUIRE(h.prec == P16 && h.recon == 12 && (h.fast & 1), CMX_ERR_UNSUPPORTED,
              "multi-source hop (matrix cores): FP16 storage with 12-real links and real boundary phases only");
  CMX_REQUIRE(M.nrhs >= 1 && M.nrhs <= MULTI_MAXR, CMX_ERR_ARG, "multi-source hop (matrix cores): 1..12 sources");
  CMX_REQUIRE(h.epi == EPI_NONE && !h.partials, CMX_ERR_ARG, "multi-source hop (matrix cores): plain hopping term only");
  CMX_REQUIRE(!h.tiled || (h.g.L[3] % 8 == 0 && h.g.L[2] % h.tile_zb == 0), CMX_ERR_ARG,
              "multi-source hop (matrix cores): tiled order needs Lt % 8 == 0 and Lz % zb == 0");
  const int Pl = h.g.Lxh * h.g.L[1];
  const dim3 grid(((Pl + 63) / 64) * h.g.L[2] * h.g.L[3]);
  const bool lnt = (h.fast & HOP_FAST_STREAM) != 0;
  if (h.isign == 1) {
    if (lnt)
      hipLaunchKernelGGL((k_hop_mfma<1, CMX_NT>), grid, dim3(256), 0, s, M);
    else
      hipLaunchKernelGGL((k_hop_mfma<1, 0>), grid, dim3(256), 0, s, M);
  } else {
    if (lnt)
      hipLaunchKernelGGL((k_hop_mfma<-1, CMX_NT>), grid, dim3(256), 0, s, M);
    else
      hipLaunchKernelGGL((k_hop_mfma<-1, 0>), grid, dim3(256), 0, s, M);
  }
  CMX_HIP(hipGetLastError());
}

}  // namespace cmx
