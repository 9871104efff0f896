// Hopping term on several right-hand sides at once (FP16 storage, 12-real links with real boundary
// phases): the 12 spin-colour sources of a propagator (quarkProp4_a, quarkprop4_w.cc:70-117) share the
// gauge field, so each lane loads a link and rebuilds its third row once per group of NB sources and
// applies it to all of them.  Per site and source the link bytes fall from 8 x 24 B to 8 x 24 B / NB.
// The per-source arithmetic is that of the single-source plane kernel (hop_fast.h), so the result is
// identical to NB single-source applications.
#include "hop_fast.h"
#include "hop_multi.h"

namespace cmx {

namespace {

template <int NB, int ISIGN, int LCP>
__global__ void __launch_bounds__(256) k_hop_multi(MultiHopLaunch M, int r0) {
  const HopLaunch& h = M.h;
  const int Lxh = h.g.Lxh, Ly = h.g.L[1], Lz = h.g.L[2], Lt = h.g.L[3];
  const int Pl = Lxh * Ly;
  const int gx = (Pl + (int)blockDim.x - 1) / (int)blockDim.x;
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
  const int plane = t * Lz + z;
  const int r = bx * blockDim.x + threadIdx.x;
  if (r >= Pl) return;
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
  __amdgpu_buffer_rsrc_t rin[NB], rnin[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    rin[k] = make_rsrc(M.in[r0 + k].v, (unsigned)(48 * M.in[r0 + k].stride));
    rnin[k] = make_rsrc(M.in[r0 + k].nrm, (unsigned)(4 * M.in[r0 + k].stride));
  }
  float acc[NB][24];
#pragma unroll
  for (int k = 0; k < NB; ++k)
#pragma unroll
    for (int c = 0; c < 24; ++c) acc[k][c] = 0.f;
  // raw loads of step STEP into buffer slot (STEP & 1); the next step's loads are issued before the
  // current step is computed, so two steps of raw data are live at a time
  FastLink<P16> lk[2];
  FastSpinor<P16> sp[2][NB];
  auto load_step = [&](int step, int slot) {
    const int mu = step >> 1;
    const bool fwd = (step & 1) == 0;
    const __amdgpu_buffer_rsrc_t rl = make_rsrc(
        static_cast<const char*>(h.gauge[fwd ? h.cb : 1 - h.cb]) + (size_t)mu * 24 * Vh, (unsigned)(24 * Vh));
    fast_load_link<P16, LCP>(rl, Vh, fwd ? i : nb[step], lk[slot]);
#pragma unroll
    for (int k = 0; k < NB; ++k) fast_load_spinor<P16>(rin[k], rnin[k], M.in[r0 + k].stride, nb[step], sp[slot][k]);
  };
#define CMX_MULTI_STEP(STEP)                                                                          \
  {                                                                                                   \
    if ((STEP) < 7) load_step((STEP) + 1, ((STEP) + 1) & 1);                                          \
    constexpr int MU_ = (STEP) >> 1;                                                                  \
    constexpr bool FWD_ = ((STEP) & 1) == 0;                                                          \
    float cr[3], ci[3];                                                                               \
    fast_row2_h(lk[(STEP) & 1], cr, ci);                                                              \
    const float sg_ = (FWD_ ? lastf[MU_] : lastb[MU_]) ? h.sgn[MU_][1] : h.sgn[MU_][0];               \
    _Pragma("unroll") for (int k = 0; k < NB; ++k) {                                                  \
      fast_step_h_body<MU_, FWD_, ISIGN>(sp[(STEP) & 1][k], lk[(STEP) & 1], cr, ci, sg_, acc[k]);     \
      pin_acc(acc[k]);                                                                                \
    }                                                                                                 \
  }
  load_step(0, 0);
  CMX_MULTI_STEP(0)
  CMX_MULTI_STEP(1)
  CMX_MULTI_STEP(2)
  CMX_MULTI_STEP(3)
  CMX_MULTI_STEP(4)
  CMX_MULTI_STEP(5)
  CMX_MULTI_STEP(6)
  CMX_MULTI_STEP(7)
#undef CMX_MULTI_STEP
  const float b = (float)h.b;
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    float out[24];
#pragma unroll
    for (int c = 0; c < 24; ++c) out[c] = b * acc[k][c];
    SpinorIO<P16>::store(M.out[r0 + k].v, M.out[r0 + k].nrm, M.out[r0 + k].stride, i, out);
  }
}

template <int NB> void launch_nb(const MultiHopLaunch& M, int r0, hipStream_t s) {
  const HopLaunch& h = M.h;
  const int Pl = h.g.Lxh * h.g.L[1];
  const dim3 grid(((Pl + h.block - 1) / h.block) * h.g.L[2] * h.g.L[3]);
  const bool lnt = (h.fast & HOP_FAST_STREAM) != 0;
  if (h.isign == 1) {
    if (lnt)
      hipLaunchKernelGGL((k_hop_multi<NB, 1, CMX_NT>), grid, dim3(h.block), 0, s, M, r0);
    else
      hipLaunchKernelGGL((k_hop_multi<NB, 1, 0>), grid, dim3(h.block), 0, s, M, r0);
  } else {
    if (lnt)
      hipLaunchKernelGGL((k_hop_multi<NB, -1, CMX_NT>), grid, dim3(h.block), 0, s, M, r0);
    else
      hipLaunchKernelGGL((k_hop_multi<NB, -1, 0>), grid, dim3(h.block), 0, s, M, r0);
  }
}

}  // namespace

void launch_hop_multi(const MultiHopLaunch& M, hipStream_t s) {
  const HopLaunch& h = M.h;
  CMX_REQUIRE(h.prec == P16 && h.recon == 12 && (h.fast & 1), CMX_ERR_UNSUPPORTED,
              "multi-source hop: FP16 storage with 12-real links and real boundary phases only");
  CMX_REQUIRE(M.nrhs >= 1 && M.nrhs <= MULTI_MAXR, CMX_ERR_ARG, "multi-source hop: 1..12 sources");
  CMX_REQUIRE(h.epi == EPI_NONE && !h.partials, CMX_ERR_ARG, "multi-source hop: plain hopping term only");
  CMX_REQUIRE(M.group >= 1 && M.group <= 4, CMX_ERR_ARG, "multi-source hop: group of 1..4 sources");
  int r0 = 0;
  while (r0 < M.nrhs) {
    const int left = M.nrhs - r0;
    const int nb = left >= M.group ? M.group : left;
    switch (nb) {
      case 4: launch_nb<4>(M, r0, s); break;
      case 3: launch_nb<3>(M, r0, s); break;
      case 2: launch_nb<2>(M, r0, s); break;
      default: launch_nb<1>(M, r0, s); break;
    }
    r0 += nb;
  }
}

}  // namespace cmx
