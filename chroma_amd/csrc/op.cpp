// Operator and gauge handles: device storage, precision copies and the fused even-odd operator.
#include "op.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

#include "batch.h"
#include "hop_multi.h"
#include "parallel.h"
#include "setup.h"

namespace cmx {

void launch_hop(const HopLaunch& h, hipStream_t s, const BatchGrid* bg) {
  if (!bg)
    if (Recorder* r = active_recorder()) {
      record_hop(*r, h);
      return;
    }
  switch (h.prec) {
    case P64: launch_hop_p64(h, s, bg); break;
    case P32: launch_hop_p32(h, s, bg); break;
    case P16: launch_hop_p16(h, s, bg); break;
    default: throw Error(CMX_ERR_ARG, "hop: bad precision");
  }
}
int hop_num_blocks(const HopLaunch& h) {
  if (h.fast & 1) return ((h.g.Lxh * h.g.L[1] + h.block - 1) / h.block) * h.g.L[2] * h.g.L[3];
  return (h.g.Vh + h.block - 1) / h.block;
}

int hop_apply_config(HopLaunch& h, const HopConfig& c) {
  h.block = c.block;
  h.prefetch = c.prefetch;
  h.tiled = 0;
  // the fast kernel needs 12-real links (or FP32/12), real boundary phases and plane indices < 2^16
  bool real_phases = true;
  for (int mu = 0; mu < 4; ++mu)
    for (int k = 0; k < 2; ++k) {
      if (std::fabs(h.ph[mu][k][1]) > 1e-12 || std::fabs(std::fabs(h.ph[mu][k][0]) - 1.0) > 1e-6) real_phases = false;
      h.sgn[mu][k] = (float)(h.ph[mu][k][0] < 0 ? -1.0 : 1.0);
    }
  // buffer offsets of the plane kernel are 32-bit: a whole field (spinor chunks or one direction's
  // links) stays below 2 GiB
  const size_t site_bytes = h.prec == P16 ? 48 : (h.prec == P32 ? 96 : 192);
  const bool plane = (c.fast & 1) && real_phases && h.recon == 12 && h.g.Lxh * h.g.L[1] < 65536 &&
                     h.g.L[2] * h.g.L[3] < 65536 && site_bytes * (size_t)h.g.Vh < ((size_t)1 << 31);
  h.fast = plane ? c.fast : (c.fast & HOP_FAST_STREAM);  // the stream bit applies to both kernels
  h.inv_lxh = 1.0f / (float)h.g.Lxh;
  if (h.fast & 1) {  // the plane-grid kernel tiles whole planes: any block size, Lt % 8 and Lz % zb only
    if (c.tiled && h.g.L[3] % 8 == 0 && c.zb > 0 && h.g.L[2] % c.zb == 0) {
      h.tiled = 1;
      h.tile_zb = c.zb;
      h.tile_t8 = h.g.L[3] / 8;
    }
    return hop_num_blocks(h);
  }
  const int P = h.g.Lxh * h.g.L[1];
  if (c.tiled && P % c.block == 0 && h.g.L[3] % 8 == 0 && c.zb > 0 && h.g.L[2] % c.zb == 0) {
    h.tiled = 1;
    h.tile_w = P / c.block;
    h.tile_zb = c.zb;
    h.tile_t8 = h.g.L[3] / 8;
  }
  return hop_num_blocks(h);
}

namespace {

using cd = std::complex<double>;

cd det3(const double* u) {
  auto e = [&](int i, int j) { return cd(u[2 * (3 * i + j)], u[2 * (3 * i + j) + 1]); };
  return e(0, 0) * (e(1, 1) * e(2, 2) - e(1, 2) * e(2, 1)) - e(0, 1) * (e(1, 0) * e(2, 2) - e(1, 2) * e(2, 0)) +
         e(0, 2) * (e(1, 0) * e(2, 1) - e(1, 1) * e(2, 0));
}

// deviation of U from (det phase) x SU(3): max(|U U^dag - 1|, |det U - expected|)
double link_dev(const double* u, cd expect) {
  double m = std::abs(det3(u) - expect);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cd acc = 0;
      for (int k = 0; k < 3; ++k)
        acc += cd(u[2 * (3 * i + k)], u[2 * (3 * i + k) + 1]) * std::conj(cd(u[2 * (3 * j + k)], u[2 * (3 * j + k) + 1]));
      m = std::max(m, std::abs(acc - (i == j ? 1.0 : 0.0)));
    }
  return m;
}

long full_index(const int L[4], const int x[4], long Vh) {
  const int cb = (x[0] + x[1] + x[2] + x[3]) & 1;
  const long lex = (((long)x[3] * L[2] + x[2]) * L[1] + x[1]) * L[0] + x[0];
  return cb * Vh + lex / 2;
}

}  // namespace

std::shared_ptr<GaugeData> make_gauge(const int dims[4], const double* const links[4]) {
  CMX_REQUIRE(dims && links && links[0] && links[1] && links[2] && links[3], CMX_ERR_ARG,
              "gauge_create: null argument");
  auto gd = std::make_shared<GaugeData>();
  long V = 1;
  for (int m = 0; m < 4; ++m) {
    CMX_REQUIRE(dims[m] > 0 && dims[m] % 2 == 0, CMX_ERR_ARG, "gauge_create: lattice extents must be positive and even");
    gd->g.L[m] = dims[m];
    V *= dims[m];
  }
  CMX_REQUIRE(V / 2 < (1L << 30), CMX_ERR_ARG, "gauge_create: lattice too large for 32-bit site indices");
  gd->V = V;
  gd->Vh = V / 2;
  gd->g.Lxh = dims[0] / 2;
  gd->g.Vh = (int)gd->Vh;
  // one copy per direction: the caller's arrays need not be contiguous (QUDA_QDP_GAUGE_ORDER)
  const size_t nd = (size_t)4 * V * 18;
  DevBuf tmp(nd * sizeof(double));
  for (int mu = 0; mu < 4; ++mu)
    CMX_HIP(hipMemcpy(tmp.as<double>() + (size_t)mu * V * 18, links[mu], (size_t)V * 18 * sizeof(double),
                      hipMemcpyHostToDevice));
  for (int p = 0; p < 2; ++p) gd->g64[p].alloc((size_t)4 * 9 * 16 * gd->Vh);
  gauge_pack64(gd->g, tmp.as<double>(), gd->g64[0].as<double2>(), gd->g64[1].as<double2>(), 0);
  CMX_HIP(hipGetLastError());
  CMX_HIP(hipDeviceSynchronize());
  // det(U) per direction on the bulk and on the last slice, from two sample links; all links are
  // then checked against that pattern on the host (the 12-real reconstruction relies on it)
  for (int mu = 0; mu < 4; ++mu) {
    int x0[4] = {0, 0, 0, 0}, x1[4] = {0, 0, 0, 0};
    x1[mu] = dims[mu] - 1;
    const cd d0 = det3(links[mu] + (size_t)full_index(dims, x0, gd->Vh) * 18);
    const cd d1 = det3(links[mu] + (size_t)full_index(dims, x1, gd->Vh) * 18);
    gd->ph[mu][0][0] = d0.real(), gd->ph[mu][0][1] = d0.imag();
    gd->ph[mu][1][0] = d1.real(), gd->ph[mu][1][1] = d1.imag();
  }
  const double dev = parallel_max(4 * V, [&](long t) {
    const int mu = (int)(t / V);
    const long s = t % V;
    const int cb = s >= gd->Vh ? 1 : 0;
    long r = s - cb * gd->Vh;
    const int Lxh = dims[0] / 2;
    int x[4];
    const int xh = (int)(r % Lxh);
    r /= Lxh;
    x[1] = (int)(r % dims[1]);
    r /= dims[1];
    x[2] = (int)(r % dims[2]);
    x[3] = (int)(r / dims[2]);
    x[0] = 2 * xh + ((cb + x[1] + x[2] + x[3]) & 1);
    const int last = x[mu] == dims[mu] - 1;
    const cd e(gd->ph[mu][last][0], gd->ph[mu][last][1]);
    return link_dev(links[mu] + (size_t)s * 18, e);
  });
  gd->su3_dev = dev;
  return gd;
}

Op::~Op() {
  for (auto& e : graphs) (void)hipGraphExecDestroy(e.exec);
  if (stream && !shared_stream) (void)hipStreamDestroy(stream);
  if (hS) (void)hipHostFree(hS);
  if (hF) (void)hipHostFree(hF);
}

Op* make_op(std::shared_ptr<GaugeData> gd, const cmx_op_params& p) {
  CMX_REQUIRE(gd, CMX_ERR_ARG, "op_create: null gauge");
  CMX_REQUIRE(p.kind >= CMX_OP_UNPREC_WILSON && p.kind <= CMX_OP_EO_CLOVER, CMX_ERR_ARG, "op_create: unknown kind");
  CMX_REQUIRE(p.sloppy_prec == CMX_SINGLE || p.sloppy_prec == CMX_HALF, CMX_ERR_ARG,
              "op_create: sloppy precision must be CMX_SINGLE or CMX_HALF");
  CMX_REQUIRE(p.sloppy_recon == 12 || p.sloppy_recon == 18, CMX_ERR_ARG, "op_create: reconstruct must be 12 or 18");
  CMX_REQUIRE(!(p.sloppy_prec == CMX_HALF && p.sloppy_recon == 18), CMX_ERR_UNSUPPORTED,
              "op_create: FP16 links use 12-real reconstruction");
  std::unique_ptr<Op> op(new Op());
  op->gd = gd;
  op->par = p;
  op->g = gd->g;
  op->V = gd->V;
  op->Vh = gd->Vh;
  op->fact = 4.0 + p.mass;
  op->recon[P32] = p.sloppy_prec == CMX_SINGLE ? p.sloppy_recon : 12;
  op->recon[P16] = 12;
  // FP64 links as 12 reals (third row rebuilt in FP64 on use) when they are SU(3) up to a boundary
  // phase to rounding, which makes the rebuilt row as exact as a stored one
  op->recon[P64] = gd->su3_dev <= 1e-14 ? 12 : 18;
  if (gd->su3_dev > 1e-6)
    CMX_REQUIRE(p.sloppy_recon == 18 && p.sloppy_prec == CMX_SINGLE, CMX_ERR_ARG,
                "op_create: links are not SU(3) up to a per-direction boundary phase (deviation " +
                    std::to_string(gd->su3_dev) + "); use CMX_SINGLE with CMX_RECONS_18");
  CMX_HIP(hipStreamCreate(&op->stream));
  op->npartials = std::max(2 * (int)((op->Vh + 63) / 64), blas_blocks((int)op->V)) * 4;
  op->partials.alloc(sizeof(double) * op->npartials);
  op->state_S.alloc(sizeof(double) * S_NSLOTS);
  op->state_F.alloc(sizeof(int) * F_NFLAGS);
  CMX_HIP(hipHostMalloc((void**)&op->hS, sizeof(double) * S_NSLOTS, 0));
  CMX_HIP(hipHostMalloc((void**)&op->hF, sizeof(int) * F_NFLAGS, 0));
  op->staging.alloc(sizeof(double) * 24 * op->V);
  if (op->clover()) {
    // QDPCloverTermT::create (clover_term_qdp_w.h:244-294): coefficients x 1/2; planes
    // (0,1),(0,2),(0,3),(1,2),(1,3),(2,3) with t_dir = 3 (getCloverCoeff, :1525-1547)
    const double cR = 0.5 * p.clov_coeff_r, cT = 0.5 * p.clov_coeff_t;
    const double coef[6] = {cR, cR, cT, cR, cT, cT};
    const double diag_mass = 1.0 + 3.0 + p.mass;
    DevBuf F((size_t)6 * 9 * 16 * op->V);
    DevBuf c1((size_t)36 * 16 * op->Vh);
    op->clov64_even.alloc((size_t)36 * 16 * op->Vh);
    clover_build(op->g, gd->g64[0].as<double2>(), gd->g64[1].as<double2>(), F.as<double2>(), coef, diag_mass,
                 op->clov64_even.as<double2>(), c1.as<double2>(), op->stream);
    op->clov_odd[P64] = std::move(c1);
    op->clovinv_even[P64].alloc((size_t)36 * 16 * op->Vh);
    DevBuf ld(sizeof(double) * op->Vh), nn(sizeof(double) * op->Vh);
    clover_invert(op->g, op->clov64_even.as<double2>(), op->clovinv_even[P64].as<double2>(), ld.as<double>(),
                  nn.as<double>(), op->stream);
    // deterministic sums of the per-site log|det| and negative-pivot indicators
    int nb = sum_partials((int)op->Vh, ld.as<double>(), op->partials.as<double>(), op->stream);
    launch_finalize(FinLaunch{op->partials.as<double>(), nb, 1, op->state_S.as<double>(), op->state_F.as<int>(),
                              PG_STORE, 0},
                    op->stream);
    nb = sum_partials((int)op->Vh, nn.as<double>(), op->partials.as<double>(), op->stream);
    launch_finalize(FinLaunch{op->partials.as<double>(), nb, 1, op->state_S.as<double>(), op->state_F.as<int>(),
                              PG_STORE, 1},
                    op->stream);
    CMX_HIP(hipGetLastError());
    CMX_HIP(hipMemcpyAsync(op->hS, op->state_S.p, 2 * sizeof(double), hipMemcpyDeviceToHost, op->stream));
    CMX_HIP(hipStreamSynchronize(op->stream));
    op->logdet_even = op->hS[0];
    op->nneg_even = (long)std::llround(op->hS[1]);
  }
  op->ensure_prec(P64);
  op->ensure_prec(p.sloppy_prec);
  if (p.autotune_dslash) op->autotune();
  CMX_HIP(hipStreamSynchronize(op->stream));
  return op.release();
}

Op* clone_op(Op& o, int prec) {
  o.ensure_prec(prec);
  std::unique_ptr<Op> c(new Op());
  c->gd = o.gd;
  c->par = o.par;
  c->g = o.g;
  c->V = o.V;
  c->Vh = o.Vh;
  c->fact = o.fact;
  std::memcpy(c->recon, o.recon, sizeof(o.recon));
  std::memcpy(c->hcfg, o.hcfg, sizeof(o.hcfg));
  c->multi_group = o.multi_group;
  for (int p_ = 0; p_ < 3; ++p_) {
    for (int q = 0; q < 2; ++q) c->gauge[p_][q].alias(o.gauge[p_][q]);
    c->clov_odd[p_].alias(o.clov_odd[p_]);
    c->clovinv_even[p_].alias(o.clovinv_even[p_]);
  }
  c->clov64_even.alias(o.clov64_even);
  c->logdet_even = o.logdet_even;
  c->nneg_even = o.nneg_even;
  c->stream = o.stream;
  c->shared_stream = true;
  c->npartials = o.npartials;
  c->partials.alloc(sizeof(double) * c->npartials);
  c->state_S.alloc(sizeof(double) * S_NSLOTS);
  c->state_F.alloc(sizeof(int) * F_NFLAGS);
  CMX_HIP(hipHostMalloc((void**)&c->hS, sizeof(double) * S_NSLOTS, 0));
  CMX_HIP(hipHostMalloc((void**)&c->hF, sizeof(int) * F_NFLAGS, 0));
  c->staging.alias(o.staging);
  c->mg = o.mg;  // the multigrid state is read-only in a solve, but for its coarse work space (one stream)
  return c.release();
}

void Op::ensure_prec(int prec) {
  if (prec == P64 && recon[P64] == 18) return;  // the gauge field's own FP64 links
  if (gauge[prec][0].p) return;
  const int rc = recon[prec];
  const size_t gb = (size_t)4 * link_chunks(prec, rc) * link_chunk_bytes(prec, rc) * Vh;
  for (int par_ = 0; par_ < 2; ++par_) {
    gauge[prec][par_].alloc(gb);
    gauge_pack_sloppy(g, prec, rc, gd->g64[par_].as<double2>(), gauge[prec][par_].p, stream);
  }
  if (clover() && prec != P64) {
    clov_odd[prec].alloc(clover_bytes(prec, Vh));
    clovinv_even[prec].alloc(clover_bytes(prec, Vh));
    clover_convert((int)Vh, prec, clov_odd[P64].as<double2>(), clov_odd[prec].p, stream);
    clover_convert((int)Vh, prec, clovinv_even[P64].as<double2>(), clovinv_even[prec].p, stream);
  }
  CMX_HIP(hipGetLastError());
}

Vec& Op::work(int slot, int prec, int n) {
  if ((int)ws.size() <= slot) ws.resize(slot + 1);
  auto& v = ws[slot];
  if (!v || v->prec != prec || v->n != n) {
    v.reset(new Vec(prec, n));
    ++ws_gen;
  }
  return *v;
}

Vec& Op::tmp(int prec, int k) {
  auto& v = tmp_even[prec];
  while ((int)v.size() <= k) v.emplace_back(new Vec(prec, (int)Vh));
  return *v[k];
}

HopLaunch Op::hop_base(int prec, int which) const {
  HopLaunch h;
  std::memset(&h, 0, sizeof(h));
  h.g = g;
  h.prec = prec;
  h.recon = recon[prec];
  for (int p_ = 0; p_ < 2; ++p_) h.gauge[p_] = (prec == P64 && recon[P64] == 18) ? gd->g64[p_].p : gauge[prec][p_].p;
  std::memcpy(h.ph, gd->ph, sizeof(h.ph));
  hop_apply_config(h, hcfg[prec][which]);
  return h;
}

static void set_in(HopLaunch& h, VRef r) {
  h.in = r.v, h.in_nrm = r.nrm, h.in_stride = r.stride;
}
static void set_out(HopLaunch& h, VRef r) {
  h.out = r.v, h.out_nrm = r.nrm, h.out_stride = r.stride;
}
static void set_x(HopLaunch& h, VRef r) {
  h.x = r.v, h.x_nrm = r.nrm, h.x_stride = r.stride;
}

int Op::apply(int prec, VRef out, VRef in, int isign, double* parts, const int* skip, const int* run, bool ndot) {
  CMX_REQUIRE(!ndot || (parts && eo()), CMX_ERR_ARG, "apply: <chi, psi> partials need an even-odd operator");
  ensure_prec(prec);
  HopLaunch h = hop_base(prec);
  const int nb = hop_num_blocks(h);
  h.isign = isign;
  h.skip = skip;
  h.run = run;
  if (eo()) {
    // kernel 1: tmp_e = A_ee^-1 D_eo psi_o   (clover)   or   tmp_e = D_eo psi_o   (Wilson)
    Vec& t = tmp(prec, 0);
    h.cb = 0;
    set_in(h, in);
    set_out(h, t.view());
    h.epi = clover() ? EPI_CLOV : EPI_NONE;
    h.clov = clover() ? clovinv_even[prec].p : nullptr;
    h.b = 1.0;
    h.partials = nullptr;
    launch_hop(h, stream);
    // kernel 2: chi_o = A_oo psi_o - 1/4 D_oe tmp_e   (eoprec_clover_linop_w.cc:142-171)
    //       or  chi_o = (Nd+m) psi_o - 1/(4(Nd+m)) D_oe tmp_e   (eoprec_wilson_linop_w.cc:126-166)
    h.cb = 1;
    set_in(h, t.view());
    set_out(h, out);
    set_x(h, in);
    h.epi = clover() ? EPI_CLOV_XPAY : EPI_XPAY;
    h.clov = clover() ? clov_odd[prec].p : nullptr;
    h.a = fact;
    h.b = clover() ? -0.25 : -0.25 / fact;
    h.partials = parts;
    h.ndot = ndot ? 1 : 0;  // x = psi_o: <chi_o, psi_o> from the values the epilogue holds
    launch_hop(h, stream);
    return parts ? nb : 0;
  }
  // unpreconditioned Wilson on the full lattice: chi = (Nd+m) psi - 1/2 D psi  (unprec_wilson_linop_w.cc:58-77)
  VRef in_e = in, in_o = in, out_e = out, out_o = out;
  in_o.v = static_cast<char*>(in.v) + (size_t)Vh * 16;
  in_o.nrm = in.nrm ? in.nrm + Vh : nullptr;
  out_o.v = static_cast<char*>(out.v) + (size_t)Vh * 16;
  out_o.nrm = out.nrm ? out.nrm + Vh : nullptr;
  h.epi = EPI_XPAY;
  h.a = fact;
  h.b = -0.5;
  h.cb = 0;
  set_in(h, in_o);
  set_x(h, in_e);
  set_out(h, out_e);
  h.partials = parts;
  launch_hop(h, stream);
  h.cb = 1;
  set_in(h, in_e);
  set_x(h, in_o);
  set_out(h, out_o);
  h.partials = parts ? parts + nb : nullptr;
  launch_hop(h, stream);
  return parts ? 2 * nb : 0;
}

void Op::dslash(int prec, VRef out, VRef in, int isign, int cb, double b) {
  ensure_prec(prec);
  HopLaunch h = hop_base(prec, 1);
  h.isign = isign;
  h.cb = cb;
  set_in(h, in);
  set_out(h, out);
  h.epi = EPI_NONE;
  h.b = b;
  launch_hop(h, stream);
}

bool Op::dslash_multi(int prec, const VRef* out, const VRef* in, int n, int isign, int cb, double b) {
  if (prec != P16 || recon[P16] != 12) return false;
  ensure_prec(prec);
  MultiHopLaunch M;
  std::memset(&M, 0, sizeof(M));
  M.h = hop_base(prec, 1);
  HopConfig c = hcfg[prec][1];
  // the multi-source kernel is a plane-grid kernel whatever the single-source choice; it keeps the
  // link cache policy of that choice
  c.fast = 1 | (c.fast & HOP_FAST_STREAM);
  hop_apply_config(M.h, c);
  if (!(M.h.fast & 1)) return false;
  M.h.isign = isign;
  M.h.cb = cb;
  M.h.epi = EPI_NONE;
  M.h.b = b;
  // group 1..4: the vector-unit kernel with that many sources per lane; MULTI_MFMA: the link products
  // on the matrix cores (hop_mfma.hip), all sources per launch
  auto run = [&]() {
    for (int r0 = 0; r0 < n; r0 += MULTI_MAXR) {
      M.nrhs = std::min(MULTI_MAXR, n - r0);
      for (int k = 0; k < M.nrhs; ++k) {
        M.in[k] = in[r0 + k];
        M.out[k] = out[r0 + k];
      }
      if (M.group == MULTI_MFMA)
        launch_hop_mfma(M, stream);
      else
        launch_hop_multi(M, stream);
    }
  };
  if (multi_group == 0) {  // first use: time groups of 1 to 4 sources (registers vs link reuse)
    hipEvent_t e0, e1;
    CMX_HIP(hipEventCreate(&e0));
    CMX_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    for (int gsz = 1; gsz <= 4; ++gsz) {
      M.group = gsz;
      run();
      run();
      CMX_HIP(hipEventRecord(e0, stream));
      run();
      run();
      run();
      CMX_HIP(hipEventRecord(e1, stream));
      CMX_HIP(hipEventSynchronize(e1));
      float ms = 0;
      CMX_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) best = ms, multi_group = gsz;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  M.group = multi_group;
  run();
  return true;
}

void Op::upload(const double* host, size_t n) {
  CMX_REQUIRE(n * sizeof(double) <= staging.bytes, CMX_ERR_ARG, "upload: too large");
  CMX_HIP(hipMemcpyAsync(staging.p, host, n * sizeof(double), hipMemcpyHostToDevice, stream));
}
void Op::download(double* host, size_t n) {
  CMX_REQUIRE(n * sizeof(double) <= staging.bytes, CMX_ERR_ARG, "download: too large");
  CMX_HIP(hipMemcpyAsync(host, staging.p, n * sizeof(double), hipMemcpyDeviceToHost, stream));
  CMX_HIP(hipStreamSynchronize(stream));
}
void Op::pack(const double* host, int nsites, const Vec& v, int off) {
  upload(host, (size_t)nsites * 24);
  VRef r = v.view(off);
  spinor_pack(v.prec, nsites, staging.as<double>(), r.v, r.nrm, r.stride, stream);
  CMX_HIP(hipGetLastError());
}
void Op::unpack(double* host, int nsites, const Vec& v, int off) {
  VRef r = v.view(off);
  spinor_unpack(v.prec, nsites, r.v, r.nrm, r.stride, staging.as<double>(), stream);
  CMX_HIP(hipGetLastError());
  download(host, (size_t)nsites * 24);
}

void Op::blas(int opc, std::initializer_list<const Vec*> vs, int ia, int ib, int skip, int run, int n, int off) {
  BlasLaunch b;
  std::memset(&b, 0, sizeof(b));
  int prec[4];
  const Vec* first = *vs.begin();
  int k = 0;
  for (const Vec* v : vs) {
    b.v[k] = v->view(off);
    prec[k] = v->prec;
    ++k;
  }
  for (; k < 4; ++k) {
    b.v[k] = first->view(off);
    prec[k] = first->prec;
  }
  b.n = n < 0 ? first->n - off : n;
  b.S = state_S.as<double>();
  b.F = state_F.as<int>();
  b.ia = ia;
  b.ib = ib;
  b.partials = partials.as<double>();
  b.skip_flag = skip;
  b.run_flag = run;
  if (Recorder* r = active_recorder()) {
    record_blas(*r, opc, b, prec);
    return;
  }
  launch_blas(opc, b, prec, stream);
}

void Op::finalize(int nparts, int nred, int prog, int slot) {
  const FinLaunch f{partials.as<double>(), nparts, nred, state_S.as<double>(), state_F.as<int>(), prog, slot};
  if (Recorder* r = active_recorder()) {
    record_fin(*r, f);
    return;
  }
  launch_finalize(f, stream);
}

void Op::sync() {
  CMX_REQUIRE(active_recorder() == nullptr, CMX_ERR_ARG, "synchronisation inside a recorded sequence");
  CMX_HIP(hipGetLastError());
  CMX_HIP(hipStreamSynchronize(stream));
}

namespace {
struct TuneKey {
  int L[4], kind, prec;
  bool operator<(const TuneKey& o) const { return std::memcmp(this, &o, sizeof(TuneKey)) < 0; }
};
std::map<TuneKey, HopConfig>& tune_cache() {
  static std::map<TuneKey, HopConfig> m;
  return m;
}
std::mutex& tune_mutex() {
  static std::mutex m;
  return m;
}
}  // namespace

void Op::autotune() {
  // CMX_TUNE_VERBOSE=1 prints every candidate's time (like QUDA's tuning verbosity)
  const char* ev = std::getenv("CMX_TUNE_VERBOSE");
  const bool verbose = ev && ev[0] == '1';
  std::vector<int> precs = {P64, par.sloppy_prec};
  if (par.sloppy_prec == P16) precs.push_back(P32);
  const int ns = subset_sites();
  hipEvent_t e0, e1;
  CMX_HIP(hipEventCreate(&e0));
  CMX_HIP(hipEventCreate(&e1));
  for (int prec : precs)
    for (int which = 0; which < 2; ++which) {
      TuneKey key;
      std::memset(&key, 0, sizeof(key));
      for (int m = 0; m < 4; ++m) key.L[m] = g.L[m];
      key.kind = par.kind;
      key.prec = prec + 3 * which;
      {
        std::lock_guard<std::mutex> lk(tune_mutex());
        auto it = tune_cache().find(key);
        if (it != tune_cache().end()) {
          hcfg[prec][which] = it->second;
          continue;
        }
      }
      ensure_prec(prec);
      Vec in(prec, ns), out(prec, ns);
      {  // a zero input times differently from data (up to 25 % at 48^3x96), so tune on noise
        VRef r = in.view();
        spinor_noise(prec, ns, 12345u, r.v, r.nrm, r.stride, stream);
        CMX_HIP(hipGetLastError());
      }
      std::vector<HopConfig> cands;
      for (int block : {64, 128, 256}) {
        {  // plane-grid kernel, linear or XCD-slab order, links streamed or not (where admitted)
          for (int fm : {1, 1 | HOP_FAST_STREAM}) {
            cands.push_back(HopConfig{block, 0, 1, 0, fm});
            for (int zb : {1, 4, 8}) cands.push_back(HopConfig{block, 1, zb, 0, fm});
          }
        }
        for (int pf : {0, 1}) {  // site-linear kernel; streamed loads on the single-buffered path
          if (pf && prec == P64) continue;
          for (int st : {0, HOP_FAST_STREAM}) {
            if (pf && st) continue;
            cands.push_back(HopConfig{block, 0, 1, pf, st});
            for (int zb : {2, 8}) cands.push_back(HopConfig{block, 1, zb, pf, st});
          }
        }
      }
      float best = 1e30f;
      HopConfig bestc = hcfg[prec][which];
      const HopConfig saved = hcfg[prec][which];
      for (const HopConfig& c : cands) {
        HopLaunch probe = hop_base(prec, which);
        hop_apply_config(probe, c);
        if ((c.tiled && !probe.tiled) || ((c.fast & 1) && !(probe.fast & 1))) continue;  // not admitted here
        hcfg[prec][which] = c;
        auto run = [&]() {
          if (which == 0)
            apply(prec, out.view(), in.view(), +1, nullptr);
          else
            dslash(prec, out.view(), in.view(), +1, 1);
        };
        // Sustained rate: some shapes run fast for the first few launches and settle 20 % slower after
        // a few milliseconds of back-to-back launches (the site-linear kernel at 48^3x96), so each
        // candidate runs about 4 ms untimed, then about 4 ms timed.
        CMX_HIP(hipEventRecord(e0, stream));
        run();
        CMX_HIP(hipEventRecord(e1, stream));
        CMX_HIP(hipEventSynchronize(e1));
        float t1 = 0;
        CMX_HIP(hipEventElapsedTime(&t1, e0, e1));
        const int nrep = std::max(5, std::min(400, (int)(4.0f / std::max(t1, 1e-3f))));
        for (int k = 0; k < nrep; ++k) run();
        CMX_HIP(hipEventRecord(e0, stream));
        for (int k = 0; k < nrep; ++k) run();
        CMX_HIP(hipEventRecord(e1, stream));
        CMX_HIP(hipEventSynchronize(e1));
        float ms = 0;
        CMX_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms *= 5.0f / (float)nrep;  // per five launches, as printed below
        if (verbose)
          std::fprintf(stderr, "cmx tune prec %d %s block %d tiled %d zb %d prefetch %d kernel %d: %.2f us\n", prec,
                       which ? "hop" : "operator", c.block, c.tiled, c.zb, c.prefetch, c.fast, ms * 200.0f);
        if (ms < best) {
          best = ms;
          bestc = c;
        }
      }
      hcfg[prec][which] = best < 1e29f ? bestc : saved;
      std::lock_guard<std::mutex> lk(tune_mutex());
      tune_cache()[key] = hcfg[prec][which];
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  CMX_HIP(hipGetLastError());
}

}  // namespace cmx
