// extern "C" entry points declared in include/chroma_mi355x.h.  No exception crosses this boundary:
// every call returns CMX_OK or a CMX_ERR_* code and records the message for cmx_last_error().
#include <cmath>
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "mg.h"
#include "op.h"
#include "setup.h"

struct cmx_gauge_s {
  std::shared_ptr<cmx::GaugeData> gd;
};
struct cmx_op_s {
  std::unique_ptr<cmx::Op> op;
};
struct cmx_net_s {
  std::shared_ptr<cmx::LearnedNet> net;
};

namespace {
thread_local std::string g_err;
bool g_init = false;
int g_device = -1;

template <typename Fn> int guarded(Fn&& fn) {
  try {
    fn();
    return CMX_OK;
  } catch (const cmx::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return CMX_ERR_ARG;
  } catch (...) {
    g_err = "unknown error";
    return CMX_ERR_ARG;
  }
}

void require_init() { CMX_REQUIRE(g_init, CMX_ERR_STATE, "library not initialised: call cmx_init(device) first"); }
cmx::Op& O(cmx_op h) {
  CMX_REQUIRE(h && h->op, CMX_ERR_ARG, "null operator handle");
  return *h->op;
}
cmx::LearnedNet& N(cmx_net h) {
  CMX_REQUIRE(h && h->net, CMX_ERR_ARG, "null network handle");
  return *h->net;
}
}  // namespace

extern "C" {

const char* cmx_last_error(void) { return g_err.c_str(); }
int cmx_version(void) { return 1; }

int cmx_init(int device) {
  return guarded([&] {
    int n = 0;
    CMX_HIP(hipGetDeviceCount(&n));
    CMX_REQUIRE(device >= 0 && device < n, CMX_ERR_ARG, "cmx_init: no such device");
    CMX_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    CMX_HIP(hipGetDeviceProperties(&prop, device));
    CMX_REQUIRE(std::strncmp(prop.gcnArchName, "gfx950", 6) == 0, CMX_ERR_UNSUPPORTED,
                std::string("cmx_init: built for gfx950, device is ") + prop.gcnArchName);
    g_device = device;
    cmx::meson_scratch_setup();  // the allocation-free meson contraction's work space on this device
    g_init = true;
  });
}

int cmx_finalize(void) {
  return guarded([&] {
    if (g_init) CMX_HIP(hipDeviceSynchronize());
    g_init = false;
  });
}

int cmx_gauge_create(const int dims[4], const double* links, cmx_gauge* out) {
  return guarded([&] {
    require_init();
    CMX_REQUIRE(out, CMX_ERR_ARG, "gauge_create: null output");
    CMX_REQUIRE(dims && links, CMX_ERR_ARG, "gauge_create: null argument");
    const size_t V = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
    const double* dirs[4] = {links, links + V * 18, links + 2 * V * 18, links + 3 * V * 18};
    auto h = std::make_unique<cmx_gauge_s>();
    h->gd = cmx::make_gauge(dims, dirs);
    *out = h.release();
  });
}

int cmx_gauge_create_dirs(const int dims[4], const double* const links[4], cmx_gauge* out) {
  return guarded([&] {
    require_init();
    CMX_REQUIRE(out, CMX_ERR_ARG, "gauge_create_dirs: null output");
    auto h = std::make_unique<cmx_gauge_s>();
    h->gd = cmx::make_gauge(dims, links);
    *out = h.release();
  });
}

int cmx_gauge_destroy(cmx_gauge g) {
  return guarded([&] { delete g; });
}

int cmx_op_create(cmx_gauge g, const cmx_op_params* p, cmx_op* out) {
  return guarded([&] {
    require_init();
    CMX_REQUIRE(g && p && out, CMX_ERR_ARG, "op_create: null argument");
    auto h = std::make_unique<cmx_op_s>();
    h->op.reset(cmx::make_op(g->gd, *p));
    *out = h.release();
  });
}

int cmx_op_destroy(cmx_op op) {
  return guarded([&] { delete op; });
}

long cmx_op_subset_doubles(cmx_op op) {
  if (!op || !op->op) return -1;
  return 24L * op->op->subset_sites();
}

int cmx_op_apply(cmx_op h, double* chi, const double* psi, int isign, int prec) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(chi && psi, CMX_ERR_ARG, "op_apply: null array");
    CMX_REQUIRE(isign == 1 || isign == -1, CMX_ERR_ARG, "op_apply: isign must be +1 or -1");
    CMX_REQUIRE(prec >= 0 && prec <= 2, CMX_ERR_ARG, "op_apply: bad precision");
    const int n = op.subset_sites();
    cmx::Vec in(prec, n), out(prec, n);
    op.pack(psi, n, in);
    op.apply(prec, out.view(), in.view(), isign, nullptr);
    op.unpack(chi, n, out);
  });
}

int cmx_dslash_apply(cmx_op h, double* chi, const double* psi, int isign, int cb, int prec) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(chi && psi, CMX_ERR_ARG, "dslash_apply: null array");
    CMX_REQUIRE(isign == 1 || isign == -1, CMX_ERR_ARG, "dslash_apply: isign must be +1 or -1");
    CMX_REQUIRE(cb == 0 || cb == 1, CMX_ERR_ARG, "dslash_apply: cb must be 0 or 1");
    CMX_REQUIRE(prec >= 0 && prec <= 2, CMX_ERR_ARG, "dslash_apply: bad precision");
    const int n = (int)op.Vh;
    cmx::Vec in(prec, n), out(prec, n);
    op.pack(psi, n, in);
    op.dslash(prec, out.view(), in.view(), isign, cb);
    op.unpack(chi, n, out);
  });
}

int cmx_dslash_multi(cmx_op h, int nrhs, double* chi, const double* psi, int isign, int cb, int prec) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(chi && psi && nrhs >= 1, CMX_ERR_ARG, "dslash_multi: bad argument");
    CMX_REQUIRE(isign == 1 || isign == -1, CMX_ERR_ARG, "dslash_multi: isign must be +1 or -1");
    CMX_REQUIRE(cb == 0 || cb == 1, CMX_ERR_ARG, "dslash_multi: cb must be 0 or 1");
    CMX_REQUIRE(prec >= 0 && prec <= 2, CMX_ERR_ARG, "dslash_multi: bad precision");
    const int n = (int)op.Vh;
    const size_t nd = (size_t)24 * n;
    std::vector<std::unique_ptr<cmx::Vec>> in, out;
    std::vector<cmx::VRef> vi, vo;
    for (int k = 0; k < nrhs; ++k) {
      in.emplace_back(new cmx::Vec(prec, n));
      out.emplace_back(new cmx::Vec(prec, n));
      op.pack(psi + k * nd, n, *in.back());
      vi.push_back(in.back()->view());
      vo.push_back(out.back()->view());
    }
    if (!op.dslash_multi(prec, vo.data(), vi.data(), nrhs, isign, cb))
      for (int k = 0; k < nrhs; ++k) op.dslash(prec, vo[k], vi[k], isign, cb);
    for (int k = 0; k < nrhs; ++k) op.unpack(chi + k * nd, n, *out[k]);
  });
}

int cmx_dslash_deriv(cmx_op h, double* ds_u, const double* chi, const double* psi, int isign, int cb) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(ds_u && chi && psi, CMX_ERR_ARG, "dslash_deriv: null array");
    CMX_REQUIRE(isign == 1 || isign == -1, CMX_ERR_ARG, "dslash_deriv: isign must be +1 or -1");
    CMX_REQUIRE(cb >= -1 && cb <= 1, CMX_ERR_ARG, "dslash_deriv: cb must be 0, 1 or -1");
    const long Vh = op.Vh, V = op.V;
    cmx::Vec a(cmx::P64, (int)Vh), b(cmx::P64, (int)Vh);
    cmx::DevBuf ds((size_t)4 * 9 * 16 * Vh);
    std::vector<double> host((size_t)4 * 18 * Vh);
    std::memset(ds_u, 0, sizeof(double) * 4 * 18 * V);
    for (int c = 0; c < 2; ++c) {
      if (cb >= 0 && c != cb) continue;
      op.pack(chi + (size_t)c * 24 * Vh, (int)Vh, a);
      op.pack(psi + (size_t)(1 - c) * 24 * Vh, (int)Vh, b);
      cmx::hop_deriv(op.g, c, isign, a.view(), b.view(), ds.p, op.stream);
      CMX_HIP(hipGetLastError());
      CMX_HIP(hipMemcpyAsync(host.data(), ds.p, sizeof(double) * host.size(), hipMemcpyDeviceToHost, op.stream));
      op.sync();
      for (int mu = 0; mu < 4; ++mu)
        std::memcpy(ds_u + ((size_t)mu * V + (size_t)c * Vh) * 18, host.data() + (size_t)mu * Vh * 18,
                    sizeof(double) * 18 * Vh);
    }
  });
}

namespace {
// ds (device, [4][V][9] double2) -> host ds_u [4][V][18] doubles
void download_force(cmx::Op& op, const cmx::DevBuf& ds, double* ds_u) {
  CMX_HIP(hipMemcpyAsync(ds_u, ds.p, sizeof(double) * 4 * 18 * op.V, hipMemcpyDeviceToHost, op.stream));
  op.sync();
}
}  // namespace

int cmx_clover_deriv(cmx_op h, double* ds_u, const double* chi, const double* psi, int isign, int cb) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(op.clover(), CMX_ERR_ARG, "clover_deriv: not a clover operator");
    CMX_REQUIRE(ds_u && chi && psi, CMX_ERR_ARG, "clover_deriv: null array");
    CMX_REQUIRE(isign == 1 || isign == -1, CMX_ERR_ARG, "clover_deriv: isign must be +1 or -1");
    CMX_REQUIRE(cb >= -1 && cb <= 1, CMX_ERR_ARG, "clover_deriv: cb must be 0, 1 or -1");
    const long Vh = op.Vh;
    cmx::Vec a(cmx::P64, (int)Vh), b(cmx::P64, (int)Vh);
    cmx::DevBuf ds((size_t)4 * 9 * 16 * op.V);
    CMX_HIP(hipMemsetAsync(ds.p, 0, ds.bytes, op.stream));
    for (int c = 0; c < 2; ++c) {
      if (cb >= 0 && c != cb) continue;
      op.pack(chi + (size_t)c * 24 * Vh, (int)Vh, a);
      op.pack(psi + (size_t)c * 24 * Vh, (int)Vh, b);
      cmx::clover_force(op, false, c, a.view(), b.view(), 1.0, ds.p);
    }
    download_force(op, ds, ds_u);
  });
}

int cmx_clover_deriv_trln(cmx_op h, double* ds_u) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(op.clover(), CMX_ERR_ARG, "clover_deriv_trln: not a clover operator");
    CMX_REQUIRE(ds_u, CMX_ERR_ARG, "clover_deriv_trln: null array");
    cmx::DevBuf ds((size_t)4 * 9 * 16 * op.V);
    CMX_HIP(hipMemsetAsync(ds.p, 0, ds.bytes, op.stream));
    cmx::clover_force(op, true, 0, cmx::VRef{}, cmx::VRef{}, 1.0, ds.p);
    download_force(op, ds, ds_u);
  });
}

int cmx_op_deriv(cmx_op h, double* ds_u, const double* chi, const double* psi, int isign) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(op.clover(), CMX_ERR_UNSUPPORTED, "op_deriv: the even-odd clover operator only");
    CMX_REQUIRE(ds_u && chi && psi, CMX_ERR_ARG, "op_deriv: null array");
    CMX_REQUIRE(isign == 1 || isign == -1, CMX_ERR_ARG, "op_deriv: isign must be +1 or -1");
    const int Vh = (int)op.Vh;
    cmx::Vec co(cmx::P64, Vh), po(cmx::P64, Vh), t2(cmx::P64, Vh), t3(cmx::P64, Vh);
    cmx::DevBuf ds((size_t)4 * 9 * 16 * op.V), dcb((size_t)4 * 9 * 16 * Vh);
    CMX_HIP(hipMemsetAsync(ds.p, 0, ds.bytes, op.stream));
    op.pack(chi, Vh, co);
    op.pack(psi, Vh, po);
    // t = A_ee^-1 D_eo(sign) v, the first kernel of the even-odd operator at FP64
    auto ainv_d = [&](cmx::Vec& t, const cmx::Vec& v, int sign) {
      cmx::HopLaunch hh = op.hop_base(cmx::P64);
      hh.isign = sign;
      hh.cb = 0;
      const cmx::VRef vi = v.view(), vo = t.view();
      hh.in = vi.v, hh.in_nrm = nullptr, hh.in_stride = vi.stride;
      hh.out = vo.v, hh.out_nrm = nullptr, hh.out_stride = vo.stride;
      hh.epi = cmx::EPI_CLOV;
      hh.clov = op.clovinv_even[cmx::P64].p;
      hh.b = 1.0;
      hh.partials = nullptr;
      cmx::launch_hop(hh, op.stream);
      CMX_HIP(hipGetLastError());
    };
    // EvenOddPrecLogDetLinearOperator::deriv (eoprec_logdet_linop.h:114-157) with the reference's
    // tmp2 = A_ee^-1 (-1/2 D_eo psi) = -t2/2 and tmp3 = A_ee^-1 (-1/2 D_eo(-isign) chi) = -t3/2:
    //   ds  = clov.deriv(chi, psi, cb 1)
    //   ds -= -1/2 D.deriv(chi, tmp2, cb 1)     = -1/4 D.deriv(chi, t2, cb 1)
    //   ds += clov.deriv(tmp3, tmp2, cb 0)      = +1/4 clov.deriv(t3, t2, cb 0)
    //   ds -= -1/2 D.deriv(tmp3, psi, cb 0)     = -1/4 D.deriv(t3, psi, cb 0)
    cmx::clover_force(op, false, 1, co.view(), po.view(), 1.0, ds.p);
    ainv_d(t2, po, isign);
    cmx::hop_deriv(op.g, 1, isign, co.view(), t2.view(), dcb.p, op.stream);
    cmx::force_axpy(op, 1, -0.25, dcb.p, ds.p);
    ainv_d(t3, co, -isign);
    cmx::clover_force(op, false, 0, t3.view(), t2.view(), 0.25, ds.p);
    cmx::hop_deriv(op.g, 0, isign, t3.view(), po.view(), dcb.p, op.stream);
    cmx::force_axpy(op, 0, -0.25, dcb.p, ds.p);
    download_force(op, ds, ds_u);
  });
}

int cmx_gauge_stout_smear(const int dims[4], double* links, int nstep, double rho) {
  return guarded([&] {
    require_init();
    cmx::stout_smear(dims, links, nstep, rho, 15);
  });
}

int cmx_gauge_stout_smear_dirs(const int dims[4], double* links, int nstep, double rho, int smear_dirs) {
  return guarded([&] {
    require_init();
    cmx::stout_smear(dims, links, nstep, rho, smear_dirs);
  });
}

int cmx_meson_2pt(const int dims[4], const double* prop1, const double* prop2, const int origin[4], int nmom,
                  const int* moms, double* corr) {
  return guarded([&] {
    require_init();
    cmx::meson_2pt(dims, prop1, prop2, origin, nmom, moms, corr);
  });
}

int cmx_meson_2pt_noalloc(const int dims[4], const double* prop1, const double* prop2, const int origin[4],
                          int nmom, const int* moms, double* corr, void* stream) {
  return guarded([&] {
    require_init();
    cmx::meson_2pt_noalloc(dims, prop1, prop2, origin, nmom, moms, corr, static_cast<hipStream_t>(stream));
  });
}

int cmx_op_eo_piece(cmx_op h, int which, double* chi, const double* psi, int isign) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(op.eo(), CMX_ERR_ARG, "op_eo_piece: operator is not even-odd preconditioned");
    CMX_REQUIRE(chi && psi, CMX_ERR_ARG, "op_eo_piece: null array");
    CMX_REQUIRE(which >= 0 && which <= 5, CMX_ERR_ARG, "op_eo_piece: bad block");
    const int Vh = (int)op.Vh;
    cmx::Vec in(cmx::P64, 2 * Vh), out(cmx::P64, 2 * Vh);
    op.pack(psi, 2 * Vh, in);
    cmx::HopLaunch h0 = op.hop_base(cmx::P64);
    h0.isign = isign;
    auto hop = [&](int cb, int epi, const void* clov, double a, double b, bool with_x) {
      cmx::HopLaunch hh = h0;
      hh.cb = cb;
      cmx::VRef vin = in.view((1 - cb) * Vh), vout = out.view(cb * Vh), vx = in.view(cb * Vh);
      hh.in = vin.v, hh.in_nrm = nullptr, hh.in_stride = vin.stride;
      hh.out = vout.v, hh.out_nrm = nullptr, hh.out_stride = vout.stride;
      if (with_x) hh.x = vx.v, hh.x_nrm = nullptr, hh.x_stride = vx.stride;
      hh.epi = epi;
      hh.clov = clov;
      hh.a = a;
      hh.b = b;
      cmx::launch_hop(hh, op.stream);
    };
    // hop with a zero input gives b*D*0 = 0, so the "local" blocks below are the EPI_XPAY / EPI_CLOV_XPAY
    // epilogues with b = 0 applied to a zero neighbour field
    cmx::Vec zero(cmx::P64, 2 * Vh);
    op.blas(cmx::B_ZERO, {&zero});
    auto local = [&](int cb, const void* clov, double a) {
      cmx::HopLaunch hh = h0;
      hh.cb = cb;
      cmx::VRef vz = zero.view((1 - cb) * Vh), vout = out.view(cb * Vh), vx = in.view(cb * Vh);
      hh.in = vz.v, hh.in_stride = vz.stride;
      hh.out = vout.v, hh.out_stride = vout.stride;
      hh.x = vx.v, hh.x_stride = vx.stride;
      hh.epi = clov ? cmx::EPI_CLOV_XPAY : cmx::EPI_XPAY;
      hh.clov = clov;
      hh.a = a;
      hh.b = 0.0;
      cmx::launch_hop(hh, op.stream);
    };
    const bool cl = op.clover();
    switch (which) {
      case 0: local(0, cl ? op.clov64_even.p : nullptr, op.fact); break;                    // A_ee
      case 1: local(0, cl ? op.clovinv_even[cmx::P64].p : nullptr, 1.0 / op.fact); break;  // A_ee^-1
      case 2: hop(0, cmx::EPI_NONE, nullptr, 0, -0.5, false); break;                        // D_eo = -1/2 D
      case 3: hop(1, cmx::EPI_NONE, nullptr, 0, -0.5, false); break;                        // D_oe
      case 4: local(1, cl ? op.clov_odd[cmx::P64].p : nullptr, op.fact); break;             // A_oo
      case 5:  // unprecLinOp (eoprec_linop.h:147-165): chi_cb = A_cb psi_cb - 1/2 D psi_{1-cb}
        hop(0, cl ? cmx::EPI_CLOV_XPAY : cmx::EPI_XPAY, cl ? op.clov64_even.p : nullptr, op.fact, -0.5, true);
        hop(1, cl ? cmx::EPI_CLOV_XPAY : cmx::EPI_XPAY, cl ? op.clov_odd[cmx::P64].p : nullptr, op.fact, -0.5, true);
        break;
    }
    CMX_HIP(hipGetLastError());
    // write back only the target checkerboard(s)
    std::unique_ptr<double[]> full(new double[(size_t)48 * Vh]);
    op.unpack(full.get(), 2 * Vh, out);
    const int lo = (which <= 2) ? 0 : (which == 5 ? 0 : Vh);
    const int hi = (which <= 2) ? Vh : 2 * Vh;
    std::memcpy(chi + (size_t)24 * lo, full.get() + (size_t)24 * lo, sizeof(double) * 24 * (hi - lo));
  });
}

int cmx_clover_download(cmx_op h, int which, double* tri) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(op.clover(), CMX_ERR_ARG, "clover_download: not a clover operator");
    CMX_REQUIRE(tri && (which == 0 || which == 1), CMX_ERR_ARG, "clover_download: bad argument");
    const int Vh = (int)op.Vh;
    cmx::DevBuf d((size_t)72 * sizeof(double) * Vh);
    auto grab = [&](const cmx::DevBuf& src, int cb) {
      cmx::clover_unpack(Vh, src.as<double2>(), d.as<double>(), op.stream);
      CMX_HIP(hipGetLastError());
      CMX_HIP(hipMemcpyAsync(tri + (size_t)72 * cb * Vh, d.p, (size_t)72 * sizeof(double) * Vh, hipMemcpyDeviceToHost,
                             op.stream));
      op.sync();
    };
    if (which == 0) {
      grab(op.clov64_even, 0);
      grab(op.clov_odd[cmx::P64], 1);
    } else {
      grab(op.clovinv_even[cmx::P64], 0);
    }
  });
}

int cmx_clover_logdet(cmx_op h, double* logdet, long* n_neg_sites) {
  return guarded([&] {
    cmx::Op& op = O(h);
    CMX_REQUIRE(op.clover(), CMX_ERR_ARG, "clover_logdet: not a clover operator");
    if (logdet) *logdet = op.logdet_even;
    if (n_neg_sites) *n_neg_sites = op.nneg_even;
  });
}

int cmx_clover_pack_quda(cmx_op h, int which, int cb, double* out) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(op.clover(), CMX_ERR_ARG, "clover_pack_quda: not a clover operator");
    CMX_REQUIRE(out && (which == 0 || which == 1) && (cb == 0 || cb == 1), CMX_ERR_ARG,
                "clover_pack_quda: bad argument");
    const int Vh = (int)op.Vh;
    std::unique_ptr<double[]> tri(new double[(size_t)72 * Vh]);
    cmx::DevBuf d((size_t)72 * sizeof(double) * Vh);
    if (which == 1 && cb == 1) {
      // A_oo^-1 is not kept by the operator (the even-odd solve needs only A_ee^-1): LDL^dag of A_oo
      // here, as the QUDA set-up inverts both checkerboards (choles(0), choles(1))
      cmx::DevBuf inv((size_t)36 * 16 * Vh), ld(sizeof(double) * Vh), nn(sizeof(double) * Vh);
      cmx::clover_invert(op.g, op.clov_odd[cmx::P64].as<double2>(), inv.as<double2>(), ld.as<double>(),
                         nn.as<double>(), op.stream);
      cmx::clover_unpack(Vh, inv.as<double2>(), d.as<double>(), op.stream);
      op.sync();
    } else {
      const cmx::DevBuf& src = which == 1 ? op.clovinv_even[cmx::P64] : (cb == 0 ? op.clov64_even : op.clov_odd[cmx::P64]);
      cmx::clover_unpack(Vh, src.as<double2>(), d.as<double>(), op.stream);
    }
    CMX_HIP(hipMemcpyAsync(tri.get(), d.p, (size_t)72 * sizeof(double) * Vh, hipMemcpyDeviceToHost, op.stream));
    op.sync();
    // QUDAPackedClovSite: diag1[6], offDiag1[15][2], diag2[6], offDiag2[15][2]; the lower triangle
    // column by column (packForQUDA, clover_term_qdp_w.h:2185-2222), written at the full-lattice site
    // index of each site of checkerboard cb (quda_array[rb[cb].siteTable()[ssite]]); the other
    // checkerboard's entries of out are left as they are
    double* outcb = out + (size_t)72 * cb * Vh;
    for (int i = 0; i < Vh; ++i) {
      const double* t = tri.get() + (size_t)72 * i;
      double* q = outcb + (size_t)72 * i;
      for (int b = 0; b < 2; ++b) {
        double* qb = q + 36 * b;
        for (int k = 0; k < 6; ++k) qb[k] = t[6 * b + k];
        int target = 0;
        for (int col = 0; col < 5; ++col)
          for (int row = col + 1; row < 6; ++row, ++target) {
            const int src_i = row * (row - 1) / 2 + col;
            qb[6 + 2 * target] = t[12 + 30 * b + 2 * src_i];
            qb[6 + 2 * target + 1] = t[12 + 30 * b + 2 * src_i + 1];
          }
      }
    }
  });
}

int cmx_solve(cmx_op h, const cmx_solver_params* p, double* psi, const double* chi, int nrhs, cmx_solver_result* res) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(p && psi && chi && res && nrhs >= 1, CMX_ERR_ARG, "solve: bad argument");
    const long nd = 24L * op.subset_sites();
    for (int k = 0; k < nrhs; ++k) cmx::solve(op, *p, psi + k * nd, chi + k * nd, false, res[k]);
  });
}

int cmx_solve_mdagm(cmx_op h, const cmx_solver_params* p, double* psi, const double* chi, int nrhs,
                    cmx_solver_result* res) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(p && psi && chi && res && nrhs >= 1, CMX_ERR_ARG, "solve_mdagm: bad argument");
    const long nd = 24L * op.subset_sites();
    for (int k = 0; k < nrhs; ++k) cmx::solve(op, *p, psi + k * nd, chi + k * nd, true, res[k]);
  });
}

int cmx_solve_multishift(cmx_op h, const cmx_solver_params* p, int nshift, const double* shifts, const double* rsd,
                         double* psi, const double* chi, cmx_solver_result* res) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(p && shifts && rsd && psi && chi && res && nshift >= 1, CMX_ERR_ARG, "solve_multishift: bad argument");
    cmx::solve_multishift(op, *p, nshift, shifts, rsd, psi, chi, res);
  });
}

int cmx_solve_batch(const cmx_op* hs, int n, const cmx_solver_params* p, const cmx_batch_params* bp,
                    double* const* psi, const double* const* chi, int mdagm, cmx_solver_result* res,
                    cmx_batch_result* bres) {
  return guarded([&] {
    require_init();
    CMX_REQUIRE(hs && n >= 1 && p && bp && psi && chi && res && bres, CMX_ERR_ARG, "solve_batch: bad argument");
    // a repeated operator gets a solver context of its own (sharing its fields and stream)
    std::vector<std::unique_ptr<cmx::Op>> clones;
    std::vector<cmx::Op*> ops(n);
    const bool sloppy = p->solver == CMX_SOLVER_RELIABLE_CG || p->solver == CMX_SOLVER_FGMRESDR;
    const int ps = sloppy && p->sloppy_prec >= 0 && p->sloppy_prec <= 2 ? p->sloppy_prec : cmx::P64;
    for (int i = 0; i < n; ++i) {
      CMX_REQUIRE(hs[i] && psi[i] && chi[i], CMX_ERR_ARG, "solve_batch: null operator or field");
      cmx::Op& o = O(hs[i]);
      bool seen = false;
      for (int j = 0; j < i; ++j) seen = seen || hs[j] == hs[i];
      if (seen) {
        clones.emplace_back(cmx::clone_op(o, ps));
        // the learned preconditioner: the clone uses the operator's network and device buffers
        if (p->solver == CMX_SOLVER_FGMRESDR && p->precond == CMX_PRECOND_LEARNED) cmx::learned_share(o, *clones.back());
        ops[i] = clones.back().get();
      } else {
        ops[i] = &o;
      }
    }
    cmx::solve_batch(ops, *p, *bp, psi, chi, mdagm != 0, res, *bres);
  });
}

int cmx_op_time(cmx_op h, int what, int prec, int n, double* ms_per_call) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(ms_per_call && n > 0 && prec >= 0 && prec <= 2, CMX_ERR_ARG, "op_time: bad argument");
    CMX_REQUIRE(what >= 0 && what <= 5, CMX_ERR_ARG, "op_time: bad selector");
    const int Vh = (int)op.Vh;
    const int ns = op.subset_sites();
    cmx::Vec a(prec, ns), b(prec, ns);
    // what = 4: the hopping term on 12 sources (distinct inputs and outputs)
    std::vector<std::unique_ptr<cmx::Vec>> ma, mb;
    std::vector<cmx::VRef> va, vb;
    if (what == 4) {
      for (int k = 0; k < 12; ++k) {
        ma.emplace_back(new cmx::Vec(prec, Vh));
        mb.emplace_back(new cmx::Vec(prec, Vh));
        va.push_back(ma.back()->view());
        vb.push_back(mb.back()->view());
      }
    }
    // what = 5: six streams of 16-byte elements, together the hopping term's algorithmic bytes at this
    // precision with 12-real links (FP16 296, FP32 576, FP64 1152 B per site)
    std::vector<cmx::DevBuf> mix;
    long n_mix = 0;
    if (what == 5) {
      const long site_bytes = prec == cmx::P16 ? 296 : (prec == cmx::P32 ? 576 : 1152);
      n_mix = (site_bytes * (long)Vh + 95) / 96;
      for (int k = 0; k < 6; ++k) {
        mix.emplace_back();
        mix.back().alloc((size_t)n_mix * 16);
        CMX_HIP(hipMemsetAsync(mix.back().p, 0x11 * (k + 1), (size_t)n_mix * 16, op.stream));
      }
    }
    // deterministic non-zero input
    {
      std::unique_ptr<double[]> hbuf(new double[(size_t)24 * ns]);
      for (long k = 0; k < 24L * ns; ++k) hbuf[k] = std::sin(0.37 * (double)k) * 0.5;
      op.pack(hbuf.get(), ns, a);
      for (auto& v : ma) op.pack(hbuf.get(), Vh, *v);
    }
    op.ensure_prec(prec);
    auto body = [&]() {
      switch (what) {
        case 0: op.dslash(prec, b.view(), a.view(), +1, 1); break;
        case 1: op.apply(prec, b.view(), a.view(), +1, nullptr); break;
        case 2: {
          cmx::HopLaunch hh = op.hop_base(prec);
          hh.isign = 1;
          hh.cb = 1;
          cmx::Vec& t = op.tmp(prec, 0);
          cmx::VRef vi = t.view(), vo = b.view(), vx = a.view();
          hh.in = vi.v, hh.in_nrm = vi.nrm, hh.in_stride = vi.stride;
          hh.out = vo.v, hh.out_nrm = vo.nrm, hh.out_stride = vo.stride;
          hh.x = vx.v, hh.x_nrm = vx.nrm, hh.x_stride = vx.stride;
          hh.epi = op.clover() ? cmx::EPI_CLOV_XPAY : cmx::EPI_XPAY;
          hh.clov = op.clover() ? op.clov_odd[prec].p : nullptr;
          hh.a = op.fact;
          hh.b = -0.25;
          cmx::launch_hop(hh, op.stream);
          break;
        }
        case 3: cmx::stream_copy((long)cmx::spinor_chunks(prec) * Vh, a.buf.p, b.buf.p, op.stream); break;
        case 5: {
          const void* in[5] = {mix[0].p, mix[1].p, mix[2].p, mix[3].p, mix[4].p};
          cmx::stream_mix(n_mix, in, mix[5].p, op.stream);
          break;
        }
        case 4:
          if (!op.dslash_multi(prec, vb.data(), va.data(), 12, +1, 1))
            for (int k = 0; k < 12; ++k) op.dslash(prec, vb[k], va[k], +1, 1);
          break;
      }
    };
    body();  // warm-up
    op.sync();
    hipEvent_t e0, e1;
    CMX_HIP(hipEventCreate(&e0));
    CMX_HIP(hipEventCreate(&e1));
    CMX_HIP(hipEventRecord(e0, op.stream));
    for (int k = 0; k < n; ++k) body();
    CMX_HIP(hipEventRecord(e1, op.stream));
    CMX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    CMX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    CMX_HIP(hipGetLastError());
    *ms_per_call = ms / n;
  });
}

int cmx_op_get_hop_config(cmx_op h, int prec, int cfg[6]) {
  return guarded([&] {
    cmx::Op& op = O(h);
    CMX_REQUIRE(cfg && prec >= 0 && prec <= 2, CMX_ERR_ARG, "get_hop_config: bad argument");
    const int which = cfg[5] == 2 ? 1 : 0;
    const cmx::HopConfig& c = op.hcfg[prec][which];
    // the kernel actually admitted for this operator (the plane-grid kernel needs 12-real links,
    // real boundary phases and a field below 2 GiB; otherwise the site-linear kernel runs)
    cfg[0] = c.block, cfg[1] = c.tiled, cfg[2] = c.zb, cfg[3] = c.prefetch, cfg[4] = op.hop_base(prec, which).fast;
  });
}

int cmx_op_set_multi_group(cmx_op h, int group) {
  return guarded([&] {
    CMX_REQUIRE(group >= 0 && group <= 5, CMX_ERR_ARG, "set_multi_group: group must be 0..5");
    O(h).multi_group = group;
  });
}

int cmx_op_link_recon(cmx_op h, int prec, int* recon) {
  return guarded([&] {
    cmx::Op& op = O(h);
    CMX_REQUIRE(recon && prec >= 0 && prec <= 2, CMX_ERR_ARG, "link_recon: bad argument");
    *recon = op.recon[prec];
  });
}

int cmx_op_set_hop_config(cmx_op h, int prec, const int cfg[6]) {
  return guarded([&] {
    cmx::Op& op = O(h);
    CMX_REQUIRE(cfg && prec >= 0 && prec <= 2, CMX_ERR_ARG, "set_hop_config: bad argument");
    CMX_REQUIRE(cfg[0] == 64 || cfg[0] == 128 || cfg[0] == 256, CMX_ERR_ARG, "set_hop_config: block must be 64, 128 or 256");
    CMX_REQUIRE((cfg[1] == 0 || cfg[1] == 1) && cfg[2] >= 1 && (cfg[3] == 0 || cfg[3] == 1) && cfg[4] >= 0 && cfg[4] <= 3,
                CMX_ERR_ARG,
                "set_hop_config: bad tiling / prefetch flags");
    CMX_REQUIRE(cfg[5] >= 0 && cfg[5] <= 2, CMX_ERR_ARG, "set_hop_config: target must be 0, 1 or 2");
    const cmx::HopConfig c{cfg[0], cfg[1], cfg[2], prec == cmx::P64 ? 0 : cfg[3], cfg[4]};
    if (cfg[5] != 2) op.hcfg[prec][0] = c;
    if (cfg[5] != 1) op.hcfg[prec][1] = c;
  });
}

int cmx_op_autotune(cmx_op h) {
  return guarded([&] {
    require_init();
    O(h).autotune();
  });
}

int cmx_layout_roundtrip(cmx_op h, int prec, const double* in, double* out) {
  return guarded([&] {
    require_init();
    cmx::Op& op = O(h);
    CMX_REQUIRE(in && out && prec >= 0 && prec <= 2, CMX_ERR_ARG, "layout_roundtrip: bad argument");
    const int n = (int)op.Vh;
    cmx::Vec v(prec, n);
    op.pack(in, n, v);
    op.unpack(out, n, v);
  });
}


// ---- learned preconditioner ------------------------------------------------------------------------
int cmx_net_create(cmx_net* out) {
  return guarded([&] {
    CMX_REQUIRE(out, CMX_ERR_ARG, "net_create: null output");
    auto* h = new cmx_net_s();
    h->net = std::make_shared<cmx::LearnedNet>();
    *out = h;
  });
}

int cmx_net_destroy(cmx_net net) {
  return guarded([&] { delete net; });
}

int cmx_net_init_neumann(cmx_net net, double diag) {
  return guarded([&] {
    CMX_REQUIRE(diag != 0, CMX_ERR_ARG, "net_init_neumann: diag must be non-zero");
    cmx::net_init_neumann(N(net), diag);
  });
}

int cmx_net_get_block(cmx_net net, int blk, double* w1, double* w2, double* w0) {
  return guarded([&] {
    cmx::LearnedNet& n = N(net);
    CMX_REQUIRE(blk >= 0 && blk < (int)n.blk.size(), CMX_ERR_ARG, "net_get_block: no such block");
    auto put = [](const std::vector<cmx::PtcBlock::cf>& w, double* o) {
      if (!o) return;
      for (size_t k = 0; k < w.size(); ++k) o[2 * k] = w[k].real(), o[2 * k + 1] = w[k].imag();
    };
    put(n.blk[blk].W1, w1);
    put(n.blk[blk].W2, w2);
    put(n.blk[blk].W0, w0);
  });
}

int cmx_net_set_block(cmx_net net, int blk, const double* w1, const double* w2, const double* w0) {
  return guarded([&] {
    cmx::LearnedNet& n = N(net);
    CMX_REQUIRE(blk >= 0 && blk < (int)n.blk.size(), CMX_ERR_ARG, "net_set_block: no such block");
    auto get = [](std::vector<cmx::PtcBlock::cf>& w, const double* i) {
      if (!i) return;
      for (size_t k = 0; k < w.size(); ++k) w[k] = cmx::PtcBlock::cf((float)i[2 * k], (float)i[2 * k + 1]);
    };
    get(n.blk[blk].W1, w1);
    get(n.blk[blk].W2, w2);
    get(n.blk[blk].W0, w0);
    n.version++;
  });
}

int cmx_net_get_weights(cmx_net net, double* w1, double* w2, double* w0) { return cmx_net_get_block(net, 0, w1, w2, w0); }

int cmx_net_set_weights(cmx_net net, const double* w1, const double* w2, const double* w0) {
  return cmx_net_set_block(net, 0, w1, w2, w0);
}

int cmx_net_set_depth(cmx_net net, int nblocks) {
  return guarded([&] { cmx::net_set_depth(N(net), nblocks); });
}

int cmx_net_depth(cmx_net net, int* nblocks) {
  return guarded([&] {
    CMX_REQUIRE(nblocks, CMX_ERR_ARG, "net_depth: null pointer");
    *nblocks = (int)N(net).blk.size();
  });
}

int cmx_op_set_net(cmx_op op, cmx_net net) {
  return guarded([&] {
    require_init();
    cmx::Op& o = O(op);
    o.net = net ? net->net : nullptr;
    o.net_version = -1;
  });
}

int cmx_op_set_net_variant(cmx_op op, int prec, int variant) {
  return guarded([&] {
    require_init();
    cmx::net_set_variant(O(op), prec, variant);
  });
}

int cmx_net_apply(cmx_op op, int prec, double* out, const double* in) {
  return guarded([&] {
    require_init();
    CMX_REQUIRE(out && in && prec >= 0 && prec <= 2, CMX_ERR_ARG, "net_apply: bad argument");
    cmx::net_apply_host(O(op), prec, out, in);
  });
}

int cmx_net_grad(cmx_op op, int prec, const double* v, const double* e, double* loss, double* g_wc, double* g_w0) {
  return guarded([&] {
    require_init();
    CMX_REQUIRE(v && e && loss && g_wc && g_w0 && prec >= 0 && prec <= 2, CMX_ERR_ARG, "net_grad: bad argument");
    cmx::net_grad_host(O(op), prec, v, e, loss, g_wc, g_w0);
  });
}

int cmx_net_train(cmx_op op, const cmx_train_params* p, cmx_train_result* res) {
  return guarded([&] {
    require_init();
    CMX_REQUIRE(p && res, CMX_ERR_ARG, "net_train: bad argument");
    cmx::net_train(O(op), *p, *res);
  });
}

int cmx_field_write_scidac(const char* path, const int dims[4], int kind, int prec, const double* field,
                           const char* file_xml, const char* record_xml) {
  return guarded([&] {
    CMX_REQUIRE(dims, CMX_ERR_ARG, "field_write_scidac: null dims");
    cmx::scidac_write(path, dims, kind, prec, field, file_xml, record_xml);
  });
}

int cmx_field_read_scidac(const char* path, const int dims[4], int kind, double* field, char* file_xml,
                          int file_xml_cap, char* record_xml, int record_xml_cap) {
  return guarded([&] {
    CMX_REQUIRE(dims, CMX_ERR_ARG, "field_read_scidac: null dims");
    std::string fx, rx;
    cmx::scidac_read(path, dims, kind, field, file_xml ? &fx : nullptr, record_xml ? &rx : nullptr);
    auto give = [](const std::string& v, char* out, int cap) {
      if (!out || cap <= 0) return;
      const size_t n = std::min(v.size(), (size_t)cap - 1);
      std::memcpy(out, v.data(), n);
      out[n] = 0;
    };
    give(fx, file_xml, file_xml_cap);
    give(rx, record_xml, record_xml_cap);
  });
}

int cmx_mg_params_default(cmx_mg_params* p) {
  return guarded([&] {
    CMX_REQUIRE(p, CMX_ERR_ARG, "mg_params_default: null argument");
    std::memset(p, 0, sizeof(*p));
    for (int mu = 0; mu < 4; ++mu) p->blocking[mu] = 4;
    p->nvec = 16;
    p->nu_pre = 2;
    p->nu_post = 2;
    p->omega = 1.0;
    p->coarse_tol = 1e-4;
    p->coarse_maxit = 12;
    p->coarse_nkrylov = 12;
    p->setup_maxit = 500;
    p->seed = 1;
    p->prec = CMX_SINGLE;
  });
}

int cmx_mg_setup(cmx_op h, const cmx_mg_params* p, cmx_mg_result* res) {
  return guarded([&] {
    require_init();
    CMX_REQUIRE(p && res, CMX_ERR_ARG, "mg_setup: null argument");
    cmx::mg_setup(O(h), *p, *res);
  });
}

int cmx_mg_update(cmx_op h, const cmx_mg_params* p) {
  return guarded([&] {
    require_init();
    CMX_REQUIRE(p, CMX_ERR_ARG, "mg_update: null argument");
    cmx::mg_update(O(h), *p);
  });
}

int cmx_mg_get_nullvecs(cmx_op h, double* out) {
  return guarded([&] {
    require_init();
    CMX_REQUIRE(out, CMX_ERR_ARG, "mg_get_nullvecs: null argument");
    cmx::mg_get_nullvecs(O(h), out);
  });
}

int cmx_mg_get_coarse_op(cmx_op h, double* out) {
  return guarded([&] {
    require_init();
    CMX_REQUIRE(out, CMX_ERR_ARG, "mg_get_coarse_op: null argument");
    cmx::mg_get_coarse_op(O(h), out);
  });
}

int cmx_mg_restrict(cmx_op h, const double* full_in, double* coarse_out) {
  return guarded([&] {
    require_init();
    CMX_REQUIRE(full_in && coarse_out, CMX_ERR_ARG, "mg_restrict: null argument");
    cmx::mg_restrict_host(O(h), full_in, coarse_out);
  });
}

int cmx_mg_prolong(cmx_op h, const double* coarse_in, double* full_out) {
  return guarded([&] {
    require_init();
    CMX_REQUIRE(coarse_in && full_out, CMX_ERR_ARG, "mg_prolong: null argument");
    cmx::mg_prolong_host(O(h), coarse_in, full_out);
  });
}

int cmx_mg_coarse_apply(cmx_op h, const double* in, double* out) {
  return guarded([&] {
    require_init();
    CMX_REQUIRE(in && out, CMX_ERR_ARG, "mg_coarse_apply: null argument");
    cmx::mg_coarse_apply_host(O(h), in, out);
  });
}

}  // extern "C"
