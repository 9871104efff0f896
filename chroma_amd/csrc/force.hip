// Clover force terms on the device, in FP64 (CloverTermBase, clover_term_base_w.h:320-870), and the
// derivative of the even-odd clover operator (EvenOddPrecLogDetLinearOperator::deriv,
// eoprec_logdet_linop.h:114-157, with the pieces of eoprec_clover_linop_w.cc:192-260).
//
// For every plane mu < nu the insertion matrix Lambda lives on the sites of one checkerboard cb:
//   deriv:     Lambda(x) = -(1/8) c_munu traceSpin(outerProduct(gamma_mu gamma_nu psi, chi))(x)
//   derivTrLn: Lambda(x) = -(1/8) c_munu traceSpin(gamma_mu gamma_nu A^-1(x))   (triacntr, the inverted term)
// with c_munu the 1/2-scaled clover coefficient of the plane (clovCoeffT when nu is time).  The link
// derivative is the sum over the four clover leaves of the plane with Lambda inserted at each corner of
// cb (deriv_loops): per output link U_mu(x) two terms at x and two terms of the leaf below it (at
// x - nu), and per U_nu(x) two at x and two of the leaf to its left (at x - mu).  One lane evaluates
// all of them for its site, so each plane is one launch that writes ds_mu(x) and ds_nu(x) of its own
// site only: no atomics, deterministic.  ds is [4][V][3][3] complex in the host site order (the layout
// of the C-ABI), accumulated as ds[mu] += ..., ds[nu] -= ... .
#include <cstring>

#include "op.h"
#include "su3_dev.h"

namespace cmx {

namespace {

// gamma_mu gamma_nu of the DeGrand-Rossi basis as a signed permutation: row s has its one entry ph[s]
// in column p[s]; it maps each chiral half (spins 0,1 / 2,3) onto itself
struct Gmn {
  int p[4];
  double re[4], im[4];
};

struct ForcePlanes {
  Gmn G[6];
  double f[6];  // -(1/8) c_munu
};

__device__ __forceinline__ int plane_index(int mu, int nu) { return mu == 0 ? nu - 1 : mu == 1 ? nu + 1 : 5; }

__device__ __forceinline__ M3 load_m3(const double2* b, size_t stride, size_t i) {
  M3 m;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const double2 v = b[(size_t)k * stride + i];
    m.a[2 * k] = v.x;
    m.a[2 * k + 1] = v.y;
  }
  return m;
}
__device__ __forceinline__ M3 madd(const M3& A, const M3& B) {
  M3 C;
#pragma unroll
  for (int k = 0; k < 18; ++k) C.a[k] = A.a[k] + B.a[k];
  return C;
}

// Lambda for the six planes at site i of cb, from chi and psi (deriv) or the inverted clover (derivTrLn).
// Lam: [plane][9][Vh] double2
template <bool TRLN>
__global__ void __launch_bounds__(256) k_force_lambda(Geom g, ForcePlanes P, VRef chi, VRef psi,
                                                      const double2* __restrict__ clov, double2* __restrict__ Lam) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
  double c[24], q[24];
  double Are[12][12], Aim[12][12];
  if constexpr (!TRLN) {
    SpinorIO<P64>::load(chi.v, chi.nrm, chi.stride, i, c);
    SpinorIO<P64>::load(psi.v, psi.nrm, psi.stride, i, q);
  } else {
    // the two chiral 6x6 blocks {diag[6], lower offd[15] complex} of the inverted term
#pragma unroll
    for (int r = 0; r < 12; ++r)
#pragma unroll
      for (int s = 0; s < 12; ++s) Are[r][s] = Aim[r][s] = 0;
    for (int b = 0; b < 2; ++b) {
      double t[36];
      for (int k = 0; k < 18; ++k) {
        const double2 v = clov[(size_t)(18 * b + k) * g.Vh + i];
        t[2 * k] = v.x;
        t[2 * k + 1] = v.y;
      }
      for (int r = 0; r < 6; ++r) Are[6 * b + r][6 * b + r] = t[r];
      for (int r = 1; r < 6; ++r)
        for (int s = 0; s < r; ++s) {
          const int e = r * (r - 1) / 2 + s;
          Are[6 * b + r][6 * b + s] = t[6 + 2 * e];
          Aim[6 * b + r][6 * b + s] = t[6 + 2 * e + 1];
          Are[6 * b + s][6 * b + r] = t[6 + 2 * e];
          Aim[6 * b + s][6 * b + r] = -t[6 + 2 * e + 1];
        }
    }
  }
  for (int pl = 0; pl < 6; ++pl) {
    const Gmn& G = P.G[pl];
    double2* L = Lam + (size_t)pl * 9 * g.Vh;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double re = 0, im = 0;
        for (int s = 0; s < 4; ++s) {
          const int t = G.p[s];
          double xr, xi;  // deriv: (Gamma psi)_{s,a} conj(chi_{s,b});  derivTrLn: Gamma_{s,t} A_{(t,a),(s,b)}
          if constexpr (!TRLN) {
            const double pr = q[2 * (3 * t + a)], pi = q[2 * (3 * t + a) + 1];
            const double gr = G.re[s] * pr - G.im[s] * pi, gi = G.re[s] * pi + G.im[s] * pr;
            const double cr = c[2 * (3 * s + b)], ci = c[2 * (3 * s + b) + 1];
            xr = gr * cr + gi * ci;
            xi = gi * cr - gr * ci;
          } else {
            const double ar = Are[3 * t + a][3 * s + b], ai = Aim[3 * t + a][3 * s + b];
            xr = G.re[s] * ar - G.im[s] * ai;
            xi = G.re[s] * ai + G.im[s] * ar;
          }
          re += xr;
          im += xi;
        }
        L[(size_t)(3 * a + b) * g.Vh + i] = make_double2(P.f[pl] * re, P.f[pl] * im);
      }
  }
}

// The pieces of deriv_loops evaluated at site y (coordinates), Lambda on the sites of parity cb.
// part = 0: the ds_mu terms at y; 1: the tmp_mu terms (shifted to y + nu); 2: the ds_nu terms;
// 3: the tmp_nu terms (shifted to y + mu)
struct LoopCtx {
  Geom g;
  const double2* g64[2];
  const double2* Lam;  // this plane, [9][Vh]
  int mu, nu, cb;
};

__device__ __forceinline__ M3 lam_at(const LoopCtx& C, const int (&x)[4]) {
  if (parity(x) != C.cb) {
    M3 z;
#pragma unroll
    for (int k = 0; k < 18; ++k) z.a[k] = 0;
    return z;
  }
  return load_m3(C.Lam, C.g.Vh, cb_index(C.g, x));
}

__device__ M3 loop_terms(const LoopCtx& C, const int (&y)[4], int part) {
  const Geom& g = C.g;
  const int mu = C.mu, nu = C.nu;
  int ym[4], yn[4], ymn[4];
  shiftc(g, y, mu, 1, ym);
  shiftc(g, y, nu, 1, yn);
  shiftc(g, ym, nu, 1, ymn);
  const M3 unu_fmu = get_link(g, C.g64, nu, ym);  // u_nu(y + mu)
  const M3 umu_fnu = get_link(g, C.g64, mu, yn);  // u_mu(y + nu)
  const M3 umu = get_link(g, C.g64, mu, y);
  const M3 unu = get_link(g, C.g64, nu, y);
  if (parity(y) == C.cb) {
    const M3 Lx = lam_at(C, y), Lmn = lam_at(C, ymn);
    switch (part) {
      case 0: {  // 1) u_nu(y+mu) L(y+mu+nu) ulc;  5) staple_for L(y)
        const M3 ulc = mul(dag(umu_fnu), dag(unu));
        return madd(mul(mul(unu_fmu, Lmn), ulc), mul(mul(unu_fmu, ulc), Lx));
      }
      case 1: {  // 4) lrc L(y) u_nu;  8) L(y+mu+nu) staple_back
        const M3 lrc = mul(dag(unu_fmu), dag(umu));
        const M3 llc = mul(dag(umu), unu);
        return madd(mul(mul(lrc, Lx), unu), mul(Lmn, mul(dag(unu_fmu), llc)));
      }
      case 2: {  // 4') u_mu(y+nu) L(y+mu+nu) lrc;  7') staple_left L(y)
        const M3 lrc = mul(dag(unu_fmu), dag(umu));
        return madd(mul(mul(umu_fnu, Lmn), lrc), mul(mul(umu_fnu, lrc), Lx));
      }
      default: {  // 2') ulc L(y) u_mu;  6') L(y+mu+nu) staple_right
        const M3 ulc = mul(dag(umu_fnu), dag(unu));
        return madd(mul(mul(ulc, Lx), umu), mul(Lmn, mul(ulc, umu)));
      }
    }
  } else {
    const M3 Lm = lam_at(C, ym), Ln = lam_at(C, yn);
    switch (part) {
      case 0: {  // 2) urc L(y+nu) u_nu^dag;  6) L(y+mu) staple_for
        const M3 urc = mul(unu_fmu, dag(umu_fnu));
        const M3 ulc = mul(dag(umu_fnu), dag(unu));
        return madd(mul(mul(urc, Ln), dag(unu)), mul(Lm, mul(unu_fmu, ulc)));
      }
      case 1: {  // 3) u_nu(y+mu)^dag L(y+mu) llc;  7) staple_back L(y+nu)
        const M3 llc = mul(dag(umu), unu);
        return madd(mul(mul(dag(unu_fmu), Lm), llc), mul(mul(dag(unu_fmu), llc), Ln));
      }
      case 2: {  // 3') urc^dag L(y+mu) u_mu^dag;  8') L(y+nu) staple_left
        const M3 urc = mul(unu_fmu, dag(umu_fnu));
        const M3 lrc = mul(dag(unu_fmu), dag(umu));
        return madd(mul(mul(dag(urc), Lm), dag(umu)), mul(Ln, mul(umu_fnu, lrc)));
      }
      default: {  // 1') u_mu(y+nu)^dag L(y+nu) llc^dag;  5') staple_right L(y+mu)
        const M3 llc = mul(dag(umu), unu);
        const M3 ulc = mul(dag(umu_fnu), dag(unu));
        return madd(mul(mul(dag(umu_fnu), Ln), dag(llc)), mul(mul(ulc, umu), Lm));
      }
    }
  }
}

// ds[mu](x) += ds_mu(x) - tmp_mu(x - nu);  ds[nu](x) -= ds_nu(x) - tmp_nu(x - mu), scaled by `scale`
__global__ void __launch_bounds__(128) k_force_loops(LoopCtx C, double scale, double2* __restrict__ ds) {
  const long V = 2L * C.g.Vh;
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  const int p = s >= C.g.Vh ? 1 : 0;
  int x[4];
  cb_coords(C.g, p, (int)(s - (long)p * C.g.Vh), x);
  int xmn[4], xmm[4];
  shiftc(C.g, x, C.nu, -1, xmn);
  shiftc(C.g, x, C.mu, -1, xmm);
  {
    const M3 a = loop_terms(C, x, 0), b = loop_terms(C, xmn, 1);
    double2* d = ds + ((size_t)C.mu * V + s) * 9;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double2 o = d[k];
      d[k] = make_double2(o.x + scale * (a.a[2 * k] - b.a[2 * k]), o.y + scale * (a.a[2 * k + 1] - b.a[2 * k + 1]));
    }
  }
  {
    const M3 a = loop_terms(C, x, 2), b = loop_terms(C, xmm, 3);
    double2* d = ds + ((size_t)C.nu * V + s) * 9;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double2 o = d[k];
      d[k] = make_double2(o.x - scale * (a.a[2 * k] - b.a[2 * k]), o.y - scale * (a.a[2 * k + 1] - b.a[2 * k + 1]));
    }
  }
}

// ds_full[mu][cb*Vh + i] += f * ds_cb[mu][i]  (the hopping-term derivative's per-checkerboard layout)
__global__ void __launch_bounds__(256) k_force_axpy(long Vh, int cb, double f, const double2* __restrict__ src,
                                                    double2* __restrict__ dst) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * Vh * 9) return;
  const long k = t % 9, site = (t / 9) % Vh, mu = t / (9 * Vh);
  const double2 a = src[t];
  double2& b = dst[(mu * 2 * Vh + (long)cb * Vh + site) * 9 + k];
  b = make_double2(b.x + f * a.x, b.y + f * a.y);
}

ForcePlanes force_planes(const cmx_op_params& par) {
  // DeGrand-Rossi gamma_1..gamma_4 as signed permutations
  static const int perm[4][4] = {{3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}, {2, 3, 0, 1}};
  static const double gre[4][4] = {{0, 0, 0, 0}, {-1, 1, 1, -1}, {0, 0, 0, 0}, {1, 1, 1, 1}};
  static const double gim[4][4] = {{1, 1, -1, -1}, {0, 0, 0, 0}, {1, -1, -1, 1}, {0, 0, 0, 0}};
  ForcePlanes P;
  int pl = 0;
  for (int mu = 0; mu < 3; ++mu)
    for (int nu = mu + 1; nu < 4; ++nu, ++pl) {
      Gmn& G = P.G[pl];
      for (int s = 0; s < 4; ++s) {  // (gamma_mu gamma_nu)[s][p] = gmu[s][k] gnu[k][p], k = perm_mu[s]
        const int k = perm[mu][s];
        G.p[s] = perm[nu][k];
        G.re[s] = gre[mu][s] * gre[nu][k] - gim[mu][s] * gim[nu][k];
        G.im[s] = gre[mu][s] * gim[nu][k] + gim[mu][s] * gre[nu][k];
      }
      const double c = 0.5 * (nu == 3 ? par.clov_coeff_t : par.clov_coeff_r);  // getCloverCoeff
      P.f[pl] = -c / 8.0;
    }
  return P;
}

void launch_loops(Op& op, const double2* Lam, int cb, double scale, double2* ds) {
  LoopCtx C;
  C.g = op.g;
  C.g64[0] = op.gd->g64[0].as<double2>();
  C.g64[1] = op.gd->g64[1].as<double2>();
  C.cb = cb;
  int pl = 0;
  for (int mu = 0; mu < 3; ++mu)
    for (int nu = mu + 1; nu < 4; ++nu, ++pl) {
      C.mu = mu;
      C.nu = nu;
      C.Lam = Lam + (size_t)pl * 9 * op.Vh;
      hipLaunchKernelGGL(k_force_loops, dim3((unsigned)((2 * op.Vh + 127) / 128)), dim3(128), 0, op.stream, C, scale, ds);
      CMX_HIP(hipGetLastError());
    }
}

}  // namespace

void clover_force(Op& op, bool trln, int cb, VRef chi, VRef psi, double scale, void* ds_full) {
  CMX_REQUIRE(op.clover(), CMX_ERR_ARG, "clover force: not a clover operator");
  CMX_REQUIRE(cb == 0 || cb == 1, CMX_ERR_ARG, "clover force: cb must be 0 or 1");
  CMX_REQUIRE(!trln || cb == 0, CMX_ERR_ARG, "clover force: derivTrLn of the inverted even-even term only");
  const ForcePlanes P = force_planes(op.par);
  DevBuf lam((size_t)6 * 9 * 16 * op.Vh);
  const dim3 grid((unsigned)((op.Vh + 255) / 256));
  if (trln)
    hipLaunchKernelGGL(k_force_lambda<true>, grid, dim3(256), 0, op.stream, op.g, P, chi, psi,
                       op.clovinv_even[P64].as<double2>(), lam.as<double2>());
  else
    hipLaunchKernelGGL(k_force_lambda<false>, grid, dim3(256), 0, op.stream, op.g, P, chi, psi, nullptr,
                       lam.as<double2>());
  CMX_HIP(hipGetLastError());
  launch_loops(op, lam.as<double2>(), cb, scale, static_cast<double2*>(ds_full));
  op.sync();  // the Lambda buffer is released on return
}

void force_axpy(Op& op, int cb, double f, const void* ds_cb, void* ds_full) {
  const long n = 4 * op.Vh * 9;
  hipLaunchKernelGGL(k_force_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, op.stream, op.Vh, cb, f,
                     static_cast<const double2*>(ds_cb), static_cast<double2*>(ds_full));
  CMX_HIP(hipGetLastError());
}

}  // namespace cmx
