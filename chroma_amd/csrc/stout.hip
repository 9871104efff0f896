// Stout smearing on the device, in FP64: Stouting::smear_links (stout_utils.cc:936-962) with the
// parameters of StoutLinkSmearing (stout_link_smearing.cc:62-120): a set of smeared directions and
// rho(mu,nu) = rho for mu != nu both smeared, else 0.  Links of an unsmeared direction are copied, and
// only smeared directions contribute staples; the staple sum C and Q of getQsandCs (:105-180):
//   C_mu(x) = rho sum_{nu != mu} [ U_nu(x) U_mu(x+nu) U_nu^dag(x+mu) + U_nu^dag(x-nu) U_mu(x-nu) U_nu(x-nu+mu) ],
//   Omega = C U_mu^dag(x),  Q = (i/2) [ (Omega^dag - Omega) - tr(Omega^dag - Omega)/3 ],  U'_mu(x) = exp(iQ) U_mu(x).
// The reference evaluates exp(iQ) by Cayley-Hamilton (getFsAndBs, :383-874); here it is a degree-14
// Taylor polynomial with scaling and squaring (Moler & Van Loan, "Nineteen dubious ways to compute the
// exponential of a matrix", method 3), exact to rounding with no small-eigenvalue special case.  The
// product is re-projected onto SU(3) (rows 0 and 1 orthonormalised, row 2 = conj(row0 x row1)), which
// moves it only at rounding level and keeps many steps from drifting off the group.
//
// Layout: the host gauge layout ([mu][site][3][3] complex, sites checkerboard-major, x fastest within a
// checkerboard) is transposed on the device into [mu][element][site], so that each of a lane's loads and
// stores is one 16 B element and a wave's 64 lanes touch 1 KiB of consecutive sites; the steps run in
// that layout and the result is transposed back.  One lane per (mu, site); each link is read by 13 lanes
// (its own, and as a staple leg of 12 others), mostly from L2.
#include <cmath>
#include <string>
#include <utility>

#include "op.h"

namespace cmx {

namespace {

struct M3 {
  double2 e[9];
};

__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {  // c + a b
  return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
__device__ __forceinline__ double2 cfmac(double2 a, double2 b, double2 c) {  // c + a conj(b)
  return make_double2(fma(a.x, b.x, fma(a.y, b.y, c.x)), fma(a.y, b.x, fma(-a.x, b.y, c.y)));
}
__device__ __forceinline__ double2 cfmca(double2 a, double2 b, double2 c) {  // c + conj(a) b
  return make_double2(fma(a.x, b.x, fma(a.y, b.y, c.x)), fma(a.x, b.y, fma(-a.y, b.x, c.y)));
}

__device__ __forceinline__ M3 mul(const M3& a, const M3& b) {  // a b
  M3 c;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 t = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 3; ++k) t = cfma(a.e[3 * i + k], b.e[3 * k + j], t);
      c.e[3 * i + j] = t;
    }
  return c;
}
__device__ __forceinline__ M3 mul_nd(const M3& a, const M3& b) {  // a b^dag
  M3 c;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 t = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 3; ++k) t = cfmac(a.e[3 * i + k], b.e[3 * j + k], t);
      c.e[3 * i + j] = t;
    }
  return c;
}
__device__ __forceinline__ M3 mul_dn(const M3& a, const M3& b) {  // a^dag b
  M3 c;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 t = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 3; ++k) t = cfmca(a.e[3 * k + i], b.e[3 * k + j], t);
      c.e[3 * i + j] = t;
    }
  return c;
}

struct StoutGeom {
  int L[4];
  long V, Vh;
};

__device__ __forceinline__ long site_of(const StoutGeom& g, const int (&x)[4]) {
  const int cb = (x[0] + x[1] + x[2] + x[3]) & 1;
  return cb * g.Vh + ((((long)x[3] * g.L[2] + x[2]) * g.L[1] + x[1]) * g.L[0] + x[0]) / 2;
}

// [mu][element][site] layout on the device
__device__ __forceinline__ M3 load(const double2* __restrict__ u, const StoutGeom& g, int mu, long s) {
  const double2* p = u + (size_t)mu * 9 * g.V + s;
  M3 m;
#pragma unroll
  for (int k = 0; k < 9; ++k) m.e[k] = p[(size_t)k * g.V];
  return m;
}

// host layout [mu][site][9] <-> device layout [mu][9][site], one lane per link
__global__ void __launch_bounds__(256) k_links_to_soa(long V, const double2* __restrict__ in, double2* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * V) return;
  const long mu = t / V, s = t - mu * V;
#pragma unroll
  for (int k = 0; k < 9; ++k) out[(mu * 9 + k) * V + s] = in[t * 9 + k];
}
__global__ void __launch_bounds__(256) k_links_from_soa(long V, const double2* __restrict__ in, double2* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * V) return;
  const long mu = t / V, s = t - mu * V;
#pragma unroll
  for (int k = 0; k < 9; ++k) out[t * 9 + k] = in[(mu * 9 + k) * V + s];
}

// exp(X) for a 3x3 X with |X|_F = O(1): X 2^-s with |X 2^-s|_F <= 1/4, Horner form of the degree-14
// Taylor polynomial (truncation below 1e-21 relative), then s squarings
__device__ M3 expm(const M3& X) {
  double nrm = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) nrm += X.e[k].x * X.e[k].x + X.e[k].y * X.e[k].y;
  nrm = sqrt(nrm);
  int s = 0;
  while (nrm > 0.25 && s < 60) {
    nrm *= 0.5;
    ++s;
  }
  const double sc = ldexp(1.0, -s);
  M3 Y;
#pragma unroll
  for (int k = 0; k < 9; ++k) Y.e[k] = make_double2(X.e[k].x * sc, X.e[k].y * sc);
  // E = I + Y/1 (I + Y/2 (I + ... (I + Y/14)))
  M3 E;
#pragma unroll
  for (int k = 0; k < 9; ++k) E.e[k] = make_double2(k % 4 == 0 ? 1.0 : 0.0, 0.0);
  for (int n = 14; n >= 1; --n) {
    M3 T = mul(Y, E);
    const double f = 1.0 / n;
#pragma unroll
    for (int k = 0; k < 9; ++k) E.e[k] = make_double2(T.e[k].x * f + (k % 4 == 0 ? 1.0 : 0.0), T.e[k].y * f);
  }
  for (int k = 0; k < s; ++k) E = mul(E, E);
  return E;
}

__device__ void reunit(M3& U) {
  double n0 = 0;
#pragma unroll
  for (int j = 0; j < 3; ++j) n0 += U.e[j].x * U.e[j].x + U.e[j].y * U.e[j].y;
  n0 = 1.0 / sqrt(n0);
#pragma unroll
  for (int j = 0; j < 3; ++j) U.e[j] = make_double2(U.e[j].x * n0, U.e[j].y * n0);
  double2 d = make_double2(0, 0);
#pragma unroll
  for (int j = 0; j < 3; ++j) d = cfmca(U.e[j], U.e[3 + j], d);  // <row0, row1>
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double2 p = cfma(d, U.e[j], make_double2(0, 0));
    U.e[3 + j] = make_double2(U.e[3 + j].x - p.x, U.e[3 + j].y - p.y);
  }
  double n1 = 0;
#pragma unroll
  for (int j = 0; j < 3; ++j) n1 += U.e[3 + j].x * U.e[3 + j].x + U.e[3 + j].y * U.e[3 + j].y;
  n1 = 1.0 / sqrt(n1);
#pragma unroll
  for (int j = 0; j < 3; ++j) U.e[3 + j] = make_double2(U.e[3 + j].x * n1, U.e[3 + j].y * n1);
#pragma unroll
  for (int k = 0; k < 3; ++k) {  // row2_k = conj(row0_a row1_b - row0_b row1_a)
    const int a = (k + 1) % 3, b = (k + 2) % 3;
    const double2 x = U.e[a], y = U.e[3 + b], z = U.e[b], w = U.e[3 + a];
    const double re = x.x * y.x - x.y * y.y - (z.x * w.x - z.y * w.y);
    const double im = x.x * y.y + x.y * y.x - (z.x * w.y + z.y * w.x);
    U.e[6 + k] = make_double2(re, -im);
  }
}

__global__ void __launch_bounds__(256) k_stout_step(StoutGeom g, const double2* __restrict__ cur,
                                                    double2* __restrict__ next, double rho, int dirs) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * g.V) return;
  const int mu = (int)(t / g.V);
  const long s = t - (long)mu * g.V;
  if (!(dirs >> mu & 1)) {  // next[mu] = current[mu]
    const double2* p = cur + (size_t)mu * 9 * g.V + s;
    double2* o = next + (size_t)mu * 9 * g.V + s;
#pragma unroll
    for (int k = 0; k < 9; ++k) o[(size_t)k * g.V] = p[(size_t)k * g.V];
    return;
  }
  int x[4];
  {
    const int cb = s >= g.Vh ? 1 : 0;
    long r = s - cb * g.Vh;
    const int Lxh = g.L[0] / 2;
    const int xh = (int)(r % Lxh);
    r /= Lxh;
    x[1] = (int)(r % g.L[1]);
    r /= g.L[1];
    x[2] = (int)(r % g.L[2]);
    x[3] = (int)(r / g.L[2]);
    x[0] = 2 * xh + ((cb + x[1] + x[2] + x[3]) & 1);
  }
  M3 C;
#pragma unroll
  for (int k = 0; k < 9; ++k) C.e[k] = make_double2(0, 0);
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu || !(dirs >> nu & 1)) continue;
    int y[4] = {x[0], x[1], x[2], x[3]};
    y[mu] = x[mu] + 1 == g.L[mu] ? 0 : x[mu] + 1;
    const long xpm = site_of(g, y);
    y[nu] = x[nu] == 0 ? g.L[nu] - 1 : x[nu] - 1;
    const long xmnpm = site_of(g, y);
    y[mu] = x[mu];
    const long xmn = site_of(g, y);
    y[nu] = x[nu] + 1 == g.L[nu] ? 0 : x[nu] + 1;
    const long xpn = site_of(g, y);
    // forward staple U_nu(x) U_mu(x+nu) U_nu^dag(x+mu)
    const M3 f = mul_nd(mul(load(cur, g, nu, s), load(cur, g, mu, xpn)), load(cur, g, nu, xpm));
    // backward staple U_nu^dag(x-nu) U_mu(x-nu) U_nu(x-nu+mu)
    const M3 b = mul(mul_dn(load(cur, g, nu, xmn), load(cur, g, mu, xmn)), load(cur, g, nu, xmnpm));
#pragma unroll
    for (int k = 0; k < 9; ++k) C.e[k] = make_double2(C.e[k].x + f.e[k].x + b.e[k].x, C.e[k].y + f.e[k].y + b.e[k].y);
  }
  const M3 U = load(cur, g, mu, s);
  const M3 Om = mul_nd(C, U);
  // i Q = i (i/2) T = -T/2 with T = (Omega^dag - Omega) - tr/3, Omega scaled by rho
  M3 X;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double2 a = Om.e[3 * j + i], c = Om.e[3 * i + j];
      X.e[3 * i + j] = make_double2(-0.5 * rho * (a.x - c.x), -0.5 * rho * (-a.y - c.y));
    }
  const double2 tr = make_double2((X.e[0].x + X.e[4].x + X.e[8].x) / 3.0, (X.e[0].y + X.e[4].y + X.e[8].y) / 3.0);
#pragma unroll
  for (int i = 0; i < 3; ++i) X.e[4 * i] = make_double2(X.e[4 * i].x - tr.x, X.e[4 * i].y - tr.y);
  M3 N = mul(expm(X), U);
  reunit(N);
  double2* o = next + (size_t)mu * 9 * g.V + s;
#pragma unroll
  for (int k = 0; k < 9; ++k) o[(size_t)k * g.V] = N.e[k];
}

}  // namespace

void stout_smear(const int dims[4], double* links, int nstep, double rho, int dirs) {
  CMX_REQUIRE(dims && links, CMX_ERR_ARG, "gauge_stout_smear: null array");
  CMX_REQUIRE(dirs >= 0 && dirs <= 15, CMX_ERR_ARG, "gauge_stout_smear: smeared directions are a 4-bit mask");
  CMX_REQUIRE(nstep >= 0, CMX_ERR_ARG, "gauge_stout_smear: nstep must be >= 0");
  CMX_REQUIRE(std::isfinite(rho), CMX_ERR_ARG, "gauge_stout_smear: rho must be finite");
  StoutGeom g;
  g.V = 1;
  for (int m = 0; m < 4; ++m) {
    CMX_REQUIRE(dims[m] > 0 && dims[m] % 2 == 0, CMX_ERR_ARG, "gauge_stout_smear: extents must be even and positive");
    g.L[m] = dims[m];
    g.V *= dims[m];
  }
  g.Vh = g.V / 2;
  CMX_REQUIRE(4 * g.V <= (long)INT32_MAX * 256, CMX_ERR_ARG, "gauge_stout_smear: lattice too large for one grid");
  if (nstep == 0) return;
  const size_t bytes = sizeof(double) * 18 * 4 * (size_t)g.V;
  DevBuf a(bytes), b(bytes);
  CMX_HIP(hipMemcpy(a.p, links, bytes, hipMemcpyHostToDevice));
  const dim3 grid((unsigned)((4 * g.V + 255) / 256));
  hipLaunchKernelGGL(k_links_to_soa, grid, dim3(256), 0, 0, g.V, a.as<double2>(), b.as<double2>());
  CMX_HIP(hipGetLastError());
  double2* cur = b.as<double2>();
  double2* nxt = a.as<double2>();
  for (int k = 0; k < nstep; ++k) {
    hipLaunchKernelGGL(k_stout_step, grid, dim3(256), 0, 0, g, cur, nxt, rho, dirs);
    CMX_HIP(hipGetLastError());
    std::swap(cur, nxt);
  }
  hipLaunchKernelGGL(k_links_from_soa, grid, dim3(256), 0, 0, g.V, cur, nxt);
  CMX_HIP(hipGetLastError());
  CMX_HIP(hipMemcpy(links, nxt, bytes, hipMemcpyDeviceToHost));
}

}  // namespace cmx
