// Set-up kernels: layout conversion, gauge packing, clover construction and LDL^dag inversion (FP64).
#include "setup.h"
#include "su3_dev.h"

namespace cmx {

namespace {

// ---- gauge packing ------------------------------------------------------------------------------------
// host order [mu][full site][18]  ->  g64[parity][mu][9 x double2][Vh]
__global__ void k_gauge_pack64(Geom g, const double* __restrict__ host, double2* g0, double2* g1) {
  const long V = 2L * g.Vh;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * V) return;
  const int mu = (int)(t / V);
  const long s = t % V;
  const int p = s >= g.Vh ? 1 : 0;
  const int i = (int)(s - (long)p * g.Vh);
  double2* dst = (p ? g1 : g0) + (size_t)mu * 9 * g.Vh;
  const double* src = host + t * 18;
#pragma unroll
  for (int k = 0; k < 9; ++k) dst[(size_t)k * g.Vh + i] = make_double2(src[2 * k], src[2 * k + 1]);
}

// g64 (per parity) -> sloppy storage at precision P, reconstruct RECON
template <int P, int RECON>
__global__ void k_gauge_pack_sloppy(Geom g, const double2* src, void* dst) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;  // over 4 * Vh
  if (t >= 4L * g.Vh) return;
  const int mu = (int)(t / g.Vh), i = (int)(t % g.Vh);
  const double2* s = src + (size_t)mu * 9 * g.Vh;
  double u[18];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    double2 v = s[(size_t)k * g.Vh + i];
    u[2 * k] = v.x;
    u[2 * k + 1] = v.y;
  }
  if constexpr (P == P64) {  // rows 0, 1 in FP64 (12-real; the third row is rebuilt on use)
    double2* d = reinterpret_cast<double2*>(dst) + (size_t)mu * 6 * g.Vh;
#pragma unroll
    for (int k = 0; k < 6; ++k) d[(size_t)k * g.Vh + i] = make_double2(u[2 * k], u[2 * k + 1]);
  } else if constexpr (P == P32 && RECON == 18) {
    float2* d = reinterpret_cast<float2*>(dst) + (size_t)mu * 9 * g.Vh;
#pragma unroll
    for (int k = 0; k < 9; ++k) d[(size_t)k * g.Vh + i] = make_float2((float)u[2 * k], (float)u[2 * k + 1]);
  } else if constexpr (P == P32 && RECON == 12) {
    float4* d = reinterpret_cast<float4*>(dst) + (size_t)mu * 3 * g.Vh;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      d[(size_t)k * g.Vh + i] = make_float4((float)u[4 * k], (float)u[4 * k + 1], (float)u[4 * k + 2], (float)u[4 * k + 3]);
  } else {
    // per direction: 16 B (reals 0-7) of every site, then 8 B (reals 8-11) of every site
    uint2* d = reinterpret_cast<uint2*>(dst) + (size_t)mu * 3 * g.Vh;
    H4 v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      v[k].h[0] = __floats2half2_rn((float)u[4 * k], (float)u[4 * k + 1]);
      v[k].h[1] = __floats2half2_rn((float)u[4 * k + 2], (float)u[4 * k + 3]);
    }
    reinterpret_cast<uint4*>(d)[i] = make_uint4(v[0].u.x, v[0].u.y, v[1].u.x, v[1].u.y);
    d[(size_t)2 * g.Vh + i] = v[2].u;
  }
}

// ---- field strength: mesField (mesfield.cc:30-78), F_{mu nu} = (1/8)(Q - Q^dag) with Q the sum of
// the four clover leaves listed in mesfield.cc:15-21, evaluated directly at each site.
// Output F[plane][9 x double2][V] over full site index s = cb*Vh + i.
__global__ void k_fieldstrength(Geom g, const double2* g0, const double2* g1, double2* F) {
  const long V = 2L * g.Vh;
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  const double2* const gg[2] = {g0, g1};
  const int cb = s >= g.Vh ? 1 : 0;
  int x[4];
  cb_coords(g, cb, (int)(s - (long)cb * g.Vh), x);
  int plane = 0;
  for (int mu = 0; mu < 3; ++mu)
    for (int nu = mu + 1; nu < 4; ++nu, ++plane) {
      int xpm[4], xpn[4], xmm[4], xmn[4], xmmpn[4], xmmmn[4], xmnpm[4];
      shiftc(g, x, mu, 1, xpm);
      shiftc(g, x, nu, 1, xpn);
      shiftc(g, x, mu, -1, xmm);
      shiftc(g, x, nu, -1, xmn);
      shiftc(g, xmm, nu, 1, xmmpn);
      shiftc(g, xmm, nu, -1, xmmmn);
      shiftc(g, xmn, mu, 1, xmnpm);
      // U_1 = u(x,mu) u(x+mu,nu) u^dag(x+nu,mu) u^dag(x,nu)
      M3 Q = mul(mul(get_link(g, gg, mu, x), get_link(g, gg, nu, xpm)),
                 mul(dag(get_link(g, gg, mu, xpn)), dag(get_link(g, gg, nu, x))));
      // U_2 = u(x,nu) u^dag(x-mu+nu,mu) u^dag(x-mu,nu) u(x-mu,mu)
      M3 T = mul(mul(get_link(g, gg, nu, x), dag(get_link(g, gg, mu, xmmpn))),
                 mul(dag(get_link(g, gg, nu, xmm)), get_link(g, gg, mu, xmm)));
#pragma unroll
      for (int k = 0; k < 18; ++k) Q.a[k] += T.a[k];
      // U_3 = u^dag(x-mu,mu) u^dag(x-mu-nu,nu) u(x-mu-nu,mu) u(x-nu,nu)
      T = mul(mul(dag(get_link(g, gg, mu, xmm)), dag(get_link(g, gg, nu, xmmmn))),
              mul(get_link(g, gg, mu, xmmmn), get_link(g, gg, nu, xmn)));
#pragma unroll
      for (int k = 0; k < 18; ++k) Q.a[k] += T.a[k];
      // U_4 = u^dag(x-nu,nu) u(x-nu,mu) u(x-nu+mu,nu) u^dag(x,mu)
      T = mul(mul(dag(get_link(g, gg, nu, xmn)), get_link(g, gg, mu, xmn)),
              mul(get_link(g, gg, nu, xmnpm), dag(get_link(g, gg, mu, x))));
#pragma unroll
      for (int k = 0; k < 18; ++k) Q.a[k] += T.a[k];
      double2* dst = F + (size_t)plane * 9 * V;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const double re = 0.125 * (Q.a[2 * (3 * i + j)] - Q.a[2 * (3 * j + i)]);
          const double im = 0.125 * (Q.a[2 * (3 * i + j) + 1] + Q.a[2 * (3 * j + i) + 1]);
          dst[(size_t)(3 * i + j) * V + s] = make_double2(re, im);
        }
    }
}

// ---- makeClovSiteLoop (clover_term_qdp_w.h:398-522) ------------------------------------------------
// clov[parity][36 x double2][Vh] in the order diag0[6], offd0[15], diag1[6], offd1[15] (reals)
__global__ void k_make_clover(Geom g, const double2* F, double coef0, double coef1, double coef2, double coef3,
                              double coef4, double coef5, double diag_mass, double2* c0, double2* c1) {
  const long V = 2L * g.Vh;
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  const double coef[6] = {coef0, coef1, coef2, coef3, coef4, coef5};
  auto f = [&](int k, int i, int j, double& re, double& im) {
    double2 v = F[((size_t)k * 9 + 3 * i + j) * V + s];
    re = v.x * coef[k];
    im = v.y * coef[k];
  };
  double diag[2][6], offd[2][15][2];
  for (int jj = 0; jj < 2; ++jj)
    for (int ii = 0; ii < 6; ++ii) diag[jj][ii] = diag_mass;
  for (int i = 0; i < 3; ++i) {
    double f5r, f5i, f0r, f0i;
    f(5, i, i, f5r, f5i);
    f(0, i, i, f0r, f0i);
    const double r0 = f5i - f0i;  // imag(f5 - f0)
    diag[0][i] += r0;
    diag[0][i + 3] -= r0;
    const double r1 = f5i + f0i;
    diag[1][i] -= r1;
    diag[1][i + 3] += r1;
  }
  for (int i = 1; i < 3; ++i)
    for (int j = 0; j < i; ++j) {
      const int eij = i * (i - 1) / 2 + j;
      const int etmp = (i + 3) * (i + 2) / 2 + j + 3;
      double f0r, f0i, f5r, f5i;
      f(0, i, j, f0r, f0i);
      f(5, i, j, f5r, f5i);
      // offd0 = i (f0 - f5); offd1 = i (f5 + f0); the lower blocks are the negatives
      offd[0][eij][0] = -(f0i - f5i);
      offd[0][eij][1] = f0r - f5r;
      offd[0][etmp][0] = -offd[0][eij][0];
      offd[0][etmp][1] = -offd[0][eij][1];
      offd[1][eij][0] = -(f5i + f0i);
      offd[1][eij][1] = f5r + f0r;
      offd[1][etmp][0] = -offd[1][eij][0];
      offd[1][etmp][1] = -offd[1][eij][1];
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const int eij = (i + 3) * (i + 2) / 2 + j;
      double f1r, f1i, f2r, f2i, f3r, f3i, f4r, f4i;
      f(1, i, j, f1r, f1i);
      f(2, i, j, f2r, f2i);
      f(3, i, j, f3r, f3i);
      f(4, i, j, f4r, f4i);
      // E_minus = i f2 + f4 ; B_minus = i f3 - f1
      const double emr = -f2i + f4r, emi = f2r + f4i;
      const double bmr = -f3i - f1r, bmi = f3r - f1i;
      offd[0][eij][0] = bmr - emr;
      offd[0][eij][1] = bmi - emi;
      offd[1][eij][0] = emr + bmr;
      offd[1][eij][1] = emi + bmi;
    }
  const int cb = s >= g.Vh ? 1 : 0;
  const int i = (int)(s - (long)cb * g.Vh);
  double2* dst = cb ? c1 : c0;
  double r[72];
  for (int b = 0; b < 2; ++b) {
    for (int k = 0; k < 6; ++k) r[36 * b + k] = diag[b][k];
    for (int k = 0; k < 15; ++k) {
      r[36 * b + 6 + 2 * k] = offd[b][k][0];
      r[36 * b + 6 + 2 * k + 1] = offd[b][k][1];
    }
  }
  for (int k = 0; k < 36; ++k) dst[(size_t)k * g.Vh + i] = make_double2(r[2 * k], r[2 * k + 1]);
}

// ---- LDL^dag inversion (LDagDLInvSiteLoop, clover_term_qdp_w.h:619-816), FP64 ---------------------
__global__ void k_clover_invert(Geom g, const double2* src, double2* dst, double* logdet, double* nneg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Vh) return;
  double site_log = 0;
  int site_neg = 0;
  for (int block = 0; block < 2; ++block) {
    double inv_d[6], diag_g[6];
    double ofr[15], ofi[15], vr[6], vi[6];
    double r[36];
    for (int k = 0; k < 18; ++k) {
      double2 v = src[(size_t)(18 * block + k) * g.Vh + i];
      r[2 * k] = v.x;
      r[2 * k + 1] = v.y;
    }
    for (int k = 0; k < 6; ++k) inv_d[k] = r[k];
    for (int k = 0; k < 15; ++k) {
      ofr[k] = r[6 + 2 * k];
      ofi[k] = r[6 + 2 * k + 1];
    }
    for (int j = 0; j < 6; ++j) {
      for (int k = 0; k < j; ++k) {  // v[k] = A_kk * conj(A(j,k))
        const int e = j * (j - 1) / 2 + k;
        vr[k] = inv_d[k] * ofr[e];
        vi[k] = -inv_d[k] * ofi[e];
      }
      vr[j] = inv_d[j];
      vi[j] = 0;
      for (int k = 0; k < j; ++k) {  // v[j] -= A(j,k) v[k]
        const int e = j * (j - 1) / 2 + k;
        vr[j] -= ofr[e] * vr[k] - ofi[e] * vi[k];
        vi[j] -= ofr[e] * vi[k] + ofi[e] * vr[k];
      }
      inv_d[j] = vr[j];
      for (int k = j + 1; k < 6; ++k) {
        const int ekj = k * (k - 1) / 2 + j;
        double ar = ofr[ekj], ai = ofi[ekj];
        for (int l = 0; l < j; ++l) {
          const int ekl = k * (k - 1) / 2 + l;
          ar -= ofr[ekl] * vr[l] - ofi[ekl] * vi[l];
          ai -= ofr[ekl] * vi[l] + ofi[ekl] * vr[l];
        }
        // /= v[j] (complex division)
        const double dd = vr[j] * vr[j] + vi[j] * vi[j];
        ofr[ekj] = (ar * vr[j] + ai * vi[j]) / dd;
        ofi[ekj] = (ai * vr[j] - ar * vi[j]) / dd;
      }
    }
    for (int k = 0; k < 6; ++k) {
      diag_g[k] = 1.0 / inv_d[k];
      site_log += log(fabs(inv_d[k]));
      if (inv_d[k] < 0) site_neg++;
    }
    for (int k = 0; k < 6; ++k) {
      for (int q = 0; q < k; ++q) vr[q] = vi[q] = 0;
      vr[k] = diag_g[k];
      vi[k] = 0;
      for (int q = k + 1; q < 6; ++q) {
        double ar = 0, ai = 0;
        for (int j = k; j < q; ++j) {
          const int e = q * (q - 1) / 2 + j;
          // v[q] -= offd(q,j) * d_j * v[j]
          const double tr = ofr[e] * inv_d[j], ti = ofi[e] * inv_d[j];
          ar -= tr * vr[j] - ti * vi[j];
          ai -= tr * vi[j] + ti * vr[j];
        }
        vr[q] = ar * diag_g[q];
        vi[q] = ai * diag_g[q];
      }
      for (int q = 4; q >= k; --q)
        for (int j = q + 1; j < 6; ++j) {
          const int e = j * (j - 1) / 2 + q;
          // v[q] -= conj(offd(j,q)) v[j]
          vr[q] -= ofr[e] * vr[j] + ofi[e] * vi[j];
          vi[q] -= ofr[e] * vi[j] - ofi[e] * vr[j];
        }
      inv_d[k] = vr[k];
      for (int q = k + 1; q < 6; ++q) {
        const int e = q * (q - 1) / 2 + k;
        ofr[e] = vr[q];
        ofi[e] = vi[q];
      }
    }
    for (int k = 0; k < 6; ++k) r[k] = inv_d[k];
    for (int k = 0; k < 15; ++k) {
      r[6 + 2 * k] = ofr[k];
      r[6 + 2 * k + 1] = ofi[k];
    }
    for (int k = 0; k < 18; ++k) dst[(size_t)(18 * block + k) * g.Vh + i] = make_double2(r[2 * k], r[2 * k + 1]);
  }
  logdet[i] = site_log;
  nneg[i] = site_neg ? 1.0 : 0.0;
}

// FP64 clover -> sloppy storage
template <int P> __global__ void k_clover_convert(int Vh, const double2* src, void* dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Vh) return;
  double r[72];
  for (int k = 0; k < 36; ++k) {
    double2 v = src[(size_t)k * Vh + i];
    r[2 * k] = v.x;
    r[2 * k + 1] = v.y;
  }
  if constexpr (P == P32) {
    float4* d = reinterpret_cast<float4*>(dst);
    for (int k = 0; k < 18; ++k)
      d[(size_t)k * Vh + i] = make_float4((float)r[4 * k], (float)r[4 * k + 1], (float)r[4 * k + 2], (float)r[4 * k + 3]);
  } else {
    // per chiral block b: reals 0-31 in four 16 B chunks (chunk 4b+k of every site), reals 32-35 in
    // one 8 B chunk after the eight 16 B chunks; then the norm
    uint2* d = reinterpret_cast<uint2*>(dst);
    uint4* d16 = reinterpret_cast<uint4*>(dst);
    double m = 0;
    for (int k = 0; k < 72; ++k) m = fmax(m, fabs(r[k]));
    const double inv = m > 0 ? 1.0 / m : 0.0;
    for (int b = 0; b < 2; ++b) {
      H4 v[9];
      for (int k = 0; k < 9; ++k) {
        const double* q = r + 36 * b + 4 * k;
        v[k].h[0] = __floats2half2_rn((float)(q[0] * inv), (float)(q[1] * inv));
        v[k].h[1] = __floats2half2_rn((float)(q[2] * inv), (float)(q[3] * inv));
      }
      for (int k = 0; k < 4; ++k)
        d16[(size_t)(4 * b + k) * Vh + i] = make_uint4(v[2 * k].u.x, v[2 * k].u.y, v[2 * k + 1].u.x, v[2 * k + 1].u.y);
      d[(size_t)(16 + b) * Vh + i] = v[8].u;
    }
    reinterpret_cast<float*>(d + (size_t)18 * Vh)[i] = (float)m;
  }
}

// clover SoA (FP64) -> host PrimitiveClovTriang order (72 doubles per site) for one checkerboard
__global__ void k_clover_unpack(int Vh, const double2* src, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Vh) return;
  // host order: diag[2][6], offd[2][15][2];  device order: per block {diag[6], offd[15][2]}
  for (int b = 0; b < 2; ++b) {
    for (int k = 0; k < 3; ++k) {
      double2 v = src[(size_t)(18 * b + k) * Vh + i];
      out[(size_t)i * 72 + 6 * b + 2 * k] = v.x;
      out[(size_t)i * 72 + 6 * b + 2 * k + 1] = v.y;
    }
    for (int k = 3; k < 18; ++k) {
      double2 v = src[(size_t)(18 * b + k) * Vh + i];
      out[(size_t)i * 72 + 12 + 30 * b + 2 * (k - 3)] = v.x;
      out[(size_t)i * 72 + 12 + 30 * b + 2 * (k - 3) + 1] = v.y;
    }
  }
}

// ---- fermion layout conversion -------------------------------------------------------------------------
template <int P> __global__ void k_spinor_pack(int n, const double* __restrict__ host, void* v, float* nrm, int stride) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  typename SpinorIO<P>::R s[24];
#pragma unroll
  for (int k = 0; k < 24; ++k) s[k] = (typename SpinorIO<P>::R)host[(size_t)i * 24 + k];
  SpinorIO<P>::store(v, nrm, stride, i, s);
}
// uniform noise in [-1/2, 1/2) from a hash of (seed, site, component): input of the launch-shape tuner
template <int P> __global__ void k_spinor_noise(int n, unsigned seed, void* v, float* nrm, int stride) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  typename SpinorIO<P>::R s[24];
#pragma unroll
  for (int k = 0; k < 24; ++k) {
    unsigned x = (unsigned)i * 24u + (unsigned)k + seed * 0x9E3779B9u;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    s[k] = (typename SpinorIO<P>::R)((float)(x >> 8) * (1.0f / 16777216.0f) - 0.5f);
  }
  SpinorIO<P>::store(v, nrm, stride, i, s);
}
template <int P> __global__ void k_spinor_unpack(int n, const void* v, const float* nrm, int stride, double* __restrict__ host) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  typename SpinorIO<P>::R s[24];
  SpinorIO<P>::load(v, nrm, stride, i, s);
#pragma unroll
  for (int k = 0; k < 24; ++k) host[(size_t)i * 24 + k] = (double)s[k];
}

// streaming copy (bandwidth reference): float4 in -> float4 out
__global__ void __launch_bounds__(256) k_stream_copy(long n, const float4* __restrict__ a, float4* __restrict__ b) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// five read streams, one write stream (the read/write mix of the hopping term), 16 B per lane each
__global__ void __launch_bounds__(256) k_stream_mix(long n, const uint4* __restrict__ a0, const uint4* __restrict__ a1,
                                                    const uint4* __restrict__ a2, const uint4* __restrict__ a3,
                                                    const uint4* __restrict__ a4, uint4* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const uint4 v0 = a0[i], v1 = a1[i], v2 = a2[i], v3 = a3[i], v4 = a4[i];
    out[i] = make_uint4(v0.x ^ v1.x ^ v2.x ^ v3.x ^ v4.x, v0.y ^ v1.y ^ v2.y ^ v3.y ^ v4.y,
                        v0.z ^ v1.z ^ v2.z ^ v3.z ^ v4.z, v0.w ^ v1.w ^ v2.w ^ v3.w ^ v4.w);
  }
}

// per-site double array -> per-block partial sums
__global__ void __launch_bounds__(256) k_sum_partials(int n, const double* a, double* partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {i < n ? a[i] : 0.0};
  block_sum<1>(v);
  if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

inline dim3 grid_for(long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

void gauge_pack64(const Geom& g, const double* dev_host_order, double2* g0, double2* g1, hipStream_t s) {
  hipLaunchKernelGGL(k_gauge_pack64, grid_for(8L * g.Vh), dim3(256), 0, s, g, dev_host_order, g0, g1);
}

void gauge_pack_sloppy(const Geom& g, int prec, int recon, const double2* src, void* dst, hipStream_t s) {
  const dim3 grid = grid_for(4L * g.Vh);
  if (prec == P32 && recon == 18)
    hipLaunchKernelGGL((k_gauge_pack_sloppy<P32, 18>), grid, dim3(256), 0, s, g, src, dst);
  else if (prec == P32 && recon == 12)
    hipLaunchKernelGGL((k_gauge_pack_sloppy<P32, 12>), grid, dim3(256), 0, s, g, src, dst);
  else if (prec == P16 && recon == 12)
    hipLaunchKernelGGL((k_gauge_pack_sloppy<P16, 12>), grid, dim3(256), 0, s, g, src, dst);
  else if (prec == P64 && recon == 12)
    hipLaunchKernelGGL((k_gauge_pack_sloppy<P64, 12>), grid, dim3(256), 0, s, g, src, dst);
  else
    throw Error(CMX_ERR_UNSUPPORTED, "gauge: unsupported precision/reconstruct combination");
}

void clover_build(const Geom& g, const double2* g0, const double2* g1, double2* Fscratch, const double coef[6],
                  double diag_mass, double2* c0, double2* c1, hipStream_t s) {
  hipLaunchKernelGGL(k_fieldstrength, grid_for(2L * g.Vh), dim3(256), 0, s, g, g0, g1, Fscratch);
  hipLaunchKernelGGL(k_make_clover, grid_for(2L * g.Vh), dim3(256), 0, s, g, (const double2*)Fscratch, coef[0], coef[1],
                     coef[2], coef[3], coef[4], coef[5], diag_mass, c0, c1);
}

void clover_invert(const Geom& g, const double2* src, double2* dst, double* logdet, double* nneg, hipStream_t s) {
  hipLaunchKernelGGL(k_clover_invert, dim3((unsigned)((g.Vh + 127) / 128)), dim3(128), 0, s, g, src, dst, logdet, nneg);
}

void clover_convert(int Vh, int prec, const double2* src, void* dst, hipStream_t s) {
  if (prec == P32)
    hipLaunchKernelGGL(k_clover_convert<P32>, grid_for(Vh), dim3(256), 0, s, Vh, src, dst);
  else if (prec == P16)
    hipLaunchKernelGGL(k_clover_convert<P16>, grid_for(Vh), dim3(256), 0, s, Vh, src, dst);
  else
    throw Error(CMX_ERR_ARG, "clover_convert: bad precision");
}

void clover_unpack(int Vh, const double2* src, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_clover_unpack, grid_for(Vh), dim3(256), 0, s, Vh, src, out);
}

void spinor_pack(int prec, int n, const double* host_dev, void* v, float* nrm, int stride, hipStream_t s) {
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_spinor_pack<P64>, grid_for(n), dim3(256), 0, s, n, host_dev, v, nrm, stride); break;
    case P32: hipLaunchKernelGGL(k_spinor_pack<P32>, grid_for(n), dim3(256), 0, s, n, host_dev, v, nrm, stride); break;
    case P16: hipLaunchKernelGGL(k_spinor_pack<P16>, grid_for(n), dim3(256), 0, s, n, host_dev, v, nrm, stride); break;
    default: throw Error(CMX_ERR_ARG, "spinor_pack: bad precision");
  }
}

void spinor_noise(int prec, int n, unsigned seed, void* v, float* nrm, int stride, hipStream_t s) {
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_spinor_noise<P64>, grid_for(n), dim3(256), 0, s, n, seed, v, nrm, stride); break;
    case P32: hipLaunchKernelGGL(k_spinor_noise<P32>, grid_for(n), dim3(256), 0, s, n, seed, v, nrm, stride); break;
    case P16: hipLaunchKernelGGL(k_spinor_noise<P16>, grid_for(n), dim3(256), 0, s, n, seed, v, nrm, stride); break;
    default: throw Error(CMX_ERR_ARG, "spinor_noise: bad precision");
  }
}

void spinor_unpack(int prec, int n, const void* v, const float* nrm, int stride, double* host_dev, hipStream_t s) {
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_spinor_unpack<P64>, grid_for(n), dim3(256), 0, s, n, v, nrm, stride, host_dev); break;
    case P32: hipLaunchKernelGGL(k_spinor_unpack<P32>, grid_for(n), dim3(256), 0, s, n, v, nrm, stride, host_dev); break;
    case P16: hipLaunchKernelGGL(k_spinor_unpack<P16>, grid_for(n), dim3(256), 0, s, n, v, nrm, stride, host_dev); break;
    default: throw Error(CMX_ERR_ARG, "spinor_unpack: bad precision");
  }
}

void stream_copy(long n_float4, const void* a, void* b, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_copy, grid_for(n_float4), dim3(256), 0, s, n_float4, (const float4*)a, (float4*)b);
}

void stream_mix(long n_uint4, const void* const a[5], void* out, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_mix, grid_for(n_uint4), dim3(256), 0, s, n_uint4, (const uint4*)a[0], (const uint4*)a[1],
                     (const uint4*)a[2], (const uint4*)a[3], (const uint4*)a[4], (uint4*)out);
}

int sum_partials(int n, const double* a, double* partials, hipStream_t s) {
  const int nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_sum_partials, dim3(nb), dim3(256), 0, s, n, a, partials);
  return nb;
}

}  // namespace cmx
