// Meson two-point functions on the device, in FP64: mesons() (mesons_w.cc:97-180) for the 16 gamma
// insertions Gamma(n) = gamma_1^n0 gamma_2^n1 gamma_3^n2 gamma_4^n3 (docs/notes/chroma_gamma_matrices.tex),
//   c_n(x) = tr[ adj(gamma_5 S_2(x) gamma_5) Gamma(n) S_1(x) Gamma(n) ]      (trace over spin and colour),
// then the sink-momentum projection of SftMom::sft (sftmom.cc:405-460, 520-530) over the spatial sites of
// each time slice, phase exp(i 2 pi sum_k n_k (x_k - origin_k) / L_k), with the time slice counted from
// the source, t_eff = (t - origin_t) mod L_t (mesons_w.cc:160-165).
//
// Every Gamma(n) and gamma_5 of the DeGrand-Rossi basis (docs/notes/gamma_conventions.tex:74-79) has one
// non-zero entry, a power of i, per row, so each of the 16 spin blocks (a, b) of the trace reads one 3x3
// colour block of S_2 and one of S_1 with a fixed coefficient; the host builds that table per Gamma.
// Propagators: [site][sink spin*3 + colour][source spin*3 + colour] complex, sites in the host order.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "op.h"

namespace cmx {

namespace {

struct Term {  // one spin block (a, b) of the trace for one Gamma
  double2 coef;
  int r2, c2, r1, c1;  // spin row / column of the S_2 and S_1 colour blocks
};

// signed permutation: row a has its one non-zero entry ph[a] in column p[a]
struct SPerm {
  int p[4];
  double ph_re[4], ph_im[4];
};

SPerm sp_mul(const SPerm& A, const SPerm& B) {
  SPerm C;
  for (int a = 0; a < 4; ++a) {
    const int k = A.p[a];
    C.p[a] = B.p[k];
    C.ph_re[a] = A.ph_re[a] * B.ph_re[k] - A.ph_im[a] * B.ph_im[k];
    C.ph_im[a] = A.ph_re[a] * B.ph_im[k] + A.ph_im[a] * B.ph_re[k];
  }
  return C;
}

SPerm gamma_dr(int mu) {  // gamma_1..gamma_4 (mu = 0..3) of the DeGrand-Rossi basis
  static const int perm[4][4] = {{3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}, {2, 3, 0, 1}};
  static const double re[4][4] = {{0, 0, 0, 0}, {-1, 1, 1, -1}, {0, 0, 0, 0}, {1, 1, 1, 1}};
  static const double im[4][4] = {{1, 1, -1, -1}, {0, 0, 0, 0}, {1, -1, -1, 1}, {0, 0, 0, 0}};
  SPerm g;
  for (int a = 0; a < 4; ++a) g.p[a] = perm[mu][a], g.ph_re[a] = re[mu][a], g.ph_im[a] = im[mu][a];
  return g;
}

SPerm gamma_n(int n) {
  SPerm g;
  for (int a = 0; a < 4; ++a) g.p[a] = a, g.ph_re[a] = 1, g.ph_im[a] = 0;
  for (int mu = 0; mu < 4; ++mu)
    if (n >> mu & 1) g = sp_mul(g, gamma_dr(mu));
  return g;
}

// c_n = sum_{a,b} sum_{ij} conj(anti_{(a,i),(b,j)}) (G S_1 G)_{(a,i),(b,j)} with anti = g5 S_2 g5:
// anti_{(a,i),(b,j)} = g5ph[a] S_2[(p5[a],i),(q,j)] g5ph[q] where p5[q] = b, and
// (G S_1 G)_{(a,i),(b,j)} = Gph[a] S_1[(pG[a],i),(r,j)] Gph[r] where pG[r] = b.
std::vector<Term> meson_terms() {
  const SPerm g5 = gamma_n(15);
  std::vector<Term> t;
  for (int n = 0; n < 16; ++n) {
    const SPerm G = gamma_n(n);
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        int q = -1, r = -1;
        for (int k = 0; k < 4; ++k) {
          if (g5.p[k] == b) q = k;
          if (G.p[k] == b) r = k;
        }
        // conj(g5ph[a] g5ph[q]) Gph[a] Gph[r]
        const double ar = g5.ph_re[a] * g5.ph_re[q] - g5.ph_im[a] * g5.ph_im[q];
        const double ai = -(g5.ph_re[a] * g5.ph_im[q] + g5.ph_im[a] * g5.ph_re[q]);
        const double br = G.ph_re[a] * G.ph_re[r] - G.ph_im[a] * G.ph_im[r];
        const double bi = G.ph_re[a] * G.ph_im[r] + G.ph_im[a] * G.ph_re[r];
        t.push_back(Term{make_double2(ar * br - ai * bi, ar * bi + ai * br), g5.p[a], q, G.p[a], r});
      }
  }
  return t;
}

// MESON_SITES sites per block, one lane per (site, Gamma): the block first copies its sites' two
// propagators (2 x 2304 B per site, contiguous) into LDS with coalesced loads, then each lane forms its
// trace from LDS, so HBM sees each propagator once instead of the 16 re-reads a lane per (site, Gamma)
// would make through the caches.  The block range is the nsite sites from s1/s2 (already offset to the
// range's first site); c[n * cstride + coff + i] for the range's site i.
constexpr int MESON_SITES = 8;

__global__ void __launch_bounds__(16 * MESON_SITES) k_meson_sites(long nsites, const double2* __restrict__ s1,
                                                                  const double2* __restrict__ s2,
                                                                  const Term* __restrict__ terms,
                                                                  double2* __restrict__ c, long cstride, long coff) {
  __shared__ double2 sh1[MESON_SITES * 144], sh2[MESON_SITES * 144];
  const long site0 = (long)blockIdx.x * MESON_SITES;
  const int nsite = nsites - site0 < MESON_SITES ? (int)(nsites - site0) : MESON_SITES;
  for (int q = threadIdx.x; q < nsite * 144; q += blockDim.x) {
    sh1[q] = s1[site0 * 144 + q];
    sh2[q] = s2[site0 * 144 + q];
  }
  __syncthreads();
  const int ls = (int)threadIdx.x >> 4;
  const int n = (int)threadIdx.x & 15;
  if (ls >= nsite) return;
  const double2* p1 = sh1 + ls * 144;
  const double2* p2 = sh2 + ls * 144;
  double cr = 0, ci = 0;
  for (int k = 0; k < 16; ++k) {
    const Term tm = terms[n * 16 + k];
    double sr = 0, si = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double2 x = p2[(tm.r2 * 3 + i) * 12 + tm.c2 * 3 + j];
        const double2 y = p1[(tm.r1 * 3 + i) * 12 + tm.c1 * 3 + j];
        sr += x.x * y.x + x.y * y.y;  // conj(x) y
        si += x.x * y.y - x.y * y.x;
      }
    cr += tm.coef.x * sr - tm.coef.y * si;
    ci += tm.coef.x * si + tm.coef.y * sr;
  }
  c[(size_t)n * cstride + coff + site0 + ls] = make_double2(cr, ci);
}

struct ProjGeom {
  int L[4];
  long V, Vh;
  int org[4];
};

// Where k_meson_project finds the correlator of a site (cb, lex) of time slice t0 + blockIdx.x:
// c[n * cstride + cb * cb_off + lex / 2 - lex_off].  Whole lattice: cstride V, cb_off Vh, lex_off 0.
struct CorrIndex {
  int t0;
  long cstride, cb_off, lex_off;
};

// one block per (time slice, momentum): out[n][m][t_eff] = sum over the slice of phase(x) c[n][x], summed in
// a fixed order (strided per lane, then a tree over the block), so the result is deterministic
__global__ void __launch_bounds__(256) k_meson_project(ProjGeom g, CorrIndex ci, const double2* __restrict__ c,
                                                       const int* __restrict__ moms, int nmom,
                                                       double2* __restrict__ out) {
  __shared__ double2 red[256];
  const int t = ci.t0 + (int)blockIdx.x, m = blockIdx.y;
  const long Vs = (long)g.L[0] * g.L[1] * g.L[2];
  const double k0 = 6.283185307179586 * moms[3 * m] / g.L[0], k1 = 6.283185307179586 * moms[3 * m + 1] / g.L[1],
               k2 = 6.283185307179586 * moms[3 * m + 2] / g.L[2];
  double2 acc[16];
#pragma unroll
  for (int n = 0; n < 16; ++n) acc[n] = make_double2(0, 0);
  for (long v = threadIdx.x; v < Vs; v += blockDim.x) {
    const int x0 = (int)(v % g.L[0]), x1 = (int)(v / g.L[0] % g.L[1]), x2 = (int)(v / ((long)g.L[0] * g.L[1]));
    const long lex = v + Vs * t;
    const long idx = ((x0 + x1 + x2 + t) & 1) * ci.cb_off + lex / 2 - ci.lex_off;
    double sn, cs;
    sincos(k0 * (x0 - g.org[0]) + k1 * (x1 - g.org[1]) + k2 * (x2 - g.org[2]), &sn, &cs);
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const double2 z = c[(size_t)n * ci.cstride + idx];
      acc[n].x += cs * z.x - sn * z.y;
      acc[n].y += cs * z.y + sn * z.x;
    }
  }
  const int teff = (t - g.org[3] % g.L[3] + g.L[3]) % g.L[3];
  for (int n = 0; n < 16; ++n) {
    red[threadIdx.x] = acc[n];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w)
        red[threadIdx.x] = make_double2(red[threadIdx.x].x + red[threadIdx.x + w].x, red[threadIdx.x].y + red[threadIdx.x + w].y);
      __syncthreads();
    }
    if (threadIdx.x == 0) out[((size_t)n * nmom + m) * g.L[3] + teff] = red[0];
    __syncthreads();
  }
}

ProjGeom proj_geom(const int dims[4], const int origin[4]) {
  ProjGeom g;
  g.V = 1;
  for (int m = 0; m < 4; ++m) {
    CMX_REQUIRE(dims[m] > 0 && dims[m] % 2 == 0, CMX_ERR_ARG, "meson_2pt: extents must be even and positive");
    g.L[m] = dims[m];
    g.V *= dims[m];
    g.org[m] = origin ? origin[m] : 0;
    CMX_REQUIRE(g.org[m] >= 0 && g.org[m] < dims[m], CMX_ERR_ARG, "meson_2pt: origin outside the lattice");
  }
  g.Vh = g.V / 2;
  CMX_REQUIRE(g.V <= (long)INT32_MAX * MESON_SITES, CMX_ERR_ARG, "meson_2pt: lattice too large for one grid");
  return g;
}

void launch_sites(long nsites, const double2* s1, const double2* s2, const Term* terms, double2* c, long cstride,
                  long coff, hipStream_t st) {
  hipLaunchKernelGGL(k_meson_sites, dim3((unsigned)((nsites + MESON_SITES - 1) / MESON_SITES)), dim3(16 * MESON_SITES), 0,
                     st, nsites, s1, s2, terms, c, cstride, coff);
  CMX_HIP(hipGetLastError());
}

// The allocation-free variant's scratch: the correlators of up to MESON_SCRATCH_SITES sites (a pass of whole
// time slices; 64 MiB) followed by the Gamma-term table (written once per device at set-up, never
// overwritten).  Host code never uses this name as an address: it uses the per-device pointer that
// hipGetSymbolAddress returns.
constexpr long MESON_SCRATCH_SITES = 262144;  // 64^3: the largest spatial volume one pass holds
constexpr size_t MESON_SCRATCH_CORR = sizeof(double2) * 16 * MESON_SCRATCH_SITES;
constexpr size_t MESON_SCRATCH_BYTES = MESON_SCRATCH_CORR + sizeof(Term) * 256;
static __device__ __align__(256) char scratch[MESON_SCRATCH_BYTES];

constexpr int MESON_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MESON_MAX_DEVICES] = {};

}  // namespace

void meson_2pt(const int dims[4], const double* prop1, const double* prop2, const int origin[4], int nmom,
               const int* moms, double* corr) {
  CMX_REQUIRE(dims && prop1 && prop2 && corr, CMX_ERR_ARG, "meson_2pt: null array");
  CMX_REQUIRE(nmom >= 1 && nmom <= 4096 && moms, CMX_ERR_ARG, "meson_2pt: 1..4096 momenta");
  const ProjGeom g = proj_geom(dims, origin);
  const std::vector<Term> terms = meson_terms();
  const size_t pbytes = sizeof(double) * 288 * (size_t)g.V;
  DevBuf d1(pbytes), d2(pbytes), dc(sizeof(double2) * 16 * (size_t)g.V), dt(sizeof(Term) * terms.size()),
      dm(sizeof(int) * 3 * (size_t)nmom), dout(sizeof(double2) * 16 * (size_t)nmom * g.L[3]);
  CMX_HIP(hipMemcpy(d1.p, prop1, pbytes, hipMemcpyHostToDevice));
  if (prop2 == prop1)
    CMX_HIP(hipMemcpy(d2.p, d1.p, pbytes, hipMemcpyDeviceToDevice));
  else
    CMX_HIP(hipMemcpy(d2.p, prop2, pbytes, hipMemcpyHostToDevice));
  CMX_HIP(hipMemcpy(dt.p, terms.data(), dt.bytes, hipMemcpyHostToDevice));
  CMX_HIP(hipMemcpy(dm.p, moms, dm.bytes, hipMemcpyHostToDevice));
  launch_sites(g.V, d1.as<double2>(), d2.as<double2>(), dt.as<Term>(), dc.as<double2>(), g.V, 0, 0);
  const CorrIndex ci{0, g.V, g.Vh, 0};
  hipLaunchKernelGGL(k_meson_project, dim3(g.L[3], nmom), dim3(256), 0, 0, g, ci, dc.as<double2>(), dm.as<int>(), nmom,
                     dout.as<double2>());
  CMX_HIP(hipGetLastError());
  CMX_HIP(hipMemcpy(corr, dout.p, dout.bytes, hipMemcpyDeviceToHost));
}

void meson_scratch_setup() {
  int dev = 0;
  CMX_HIP(hipGetDevice(&dev));
  CMX_REQUIRE(dev >= 0 && dev < MESON_MAX_DEVICES, CMX_ERR_UNSUPPORTED, "meson_2pt_noalloc: device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return;
  void* ptr = nullptr;
  CMX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  CMX_REQUIRE(ptr != nullptr, CMX_ERR_HIP, "meson_2pt_noalloc: no device address for the scratch array");
  const std::vector<Term> terms = meson_terms();
  CMX_HIP(hipMemcpy(static_cast<char*>(ptr) + MESON_SCRATCH_CORR, terms.data(), sizeof(Term) * terms.size(),
                    hipMemcpyHostToDevice));
  g_scratch_ptr[dev] = ptr;
}

void meson_2pt_noalloc(const int dims[4], const double* prop1, const double* prop2, const int origin[4], int nmom,
                       const int* moms, double* corr, hipStream_t stream) {
  CMX_REQUIRE(dims && prop1 && prop2 && corr && moms, CMX_ERR_ARG, "meson_2pt_noalloc: null array");
  CMX_REQUIRE(nmom >= 1 && nmom <= 4096, CMX_ERR_ARG, "meson_2pt_noalloc: 1..4096 momenta");
  const ProjGeom g = proj_geom(dims, origin);
  const long Vs = (long)g.L[0] * g.L[1] * g.L[2];
  CMX_REQUIRE(Vs <= MESON_SCRATCH_SITES, CMX_ERR_ARG,
              "meson_2pt_noalloc: spatial volume above 262144 sites (64^3), the scratch array's limit");
  meson_scratch_setup();
  int dev = 0;
  CMX_HIP(hipGetDevice(&dev));
  char* base = static_cast<char*>(g_scratch_ptr[dev]);
  double2* c = reinterpret_cast<double2*>(base);
  const Term* terms = reinterpret_cast<const Term*>(base + MESON_SCRATCH_CORR);
  const double2* s1 = reinterpret_cast<const double2*>(prop1);
  const double2* s2 = reinterpret_cast<const double2*>(prop2);
  // passes of T whole time slices; the slice range [t0, t0 + T) is two contiguous site ranges, one per
  // checkerboard, stored one after the other in the scratch
  const int T = (int)std::min<long>(g.L[3], MESON_SCRATCH_SITES / Vs);
  for (int t0 = 0; t0 < g.L[3]; t0 += T) {
    const int nt = std::min(T, g.L[3] - t0);
    const long half = Vs * nt / 2, lex0 = Vs * t0 / 2;
    for (int cb = 0; cb < 2; ++cb) {
      const size_t first = (size_t)(cb * g.Vh + lex0) * 144;
      launch_sites(half, s1 + first, s2 + first, terms, c, Vs * nt, cb * half, stream);
    }
    const CorrIndex ci{t0, Vs * nt, half, lex0};
    hipLaunchKernelGGL(k_meson_project, dim3(nt, nmom), dim3(256), 0, stream, g, ci, c, moms, nmom,
                       reinterpret_cast<double2*>(corr));
    CMX_HIP(hipGetLastError());
  }
}

}  // namespace cmx
