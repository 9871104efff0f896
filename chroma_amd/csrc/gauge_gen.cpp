// Synthetic gauge fields and sources (host): counter-based RNG keyed by (seed, site, mu, component),
// so every field is reproducible bit for bit from its seed.  SU(3) exponentials use a truncated
// Taylor series with scaling and squaring (Moler & Van Loan, "Nineteen dubious ways to compute the
// exponential of a matrix", method 3), accurate to rounding for these small matrices.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#include "chroma_mi355x.h"
#include "parallel.h"

namespace {

using cd = std::complex<double>;
struct M3 {
  cd a[3][3];
};

inline uint64_t splitmix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
inline uint64_t key(uint64_t seed, uint64_t site, uint64_t mu, uint64_t k) {
  return splitmix(splitmix(splitmix(seed) ^ site) ^ ((mu << 32) | k));
}
inline double uniform(uint64_t h) { return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }  // (0,1)
inline double gauss(uint64_t seed, uint64_t site, uint64_t mu, uint64_t k) {
  const double u1 = uniform(key(seed, site, mu, 2 * k)), u2 = uniform(key(seed, site, mu, 2 * k + 1));
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

M3 mul(const M3& A, const M3& B) {
  M3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cd s = 0;
      for (int k = 0; k < 3; ++k) s += A.a[i][k] * B.a[k][j];
      C.a[i][j] = s;
    }
  return C;
}
M3 dag(const M3& A) {
  M3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C.a[i][j] = std::conj(A.a[j][i]);
  return C;
}
M3 ident() {
  M3 I{};
  for (int i = 0; i < 3; ++i) I.a[i][i] = 1.0;
  return I;
}

// exp(X) by scaling and squaring with an order-18 Taylor polynomial
M3 expm(const M3& X) {
  double nrm = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) nrm += std::norm(X.a[i][j]);
  nrm = std::sqrt(nrm);
  int s = 0;
  while (nrm > 0.25) {
    nrm *= 0.5;
    ++s;
  }
  const double sc = std::ldexp(1.0, -s);
  M3 Y;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Y.a[i][j] = X.a[i][j] * sc;
  M3 E = ident(), T = ident();
  for (int k = 1; k <= 18; ++k) {
    T = mul(T, Y);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) T.a[i][j] /= (double)k;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) E.a[i][j] += T.a[i][j];
  }
  for (int k = 0; k < s; ++k) E = mul(E, E);
  return E;
}

// Gram-Schmidt on rows 0,1 and row2 = conj(row0 x row1): the nearest-by-construction SU(3) matrix
// (the role of reunit, lib/util/gauge/reunit.cc)
M3 reunit(M3 U) {
  double n0 = 0;
  for (int j = 0; j < 3; ++j) n0 += std::norm(U.a[0][j]);
  n0 = 1.0 / std::sqrt(n0);
  for (int j = 0; j < 3; ++j) U.a[0][j] *= n0;
  cd d = 0;
  for (int j = 0; j < 3; ++j) d += std::conj(U.a[0][j]) * U.a[1][j];
  for (int j = 0; j < 3; ++j) U.a[1][j] -= d * U.a[0][j];
  double n1 = 0;
  for (int j = 0; j < 3; ++j) n1 += std::norm(U.a[1][j]);
  n1 = 1.0 / std::sqrt(n1);
  for (int j = 0; j < 3; ++j) U.a[1][j] *= n1;
  for (int k = 0; k < 3; ++k) {
    const int a = (k + 1) % 3, b = (k + 2) % 3;
    U.a[2][k] = std::conj(U.a[0][a] * U.a[1][b] - U.a[0][b] * U.a[1][a]);
  }
  return U;
}

cd det3(const M3& U) {
  const auto& e = U.a;
  return e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) - e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
         e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
}

struct Lat {
  int L[4];
  long V, Vh;
  explicit Lat(const int d[4]) {
    V = 1;
    for (int m = 0; m < 4; ++m) L[m] = d[m], V *= d[m];
    Vh = V / 2;
  }
  void coords(long s, int x[4]) const {
    const int cb = s >= Vh ? 1 : 0;
    long r = s - cb * Vh;
    const int Lxh = L[0] / 2;
    const int xh = (int)(r % Lxh);
    r /= Lxh;
    x[1] = (int)(r % L[1]);
    r /= L[1];
    x[2] = (int)(r % L[2]);
    x[3] = (int)(r / L[2]);
    x[0] = 2 * xh + ((cb + x[1] + x[2] + x[3]) & 1);
  }
  long index(const int x[4]) const {
    const int cb = (x[0] + x[1] + x[2] + x[3]) & 1;
    return cb * Vh + ((((long)x[3] * L[2] + x[2]) * L[1] + x[1]) * L[0] + x[0]) / 2;
  }
  long shift(long s, int mu, int d) const {
    int x[4];
    coords(s, x);
    x[mu] = (x[mu] + d + L[mu]) % L[mu];
    return index(x);
  }
};

M3 load(const double* links, long V, int mu, long s) {
  M3 U;
  const double* p = links + ((size_t)mu * V + s) * 18;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) U.a[i][j] = cd(p[2 * (3 * i + j)], p[2 * (3 * i + j) + 1]);
  return U;
}
void store(double* links, long V, int mu, long s, const M3& U) {
  double* p = links + ((size_t)mu * V + s) * 18;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      p[2 * (3 * i + j)] = U.a[i][j].real();
      p[2 * (3 * i + j) + 1] = U.a[i][j].imag();
    }
}

// One isotropic stout step (smear_links, stout_utils.cc:936-962, with the staple sum and Q of
// getQsandCs, :105-180): U'_mu = exp(i Q_mu) U_mu,  C = rho sum_{nu != mu} (forward + backward staples),
// Omega = C U^dag,  Q = (i/2) [ (Omega^dag - Omega) - tr(Omega^dag - Omega)/3 ].
void stout_step(const Lat& lat, const double* cur, double* next, double rho) {
  const long V = lat.V;
  cmx::parallel_for(4 * V, [&](long t) {
    const int mu = (int)(t / V);
    const long s = t % V;
    M3 C{};
    for (int nu = 0; nu < 4; ++nu) {
      if (nu == mu) continue;
      const long xpm = lat.shift(s, mu, 1), xpn = lat.shift(s, nu, 1), xmn = lat.shift(s, nu, -1);
      const long xmnpm = lat.shift(xmn, mu, 1);
      // forward: U_nu(x) U_mu(x+nu) U_nu^dag(x+mu)
      M3 f = mul(mul(load(cur, V, nu, s), load(cur, V, mu, xpn)), dag(load(cur, V, nu, xpm)));
      // backward: U_nu^dag(x-nu) U_mu(x-nu) U_nu(x-nu+mu)
      M3 b = mul(mul(dag(load(cur, V, nu, xmn)), load(cur, V, mu, xmn)), load(cur, V, nu, xmnpm));
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C.a[i][j] += rho * (f.a[i][j] + b.a[i][j]);
    }
    const M3 U = load(cur, V, mu, s);
    const M3 Om = mul(C, dag(U));
    M3 T;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) T.a[i][j] = std::conj(Om.a[j][i]) - Om.a[i][j];
    const cd tr = (T.a[0][0] + T.a[1][1] + T.a[2][2]) / 3.0;
    for (int i = 0; i < 3; ++i) T.a[i][i] -= tr;
    // i Q = i * (i/2) T = -T/2
    M3 X;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) X.a[i][j] = -0.5 * T.a[i][j];
    store(next, V, mu, s, reunit(mul(expm(X), U)));
  });
}

}  // namespace

extern "C" {

int cmx_gauge_generate(const int dims[4], int kind, unsigned long long seed, double eps, int nstout, double rho,
                       double* links) {
  if (!dims || !links || kind < 0 || kind > 2 || nstout < 0) return CMX_ERR_ARG;
  for (int m = 0; m < 4; ++m)
    if (dims[m] <= 0 || dims[m] % 2) return CMX_ERR_ARG;
  const Lat lat(dims);
  const long V = lat.V;
  cmx::parallel_for(4 * V, [&](long t) {
    const int mu = (int)(t / V);
    const long s = t % V;
    M3 U = ident();
    if (kind == 1) {
      // A_ij = (N + iN)/sqrt(2); H = (A + A^dag)/2 - tr/3; U = exp(i eps H)
      M3 A;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          A.a[i][j] = cd(gauss(seed, s, mu, 2 * (3 * i + j)), gauss(seed, s, mu, 2 * (3 * i + j) + 1)) / std::sqrt(2.0);
      M3 X;
      cd tr = 0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) X.a[i][j] = 0.5 * (A.a[i][j] + std::conj(A.a[j][i]));
      for (int i = 0; i < 3; ++i) tr += X.a[i][i];
      for (int i = 0; i < 3; ++i) X.a[i][i] -= tr / 3.0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) X.a[i][j] *= cd(0.0, eps);
      U = reunit(expm(X));
    } else if (kind == 2) {
      // Haar: QR (Gram-Schmidt on columns, positive diagonal of R) of a complex Gaussian matrix, then
      // divide by a cube root of its determinant
      M3 Z;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          Z.a[i][j] = cd(gauss(seed, s, mu, 2 * (3 * i + j)), gauss(seed, s, mu, 2 * (3 * i + j) + 1));
      M3 Q{};
      for (int c = 0; c < 3; ++c) {
        cd v[3] = {Z.a[0][c], Z.a[1][c], Z.a[2][c]};
        for (int k = 0; k < c; ++k) {
          cd d = 0;
          for (int r = 0; r < 3; ++r) d += std::conj(Q.a[r][k]) * v[r];
          for (int r = 0; r < 3; ++r) v[r] -= d * Q.a[r][k];
        }
        double n = 0;
        for (int r = 0; r < 3; ++r) n += std::norm(v[r]);
        n = 1.0 / std::sqrt(n);
        for (int r = 0; r < 3; ++r) Q.a[r][c] = v[r] * n;
      }
      const cd ph = std::exp(cd(0.0, -std::arg(det3(Q)) / 3.0));
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Q.a[i][j] *= ph;
      U = reunit(Q);
    }
    store(links, V, mu, s, U);
  });
  if (nstout > 0) {
    std::vector<double> tmp((size_t)4 * V * 18);
    double* a = links;
    double* b = tmp.data();
    for (int k = 0; k < nstout; ++k) {
      stout_step(lat, a, b, rho);
      std::swap(a, b);
    }
    if (a != links) std::memcpy(links, a, sizeof(double) * 4 * V * 18);
  }
  return CMX_OK;
}

int cmx_gauge_apply_bc(const int dims[4], double* links, const int boundary[4], const double* twist_by_pi,
                       const int* twist_dir) {
  if (!dims || !links) return CMX_ERR_ARG;
  const Lat lat(dims);
  const long V = lat.V;
  cd tw[4] = {1.0, 1.0, 1.0, 1.0};
  if (twist_by_pi && twist_dir)
    for (int i = 0; i < 3; ++i) {
      const int d = twist_dir[i];
      if (d < 0 || d > 3) return CMX_ERR_ARG;
      // TwistedFermBC::modify (twisted_fermbc_w.h:106-115): arg = phase_by_pi * pi / L[dir]
      const double arg = twist_by_pi[i] * 3.141592653589793 / dims[d];
      tw[d] *= cd(std::cos(arg), std::sin(arg));
    }
  cmx::parallel_for(4 * V, [&](long t) {
    const int mu = (int)(t / V);
    const long s = t % V;
    int x[4];
    lat.coords(s, x);
    cd f = tw[mu];
    if (boundary && boundary[mu] != 1 && x[mu] == dims[mu] - 1) f *= (double)boundary[mu];
    double* p = links + (size_t)t * 18;
    for (int k = 0; k < 9; ++k) {
      const cd z = cd(p[2 * k], p[2 * k + 1]) * f;
      p[2 * k] = z.real();
      p[2 * k + 1] = z.imag();
    }
  });
  return CMX_OK;
}

double cmx_plaquette(const int dims[4], const double* links) {
  if (!dims || !links) return NAN;
  const Lat lat(dims);
  const long V = lat.V;
  const double acc = cmx::parallel_sum(V, [&](long s) {
    double a = 0;
    for (int mu = 0; mu < 3; ++mu)
      for (int nu = mu + 1; nu < 4; ++nu) {
        const M3 P = mul(mul(load(links, V, mu, s), load(links, V, nu, lat.shift(s, mu, 1))),
                         mul(dag(load(links, V, mu, lat.shift(s, nu, 1))), dag(load(links, V, nu, s))));
        a += (P.a[0][0] + P.a[1][1] + P.a[2][2]).real() / 3.0;
      }
    return a;
  });
  return acc / (6.0 * V);
}

int cmx_fill_z2(long n, unsigned long long seed, double* out) {
  if (!out || n < 0) return CMX_ERR_ARG;
  const double h = 1.0 / std::sqrt(2.0);
  cmx::parallel_for(n, [&](long k) { out[k] = (key(seed, (uint64_t)k, 7, 0) >> 63) ? h : -h; });
  return CMX_OK;
}

}  // extern "C"
