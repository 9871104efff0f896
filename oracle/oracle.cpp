// ORACLE -- test infrastructure only (see oracle.hpp for the layout contract).
//
// Each function names the Chroma file:line it restates.  Everything here is plain site loops over
// std::complex, written for reading, not speed; OpenMP only splits the site loops so that the same
// code can be timed as the CPU baseline.
#include "oracle.hpp"

#include <omp.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

namespace orc {

using cd = std::complex<double>;
static const cd I1(0.0, 1.0);

// gamma_1..gamma_4 of the DeGrand-Rossi basis, docs/notes/gamma_conventions.tex:74-79, with
// Chroma's direction mu = 0..3 <-> gamma_1..gamma_4 (docs/notes/chroma_gamma_matrices.tex:45-61).
const cd kGamma[4][4][4] = {
    // gamma_1 = [[0, i sigma_1], [-i sigma_1, 0]]
    {{0, 0, 0, I1}, {0, 0, I1, 0}, {0, -I1, 0, 0}, {-I1, 0, 0, 0}},
    // gamma_2 = [[0, -i sigma_2], [i sigma_2, 0]]
    {{0, 0, 0, -1.0}, {0, 0, 1.0, 0}, {0, 1.0, 0, 0}, {-1.0, 0, 0, 0}},
    // gamma_3 = [[0, i sigma_3], [-i sigma_3, 0]]
    {{0, 0, I1, 0}, {0, 0, 0, -I1}, {-I1, 0, 0, 0}, {0, I1, 0, 0}},
    // gamma_4 = [[0, 1], [1, 0]]
    {{0, 0, 1.0, 0}, {0, 0, 0, 1.0}, {1.0, 0, 0, 0}, {0, 1.0, 0, 0}},
};

Geom::Geom(const int dims[4]) {
  V = 1;
  for (int m = 0; m < 4; ++m) {
    L[m] = dims[m];
    if (L[m] <= 0 || (L[m] & 1)) throw std::runtime_error("lattice extents must be positive and even");
    V *= L[m];
  }
  Vh = V / 2;
  nb.resize(V * 8);
  for (long s = 0; s < V; ++s) {
    int x[4];
    coords(s, x);
    for (int mu = 0; mu < 4; ++mu) {
      int y[4] = {x[0], x[1], x[2], x[3]};
      y[mu] = (x[mu] + 1) % L[mu];
      nb[s * 8 + 2 * mu] = index(y);
      y[mu] = (x[mu] - 1 + L[mu]) % L[mu];
      nb[s * 8 + 2 * mu + 1] = index(y);
    }
  }
}

void Geom::coords(long idx, int x[4]) const {
  int cb = idx >= Vh ? 1 : 0;
  long r = idx - cb * Vh;
  int Lxh = L[0] / 2;
  int xh = r % Lxh;
  r /= Lxh;
  x[1] = r % L[1];
  r /= L[1];
  x[2] = r % L[2];
  x[3] = r / L[2];
  x[0] = 2 * xh + ((cb + x[1] + x[2] + x[3]) & 1);
}

long Geom::index(const int x[4]) const {
  int cb = (x[0] + x[1] + x[2] + x[3]) & 1;
  long lex = (((long)x[3] * L[2] + x[2]) * L[1] + x[1]) * L[0] + x[0];
  return cb * Vh + lex / 2;
}

// ------------------------------------------------------------------------------------------------
// Small colour / spin helpers
template <typename R> using CM = cplx<R>[3][3];

template <typename R> static inline const cplx<R>* link(const cplx<R>* U, long V, int mu, long s) {
  return U + ((long)mu * V + s) * 9;
}

// (1 + sgn*gamma_mu) psi over spin, psi is [4][3]
template <typename R>
static inline void spin_proj(cplx<R> out[4][3], const cplx<R>* psi, int mu, double sgn) {
  for (int s = 0; s < 4; ++s)
    for (int c = 0; c < 3; ++c) {
      cplx<R> acc = psi[s * 3 + c];
      for (int t = 0; t < 4; ++t) {
        const cd g = kGamma[mu][s][t];
        if (g != cd(0.0)) acc += cplx<R>(R(sgn * g.real()), R(sgn * g.imag())) * psi[t * 3 + c];
      }
      out[s][c] = acc;
    }
}

// ------------------------------------------------------------------------------------------------
// Wilson hopping term.  lwldslash_w.h:226-303 (formula in the header comment :21-43):
//   chi(x) = sum_mu U_mu(x) (1 - isign gamma_mu) psi(x+mu) + U_mu^dag(x-mu) (1 + isign gamma_mu) psi(x-mu)
// on the output checkerboard cb only.  Written with the full 4x4 gamma matrices rather than the
// half-spinor projection so that the product's projection tricks are checked by it.
template <typename R>
void dslash(const Geom& g, const cplx<R>* U, cplx<R>* chi, const cplx<R>* psi, int isign, int cb) {
  const long lo = cb * g.Vh, hi = lo + g.Vh;
#pragma omp parallel for schedule(static)
  for (long s = lo; s < hi; ++s) {
    cplx<R> acc[12];
    for (int k = 0; k < 12; ++k) acc[k] = 0;
    for (int mu = 0; mu < 4; ++mu) {
      cplx<R> h[4][3];
      // forward: U_mu(x) (1 - isign gamma_mu) psi(x+mu)
      const long sp = g.fwd(s, mu);
      spin_proj(h, psi + sp * 12, mu, -double(isign));
      const cplx<R>* u = link(U, g.V, mu, s);
      for (int sp_ = 0; sp_ < 4; ++sp_)
        for (int i = 0; i < 3; ++i) {
          cplx<R> t = 0;
          for (int j = 0; j < 3; ++j) t += u[i * 3 + j] * h[sp_][j];
          acc[sp_ * 3 + i] += t;
        }
      // backward: U_mu(x-mu)^dag (1 + isign gamma_mu) psi(x-mu)
      const long sm = g.bwd(s, mu);
      spin_proj(h, psi + sm * 12, mu, double(isign));
      const cplx<R>* um = link(U, g.V, mu, sm);
      for (int sp_ = 0; sp_ < 4; ++sp_)
        for (int i = 0; i < 3; ++i) {
          cplx<R> t = 0;
          for (int j = 0; j < 3; ++j) t += std::conj(um[j * 3 + i]) * h[sp_][j];
          acc[sp_ * 3 + i] += t;
        }
    }
    for (int k = 0; k < 12; ++k) chi[s * 12 + k] = acc[k];
  }
}

// WilsonDslashBase::deriv, lwldslash_base_w.h:132-227: for x on cb and every mu,
//   ds_u[mu](x)_{ab} = sum_spin [(1 - isign gamma_mu) psi](x+mu)_{spin,a} conj(chi(x)_{spin,b})
// (traceSpin(outerProduct(shift(temp, FORWARD, mu), chi)); isotropic weights 1), zero on rb[1-cb];
// the simple and twisted fermion BCs leave it unchanged (SimpleFermBC::zero, simple_fermbc.h:123).
// ds_u is [4][V][3][3].
template <typename R>
void dslash_deriv(const Geom& g, cplx<R>* ds, const cplx<R>* chi, const cplx<R>* psi, int isign, int cb) {
  for (int mu = 0; mu < 4; ++mu)
    for (long s = (1 - cb) * g.Vh; s < (2 - cb) * g.Vh; ++s)
      for (int k = 0; k < 9; ++k) ds[((long)mu * g.V + s) * 9 + k] = 0;
  const long lo = cb * g.Vh, hi = lo + g.Vh;
#pragma omp parallel for schedule(static)
  for (long s = lo; s < hi; ++s)
    for (int mu = 0; mu < 4; ++mu) {
      cplx<R> h[4][3];
      spin_proj(h, psi + g.fwd(s, mu) * 12, mu, -double(isign));
      cplx<R>* d = ds + ((long)mu * g.V + s) * 9;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          cplx<R> t = 0;
          for (int sp = 0; sp < 4; ++sp) t += h[sp][a] * std::conj(chi[s * 12 + sp * 3 + b]);
          d[a * 3 + b] = t;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Field strength, mesfield.cc:30-78 (f[offset] for (mu,nu) = (0,1),(0,2),(0,3),(1,2),(1,3),(2,3)).
// F is [6][V][3][3].  shift(f, FORWARD, mu)(x) = f(x+mu); shift(f, BACKWARD, mu)(x) = f(x-mu).
template <typename R> static inline void mm(cplx<R>* c, const cplx<R>* a, const cplx<R>* b) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cplx<R> t = 0;
      for (int k = 0; k < 3; ++k) t += a[i * 3 + k] * b[k * 3 + j];
      c[i * 3 + j] = t;
    }
}
template <typename R> static inline void mma(cplx<R>* c, const cplx<R>* a, const cplx<R>* b) {  // a * b^dag
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cplx<R> t = 0;
      for (int k = 0; k < 3; ++k) t += a[i * 3 + k] * std::conj(b[j * 3 + k]);
      c[i * 3 + j] = t;
    }
}
template <typename R> static inline void amm(cplx<R>* c, const cplx<R>* a, const cplx<R>* b) {  // a^dag * b
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cplx<R> t = 0;
      for (int k = 0; k < 3; ++k) t += std::conj(a[k * 3 + i]) * b[k * 3 + j];
      c[i * 3 + j] = t;
    }
}

template <typename R> void mesfield(const Geom& g, const cplx<R>* U, cplx<R>* F) {
  const long V = g.V;
  std::vector<cplx<R>> t2(V * 9), t1b(V * 9), t0b(V * 9);
  int off = 0;
  for (int mu = 0; mu < 3; ++mu)
    for (int nu = mu + 1; nu < 4; ++nu, ++off) {
      cplx<R>* f = F + (long)off * V * 9;
      // pass 1 (site-local at x): tmp_3 = u_nu(x+mu); tmp_4 = u_mu(x+nu); tmp_0 = u_nu tmp_4;
      // tmp_1 = u_mu tmp_3; f = tmp_1 tmp_0^dag; tmp_2 = tmp_0^dag tmp_1;
      // tmp_1' = tmp_4 tmp_3^dag; tmp_0' = u_nu^dag u_mu
#pragma omp parallel for schedule(static)
      for (long s = 0; s < V; ++s) {
        const cplx<R>* tmp3 = link(U, V, nu, g.fwd(s, mu));
        const cplx<R>* tmp4 = link(U, V, mu, g.fwd(s, nu));
        cplx<R> tmp0[9], tmp1[9];
        mm(tmp0, link(U, V, nu, s), tmp4);
        mm(tmp1, link(U, V, mu, s), tmp3);
        mma(f + s * 9, tmp1, tmp0);
        amm(&t2[s * 9], tmp0, tmp1);
        mma(&t1b[s * 9], tmp4, tmp3);
        amm(&t0b[s * 9], link(U, V, nu, s), link(U, V, mu, s));
      }
      // pass 2: f += shift(shift(tmp_2, BACKWARD, nu), BACKWARD, mu)
      //         f += shift(tmp_0' tmp_1'^dag, BACKWARD, nu)
      //         f += shift(tmp_1'^dag tmp_0', BACKWARD, mu)
      //         f = (f - f^dag) * 1/8
#pragma omp parallel for schedule(static)
      for (long s = 0; s < V; ++s) {
        cplx<R> acc[9];
        for (int k = 0; k < 9; ++k) acc[k] = f[s * 9 + k];
        const long smn = g.bwd(g.bwd(s, nu), mu);
        for (int k = 0; k < 9; ++k) acc[k] += t2[smn * 9 + k];
        cplx<R> tmp[9];
        const long sn = g.bwd(s, nu);
        mma(tmp, &t0b[sn * 9], &t1b[sn * 9]);
        for (int k = 0; k < 9; ++k) acc[k] += tmp[k];
        const long sm = g.bwd(s, mu);
        amm(tmp, &t1b[sm * 9], &t0b[sm * 9]);
        for (int k = 0; k < 9; ++k) acc[k] += tmp[k];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) f[s * 9 + i * 3 + j] = (acc[i * 3 + j] - std::conj(acc[j * 3 + i])) * R(0.125);
      }
    }
}

// ------------------------------------------------------------------------------------------------
// One isotropic stout step: Stouting::smear_links (stout_utils.cc:936-962) with every direction smeared
// and rho(mu,nu) = rho, the staples and Q of getQsandCs (:105-180):
//   C = rho sum_{nu != mu} [u_nu(x) u_mu(x+nu) u_nu^dag(x+mu) + shift(u_nu^dag u_mu u_nu(.+mu), BACKWARD, nu)],
//   Omega = C u_mu^dag,  Q = i/2 [(Omega^dag - Omega) - tr(Omega^dag - Omega)/Nc],  next_mu = exp(iQ) u_mu.
// The reference forms exp(iQ) = f0 + f1 Q + f2 Q^2 by Cayley-Hamilton (getFsAndBs, :383-874); here it
// is the plain Taylor series of exp(iQ), summed term by term until a term is below 1e-17 of the
// running sum (|iQ| stays O(1)), with no scaling, no squaring and no re-projection onto SU(3).
static void exp_taylor(cd* E, const cd* X) {
  cd T[9], N[9];
  for (int k = 0; k < 9; ++k) E[k] = T[k] = (k % 4 == 0) ? cd(1.0) : cd(0.0);
  for (int n = 1; n < 200; ++n) {
    mm(N, T, X);
    double tn = 0, en = 0;
    for (int k = 0; k < 9; ++k) {
      T[k] = N[k] / double(n);
      E[k] += T[k];
      tn += std::norm(T[k]);
      en += std::norm(E[k]);
    }
    if (n > 3 && tn < 1e-34 * en) return;
  }
  throw std::runtime_error("stout: exp series did not converge");
}

// dirs: bit mu set = direction mu smeared (smear_in_this_dirP); an unsmeared direction is copied and
// gives no staples (smear_links :945-958, getQsandCs :120).
void stout_step(const Geom& g, const cd* U, cd* out, double rho, int dirs) {
  const long V = g.V;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < 4 * V; ++t) {
    const int mu = (int)(t / V);
    const long s = t % V;
    if (!(dirs >> mu & 1)) {
      for (int k = 0; k < 9; ++k) out[t * 9 + k] = U[t * 9 + k];
      continue;
    }
    cd C[9] = {}, a[9], b[9];
    for (int nu = 0; nu < 4; ++nu) {
      if (nu == mu || !(dirs >> nu & 1)) continue;
      // forward staple
      mm(a, link(U, V, nu, s), link(U, V, mu, g.fwd(s, nu)));
      mma(b, a, link(U, V, nu, g.fwd(s, mu)));
      for (int k = 0; k < 9; ++k) C[k] += rho * b[k];
      // backward staple, built on y = x - nu
      const long y = g.bwd(s, nu);
      amm(a, link(U, V, nu, y), link(U, V, mu, y));
      mm(b, a, link(U, V, nu, g.fwd(y, mu)));
      for (int k = 0; k < 9; ++k) C[k] += rho * b[k];
    }
    cd Om[9], X[9], E[9];
    mma(Om, C, link(U, V, mu, s));
    cd tr = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) X[i * 3 + j] = std::conj(Om[j * 3 + i]) - Om[i * 3 + j];
    for (int i = 0; i < 3; ++i) tr += X[i * 4];
    for (int i = 0; i < 3; ++i) X[i * 4] -= tr / 3.0;
    for (int k = 0; k < 9; ++k) X[k] *= I1 * (0.5 * I1);  // i Q with Q = (i/2) X
    exp_taylor(E, X);
    mm(out + t * 9, E, link(U, V, mu, s));
  }
}

// ------------------------------------------------------------------------------------------------
// Meson two-point functions, mesons() (mesons_w.cc:97-180), for gamma_value n = 0..15 with
//   Gamma(n) = gamma_1^n0 gamma_2^n1 gamma_3^n2 gamma_4^n3 (docs/notes/chroma_gamma_matrices.tex),
//   anti_quark_prop = Gamma(15) S2 Gamma(15),  corr_fn = trace(adj(anti_quark_prop) Gamma(n) S1 Gamma(n)),
//   hsum[m][t] = sum_{x in slice t} exp(i p_m.(x - origin)) corr_fn(x),  p_k = 2 pi mom_k / L_k
//   (SftMom phases, sftmom.cc:405-460; decay direction 3), stored at t_eff = (t - t0) mod L_t
//   (mesons_w.cc:160-165).  Dense 12x12 spin-colour matrices per site; S[s] is [12 sink][12 source].
static void gamma_product(cd G[4][4], int n) {
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) G[a][b] = a == b ? 1.0 : 0.0;
  for (int mu = 0; mu < 4; ++mu) {
    if (!(n >> mu & 1)) continue;
    cd T[4][4];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        T[a][b] = 0;
        for (int k = 0; k < 4; ++k) T[a][b] += G[a][k] * kGamma[mu][k][b];
      }
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) G[a][b] = T[a][b];
  }
}

// (G x 1_colour) S (G x 1_colour) for a 12x12 S
static void sandwich(cd* out, const cd G[4][4], const cd* S) {
  cd T[144];
  for (int r = 0; r < 12; ++r)
    for (int c = 0; c < 12; ++c) {
      cd t = 0;
      for (int k = 0; k < 4; ++k) t += G[r / 3][k] * S[(k * 3 + r % 3) * 12 + c];
      T[r * 12 + c] = t;
    }
  for (int r = 0; r < 12; ++r)
    for (int c = 0; c < 12; ++c) {
      cd t = 0;
      for (int k = 0; k < 4; ++k) t += T[r * 12 + k * 3 + c % 3] * G[k][c / 3];
      out[r * 12 + c] = t;
    }
}

void mesons(const Geom& g, const cd* S1, const cd* S2, const int org[4], int nmom, const int* moms, cd* out) {
  const int Lt = g.L[3];
  std::vector<cd> corr((size_t)16 * g.V);
  cd G5[4][4];
  gamma_product(G5, 15);
#pragma omp parallel for schedule(static)
  for (long s = 0; s < g.V; ++s) {
    cd A[144], B[144], G[4][4];
    sandwich(A, G5, S2 + s * 144);
    for (int n = 0; n < 16; ++n) {
      gamma_product(G, n);
      sandwich(B, G, S1 + s * 144);
      cd t = 0;
      for (int k = 0; k < 144; ++k) t += std::conj(A[k]) * B[k];
      corr[(size_t)n * g.V + s] = t;
    }
  }
  for (size_t k = 0; k < (size_t)16 * nmom * Lt; ++k) out[k] = 0;
  for (long s = 0; s < g.V; ++s) {
    int x[4];
    g.coords(s, x);
    const int teff = ((x[3] - org[3]) % Lt + Lt) % Lt;
    for (int m = 0; m < nmom; ++m) {
      double pdx = 0;
      for (int k = 0; k < 3; ++k) pdx += 6.283185307179586 * moms[3 * m + k] * (x[k] - org[k]) / g.L[k];
      const cd ph(std::cos(pdx), std::sin(pdx));
      for (int n = 0; n < 16; ++n) out[((size_t)n * nmom + m) * Lt + teff] += ph * corr[(size_t)n * g.V + s];
    }
  }
}

// ------------------------------------------------------------------------------------------------
// Clover build: QDPCloverTermT::create (clover_term_qdp_w.h:244-294) + makeClov (:525-553)
// + makeClovSiteLoop (:398-522).  Isotropic only: clovCoeffR = clovCoeffT = csw (clover_fermact_
// params_w.cc:65-77), then both scaled by 1/2 (:263-267); diag_mass = Nd + Mass (:274-278).
template <typename R>
void make_clover(const Geom& g, const cplx<R>* U, R csw_r, R csw_t, R mass, Clov<R>* tri) {
  const long V = g.V;
  std::vector<cplx<R>> F((long)6 * V * 9);
  mesfield(g, U, F.data());
  const R cR = csw_r * R(0.5), cT = csw_t * R(0.5);
  // getCloverCoeff (:1525-1547): t_dir = 3 uses the temporal coefficient
  const R coef[6] = {cR, cR, cT, cR, cT, cT};
  const R diag_mass = R(1) + R(3) + mass;
#pragma omp parallel for schedule(static)
  for (long s = 0; s < V; ++s) {
    const cplx<R>* fp[6];
    cplx<R> fs[6][9];
    for (int k = 0; k < 6; ++k) {
      for (int e = 0; e < 9; ++e) fs[k][e] = F[((long)k * V + s) * 9 + e] * coef[k];
      fp[k] = fs[k];
    }
    auto f = [&](int k, int i, int j) { return fp[k][i * 3 + j]; };
    Clov<R>& t = tri[s];
    for (int jj = 0; jj < 2; ++jj)
      for (int ii = 0; ii < 6; ++ii) t.diag[jj][ii] = diag_mass;
    for (int i = 0; i < 3; ++i) {
      R r0 = std::imag(f(5, i, i) - f(0, i, i));
      t.diag[0][i] += r0;
      t.diag[0][i + 3] -= r0;
      R r1 = std::imag(f(5, i, i) + f(0, i, i));
      t.diag[1][i] -= r1;
      t.diag[1][i + 3] += r1;
    }
    const cplx<R> iu(0, 1);
    for (int i = 1; i < 3; ++i)
      for (int j = 0; j < i; ++j) {
        int eij = i * (i - 1) / 2 + j;
        int etmp = (i + 3) * (i + 2) / 2 + j + 3;
        t.offd[0][eij] = iu * (f(0, i, j) - f(5, i, j));
        t.offd[0][etmp] = -t.offd[0][eij];
        t.offd[1][eij] = iu * (f(5, i, j) + f(0, i, j));
        t.offd[1][etmp] = -t.offd[1][eij];
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        int eij = (i + 3) * (i + 2) / 2 + j;
        cplx<R> Em = iu * f(2, i, j) + f(4, i, j);
        cplx<R> Bm = iu * f(3, i, j) - f(1, i, j);
        t.offd[0][eij] = Bm - Em;
        t.offd[1][eij] = Em + Bm;
      }
  }
}

// LDL^dag inversion of each 6x6 block on checkerboard cb, in place.
// ldagdlinv / LDagDLInvSiteLoop, clover_term_qdp_w.h:619-846.  Returns sum over sites of
// sum_i log|d_i| (cholesDet, :580-606) and the number of sites with negative pivots (:810-814).
template <typename R> double clover_invert(const Geom& g, Clov<R>* tri, int cb, long* nneg_sites) {
  const long lo = cb * g.Vh, hi = lo + g.Vh;
  double logdet = 0.0;
  long nneg = 0;
  const int N = 6;
#pragma omp parallel for schedule(static) reduction(+ : logdet, nneg)
  for (long s = lo; s < hi; ++s) {
    int site_neg = 0;
    double site_log = 0.0;
    for (int block = 0; block < 2; ++block) {
      R inv_d[6], diag_g[6];
      cplx<R> inv_offd[15], v[6];
      for (int i = 0; i < N; ++i) inv_d[i] = tri[s].diag[block][i];
      for (int i = 0; i < 15; ++i) inv_offd[i] = tri[s].offd[block][i];
      // Golub & van Loan algorithm 4.1.2 (:649-718)
      for (int j = 0; j < N; ++j) {
        for (int i = 0; i < j; ++i) {
          int eji = j * (j - 1) / 2 + i;
          v[i] = cplx<R>(inv_d[i], 0) * std::conj(inv_offd[eji]);
        }
        v[j] = cplx<R>(inv_d[j], 0);
        for (int k = 0; k < j; ++k) {
          int ejk = j * (j - 1) / 2 + k;
          v[j] -= inv_offd[ejk] * v[k];
        }
        inv_d[j] = std::real(v[j]);
        for (int k = j + 1; k < N; ++k) {
          int ekj = k * (k - 1) / 2 + j;
          for (int l = 0; l < j; ++l) {
            int ekl = k * (k - 1) / 2 + l;
            inv_offd[ekj] -= inv_offd[ekl] * v[l];
          }
          inv_offd[ekj] /= v[j];
        }
      }
      for (int i = 0; i < N; ++i) {
        diag_g[i] = R(1) / inv_d[i];
        site_log += std::log(std::fabs((double)inv_d[i]));
        if (inv_d[i] < 0) site_neg++;
      }
      // forward / backward substitution to the inverse (:738-807)
      for (int k = 0; k < N; ++k) {
        for (int i = 0; i < k; ++i) v[i] = 0;
        v[k] = cplx<R>(diag_g[k], 0);
        for (int i = k + 1; i < N; ++i) {
          v[i] = 0;
          for (int j = k; j < i; ++j) {
            int eij = i * (i - 1) / 2 + j;
            v[i] -= inv_offd[eij] * inv_d[j] * v[j];
          }
          v[i] *= diag_g[i];
        }
        for (int i = N - 2; i >= k; --i)
          for (int j = i + 1; j < N; ++j) {
            int eji = j * (j - 1) / 2 + i;
            v[i] -= std::conj(inv_offd[eji]) * v[j];
          }
        inv_d[k] = std::real(v[k]);
        for (int i = k + 1; i < N; ++i) {
          int eik = i * (i - 1) / 2 + k;
          inv_offd[eik] = v[i];
        }
      }
      for (int i = 0; i < N; ++i) tri[s].diag[block][i] = inv_d[i];
      for (int i = 0; i < 15; ++i) tri[s].offd[block][i] = inv_offd[i];
    }
    logdet += site_log;
    if (site_neg) nneg++;
  }
  if (nneg_sites) *nneg_sites = nneg;
  return logdet;
}

// chi = (L + D + L^dag) psi per chiral block on checkerboard cb.  applySite, clover_term_qdp_w.h:
// 1056-1162: block 0 acts on fermion components 0..5 (spins 0,1), block 1 on 6..11 (spins 2,3).
template <typename R> void clover_apply(const Geom& g, const Clov<R>* tri, cplx<R>* chi, const cplx<R>* psi, int cb) {
  const long lo = cb * g.Vh, hi = lo + g.Vh;
#pragma omp parallel for schedule(static)
  for (long s = lo; s < hi; ++s) {
    const cplx<R>* p = psi + s * 12;
    cplx<R> out[12];
    for (int b = 0; b < 2; ++b)
      for (int i = 0; i < 6; ++i) {
        cplx<R> acc = tri[s].diag[b][i] * p[b * 6 + i];
        for (int j = 0; j < 6; ++j) {
          if (j < i) acc += tri[s].offd[b][i * (i - 1) / 2 + j] * p[b * 6 + j];
          if (j > i) acc += std::conj(tri[s].offd[b][j * (j - 1) / 2 + i]) * p[b * 6 + j];
        }
        out[b * 6 + i] = acc;
      }
    for (int k = 0; k < 12; ++k) chi[s * 12 + k] = out[k];
  }
}

// ------------------------------------------------------------------------------------------------
// Clover force terms (CloverTermBase, clover_term_base_w.h).
//   deriv (:691-745): for every plane mu < nu,
//     Lambda(x) = -(1/8) c_munu traceSpin(outerProduct(Gamma(2^mu + 2^nu) psi, chi))(x)  on cb, 0 elsewhere,
//     c_munu = getCloverCoeff(mu, nu) (clover_term_qdp_w.h:1525-1547: the 1/2-scaled clovCoeffR, or
//     clovCoeffT when nu is the time direction), then deriv_loops(mu, nu, cb, Lambda) -> ds_mu, ds_nu and
//     ds[mu] += ds_mu, ds[nu] -= ds_nu; the clover term is Hermitian, so isign does not enter.
//   derivTrLn (:825-870): the same with Lambda = -(1/8) c_munu triacntr(2^mu + 2^nu) of the (inverted)
//     clover term on cb, triacntr(mat)(x) = traceSpin(Gamma(mat) A(x)) (clover_term_qdp_w.h:1213-1500:
//     B_ij = sum_{s,t} Gamma_st A_(t,i),(s,j) for the six mu-nu matrices).
// deriv_loops (:320-683): the 16 insertions of Lambda into the four clover leaves of the plane, with
// Lambda on cb only; the corners and staples on all sites, then the terms on cb and on 1-cb, and the
// two accumulators shifted back by nu and mu.
template <typename R> static inline void aam(cplx<R>* c, const cplx<R>* a, const cplx<R>* b) {  // a^dag b^dag
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cplx<R> t = 0;
      for (int k = 0; k < 3; ++k) t += std::conj(a[k * 3 + i]) * std::conj(b[j * 3 + k]);
      c[i * 3 + j] = t;
    }
}
template <typename R> static inline void adj3(cplx<R>* c, const cplx<R>* a) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c[i * 3 + j] = std::conj(a[j * 3 + i]);
}
template <typename R> static inline void mm3(cplx<R>* d, const cplx<R>* a, const cplx<R>* b, const cplx<R>* c) {  // a b c
  cplx<R> t[9];
  mm(t, a, b);
  mm(d, t, c);
}

template <typename R>
static void deriv_loops(const Geom& g, const cplx<R>* U, int mu, int nu, int cb, const cplx<R>* Lam, cplx<R>* ds_mu,
                        cplx<R>* ds_nu) {
  const long V = g.V;
  std::vector<cplx<R>> ulc(V * 9), urc(V * 9), lrc(V * 9), llc(V * 9), sfor(V * 9), sright(V * 9), sleft(V * 9),
      sback(V * 9), tmu(V * 9), tnu(V * 9);
#pragma omp parallel for schedule(static)
  for (long s = 0; s < V; ++s) {
    const cplx<R>* unu_fmu = link(U, V, nu, g.fwd(s, mu));  // u_nu_for_mu
    const cplx<R>* umu_fnu = link(U, V, mu, g.fwd(s, nu));  // u_mu_for_nu
    const cplx<R>* umu = link(U, V, mu, s);
    const cplx<R>* unu = link(U, V, nu, s);
    aam(&ulc[s * 9], umu_fnu, unu);                 // up_left_corner  = adj(u_mu_for_nu) adj(u_nu)
    mma(&urc[s * 9], unu_fmu, umu_fnu);             // up_right_corner = u_nu_for_mu adj(u_mu_for_nu)
    aam(&lrc[s * 9], unu_fmu, umu);                 // low_right_corner = adj(u_nu_for_mu) adj(u_mu)
    amm(&llc[s * 9], umu, unu);                     // low_left_corner = adj(u_mu) u_nu
    mm(&sfor[s * 9], unu_fmu, &ulc[s * 9]);         // staple_for
    mm(&sright[s * 9], &ulc[s * 9], umu);           // staple_right
    mm(&sleft[s * 9], umu_fnu, &lrc[s * 9]);        // staple_left
    amm(&sback[s * 9], unu_fmu, &llc[s * 9]);       // staple_back = adj(u_nu_for_mu) low_left_corner
  }
#pragma omp parallel for schedule(static)
  for (long s = 0; s < V; ++s) {
    const long sm = g.fwd(s, mu), sn = g.fwd(s, nu), smn = g.fwd(sm, nu);
    const cplx<R>* Lx = Lam + s * 9;
    const cplx<R>* Lxm = Lam + sm * 9;
    const cplx<R>* Lxn = Lam + sn * 9;
    const cplx<R>* Lxmn = Lam + smn * 9;
    const cplx<R>* unu_fmu = link(U, V, nu, sm);
    const cplx<R>* umu_fnu = link(U, V, mu, sn);
    const cplx<R>* umu = link(U, V, mu, s);
    const cplx<R>* unu = link(U, V, nu, s);
    cplx<R> a[9], b[9], c[9], d[9];
    if (g.parity(s) == cb) {
      mm3(a, unu_fmu, Lxmn, &ulc[s * 9]);           // 1) mu
      mm(b, &sfor[s * 9], Lx);                      // 5) mu
      for (int k = 0; k < 9; ++k) ds_mu[s * 9 + k] = a[k] + b[k];
      mm3(a, &ulc[s * 9], Lx, umu);                 // 2) nu
      mm(b, Lxmn, &sright[s * 9]);                  // 6) nu
      for (int k = 0; k < 9; ++k) tnu[s * 9 + k] = a[k] + b[k];
      mm3(a, &lrc[s * 9], Lx, unu);                 // 4) mu
      mm(b, Lxmn, &sback[s * 9]);                   // 8) mu
      for (int k = 0; k < 9; ++k) tmu[s * 9 + k] = a[k] + b[k];
      mm3(a, umu_fnu, Lxmn, &lrc[s * 9]);           // 4) nu
      mm(b, &sleft[s * 9], Lx);                     // 7) nu
      for (int k = 0; k < 9; ++k) ds_nu[s * 9 + k] = a[k] + b[k];
    } else {
      adj3(c, umu_fnu);
      adj3(d, &llc[s * 9]);
      mm3(a, c, Lxn, d);                            // 1) nu
      mm(b, &sright[s * 9], Lxm);                   // 5) nu
      for (int k = 0; k < 9; ++k) tnu[s * 9 + k] = a[k] + b[k];
      adj3(c, unu);
      mm3(a, &urc[s * 9], Lxn, c);                  // 2) mu
      mm(b, Lxm, &sfor[s * 9]);                     // 6) mu
      for (int k = 0; k < 9; ++k) ds_mu[s * 9 + k] = a[k] + b[k];
      adj3(c, unu_fmu);
      mm3(a, c, Lxm, &llc[s * 9]);                  // 3) mu
      mm(b, &sback[s * 9], Lxn);                    // 7) mu
      for (int k = 0; k < 9; ++k) tmu[s * 9 + k] = a[k] + b[k];
      adj3(c, &urc[s * 9]);
      adj3(d, umu);
      mm3(a, c, Lxm, d);                            // 3) nu
      mm(b, Lxn, &sleft[s * 9]);                    // 8) nu
      for (int k = 0; k < 9; ++k) ds_nu[s * 9 + k] = a[k] + b[k];
    }
  }
  // ds_mu -= shift(tmp_mu, BACKWARD, nu); ds_nu -= shift(tmp_nu, BACKWARD, mu)
#pragma omp parallel for schedule(static)
  for (long s = 0; s < V; ++s) {
    const long bn = g.bwd(s, nu), bm = g.bwd(s, mu);
    for (int k = 0; k < 9; ++k) {
      ds_mu[s * 9 + k] -= tmu[bn * 9 + k];
      ds_nu[s * 9 + k] -= tnu[bm * 9 + k];
    }
  }
}

static inline double clover_coeff(int mu, int nu, double csw_r, double csw_t) {  // getCloverCoeff, 1/2 scaled
  return 0.5 * ((mu == 3 || nu == 3) ? csw_t : csw_r);
}

static void gamma_munu(cd G[4][4], int mu, int nu) {
  for (int s = 0; s < 4; ++s)
    for (int t = 0; t < 4; ++t) {
      cd a = 0;
      for (int k = 0; k < 4; ++k) a += kGamma[mu][s][k] * kGamma[nu][k][t];
      G[s][t] = a;
    }
}

// ds += the planes' insertions of Lambda_munu (given per plane by lam(mu, nu, Lambda[V*9]))
template <typename R, typename F>
static void clover_force(const Geom& g, const cplx<R>* U, int cb, cplx<R>* ds, F&& lam) {
  const long V = g.V;
  for (long k = 0; k < 4 * V * 9; ++k) ds[k] = 0;
  std::vector<cplx<R>> L(V * 9), dmu(V * 9), dnu(V * 9);
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu + 1; nu < 4; ++nu) {
      std::fill(L.begin(), L.end(), cplx<R>(0));
      lam(mu, nu, L.data());
      deriv_loops<R>(g, U, mu, nu, cb, L.data(), dmu.data(), dnu.data());
      for (long k = 0; k < V * 9; ++k) {
        ds[(long)mu * V * 9 + k] += dmu[k];
        ds[(long)nu * V * 9 + k] -= dnu[k];
      }
    }
}

template <typename R>
void clover_deriv(const Geom& g, const cplx<R>* U, double csw_r, double csw_t, const cplx<R>* chi, const cplx<R>* psi,
                  int cb, cplx<R>* ds) {
  clover_force<R>(g, U, cb, ds, [&](int mu, int nu, cplx<R>* L) {
    cd G[4][4];
    gamma_munu(G, mu, nu);
    const R f = R(-clover_coeff(mu, nu, csw_r, csw_t) / 8.0);
    for (long s = cb * g.Vh; s < (cb + 1) * g.Vh; ++s) {
      cplx<R> gp[4][3];  // Gamma psi
      for (int sp = 0; sp < 4; ++sp)
        for (int c = 0; c < 3; ++c) {
          cplx<R> t = 0;
          for (int k = 0; k < 4; ++k) t += cplx<R>(G[sp][k]) * psi[s * 12 + k * 3 + c];
          gp[sp][c] = t;
        }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          cplx<R> t = 0;
          for (int sp = 0; sp < 4; ++sp) t += gp[sp][i] * std::conj(chi[s * 12 + sp * 3 + j]);
          L[s * 9 + i * 3 + j] = f * t;
        }
    }
  });
}

// the 12x12 (spin x colour) matrix of a packed clover site (two Hermitian 6x6 chiral blocks)
template <typename R> static void clover_dense(const Clov<R>& t, cplx<R> A[12][12]) {
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j < 12; ++j) A[i][j] = 0;
  for (int b = 0; b < 2; ++b) {
    for (int i = 0; i < 6; ++i) A[6 * b + i][6 * b + i] = t.diag[b][i];
    for (int i = 1; i < 6; ++i)
      for (int j = 0; j < i; ++j) {
        const cplx<R> v = t.offd[b][i * (i - 1) / 2 + j];
        A[6 * b + i][6 * b + j] = v;
        A[6 * b + j][6 * b + i] = std::conj(v);
      }
  }
}

template <typename R>
void clover_deriv_trln(const Geom& g, const cplx<R>* U, double csw_r, double csw_t, const Clov<R>* tri_inv, int cb,
                       cplx<R>* ds) {
  clover_force<R>(g, U, cb, ds, [&](int mu, int nu, cplx<R>* L) {
    cd G[4][4];
    gamma_munu(G, mu, nu);
    const R f = R(-clover_coeff(mu, nu, csw_r, csw_t) / 8.0);
    for (long s = cb * g.Vh; s < (cb + 1) * g.Vh; ++s) {
      cplx<R> A[12][12];
      clover_dense(tri_inv[s], A);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          cplx<R> t = 0;  // triacntr: B_ij = sum_{s,t} Gamma_st A_(t,i),(s,j)
          for (int sp = 0; sp < 4; ++sp)
            for (int tp = 0; tp < 4; ++tp) t += cplx<R>(G[sp][tp]) * A[tp * 3 + i][sp * 3 + j];
          L[s * 9 + i * 3 + j] = f * t;
        }
    }
  });
}

// ------------------------------------------------------------------------------------------------
// BLAS over a checkerboard range.  QDP++ norm2 / innerProduct return Double / DComplex.
struct Range {
  long lo, hi;  // site range
};
template <typename R> double norm2(const cplx<R>* x, Range r) {
  double acc = 0;
#pragma omp parallel for reduction(+ : acc) schedule(static)
  for (long k = r.lo * 12; k < r.hi * 12; ++k) acc += (double)std::norm(x[k]);
  return acc;
}
template <typename R> cd inner(const cplx<R>* x, const cplx<R>* y, Range r) {  // <x|y> = sum conj(x) y
  double re = 0, im = 0;
#pragma omp parallel for reduction(+ : re, im) schedule(static)
  for (long k = r.lo * 12; k < r.hi * 12; ++k) {
    cd a = cd(x[k].real(), x[k].imag()), b = cd(y[k].real(), y[k].imag());
    cd p = std::conj(a) * b;
    re += p.real();
    im += p.imag();
  }
  return cd(re, im);
}
// y = a*x + b*y over the range (a, b given in the precision of the fields, as RT ar = a in Chroma)
template <typename R> void axpby(cplx<R>* y, cplx<R> a, const cplx<R>* x, cplx<R> b, Range r) {
#pragma omp parallel for schedule(static)
  for (long k = r.lo * 12; k < r.hi * 12; ++k) y[k] = a * x[k] + b * y[k];
}
template <typename R> void copy(cplx<R>* y, const cplx<R>* x, Range r) {
  std::memcpy(y + r.lo * 12, x + r.lo * 12, sizeof(cplx<R>) * 12 * (r.hi - r.lo));
}
template <typename R> void zero(cplx<R>* y, Range r) { std::memset((void*)(y + r.lo * 12), 0, sizeof(cplx<R>) * 12 * (r.hi - r.lo)); }

// ------------------------------------------------------------------------------------------------
// Linear operators (lib/linearop.h:25-49, lib/eoprec_linop.h:90-165)
using Field = std::vector<cd>;

template <typename R> struct LinOp {
  using F = std::vector<cplx<R>>;
  const Geom* g;
  virtual ~LinOp() {}
  virtual void apply(cplx<R>* chi, const cplx<R>* psi, int isign) const = 0;
  virtual Range subset() const = 0;
};

template <typename R> struct EOLinOp : LinOp<R> {
  using typename LinOp<R>::F;
  Range subset() const override { return Range{this->g->Vh, this->g->V}; }  // rb[1]
  virtual void evenEvenLinOp(cplx<R>* chi, const cplx<R>* psi, int isign) const = 0;
  virtual void evenEvenInvLinOp(cplx<R>* chi, const cplx<R>* psi, int isign) const = 0;
  virtual void evenOddLinOp(cplx<R>* chi, const cplx<R>* psi, int isign) const = 0;
  virtual void oddEvenLinOp(cplx<R>* chi, const cplx<R>* psi, int isign) const = 0;
  virtual void oddOddLinOp(cplx<R>* chi, const cplx<R>* psi, int isign) const = 0;
  // eoprec_linop.h:125-145
  void apply(cplx<R>* chi, const cplx<R>* psi, int isign) const override {
    F t1(this->g->V * 12), t2(this->g->V * 12);
    evenOddLinOp(t1.data(), psi, isign);
    evenEvenInvLinOp(t2.data(), t1.data(), isign);
    oddEvenLinOp(t1.data(), t2.data(), isign);
    oddOddLinOp(chi, psi, isign);
    axpby<R>(chi, R(-1), t1.data(), R(1), subset());
  }
  // eoprec_linop.h:147-165
  void unprecLinOp(cplx<R>* chi, const cplx<R>* psi, int isign) const {
    const long V = this->g->V, Vh = this->g->Vh;
    F t1(V * 12), t2(V * 12);
    evenEvenLinOp(t1.data(), psi, isign);
    evenOddLinOp(t2.data(), psi, isign);
    for (long k = 0; k < Vh * 12; ++k) chi[k] = t1[k] + t2[k];
    oddEvenLinOp(t1.data(), psi, isign);
    oddOddLinOp(t2.data(), psi, isign);
    for (long k = Vh * 12; k < V * 12; ++k) chi[k] = t1[k] + t2[k];
  }
};

// EvenOddPrecCloverLinOp, eoprec_clover_linop_w.cc:19-187 (no twisted mass)
template <typename R> struct EOClover : EOLinOp<R> {
  const cplx<R>* U;
  std::vector<Clov<R>> clov, invclov;
  double logdet_e = 0;
  long nneg = 0;
  EOClover(const Geom& g_, const cplx<R>* U_, R mass, R csw_r, R csw_t) : U(U_) {
    this->g = &g_;
    clov.resize(g_.V);
    make_clover<R>(g_, U, csw_r, csw_t, mass, clov.data());  // clov.create
    invclov = clov;                                          // invclov.create(fs, param, clov)
    logdet_e = clover_invert<R>(g_, invclov.data(), 0, &nneg);  // invclov.choles(0)
  }
  void hop(cplx<R>* chi, const cplx<R>* psi, int isign, int cb, R scale) const {
    dslash<R>(*this->g, U, chi, psi, isign, cb);
    Range r{cb * this->g->Vh, (cb + 1) * this->g->Vh};
    axpby<R>(chi, R(0), chi, scale, r);
  }
  void evenEvenLinOp(cplx<R>* c, const cplx<R>* p, int) const override { clover_apply<R>(*this->g, clov.data(), c, p, 0); }
  void evenEvenInvLinOp(cplx<R>* c, const cplx<R>* p, int) const override { clover_apply<R>(*this->g, invclov.data(), c, p, 0); }
  void oddOddLinOp(cplx<R>* c, const cplx<R>* p, int) const override { clover_apply<R>(*this->g, clov.data(), c, p, 1); }
  void evenOddLinOp(cplx<R>* c, const cplx<R>* p, int isign) const override { hop(c, p, isign, 0, R(-0.5)); }
  void oddEvenLinOp(cplx<R>* c, const cplx<R>* p, int isign) const override { hop(c, p, isign, 1, R(-0.5)); }
  // eoprec_clover_linop_w.cc:142-187: tmp1 = D(cb0) psi; tmp2 = A_ee^-1 tmp1; tmp1 = D(cb1) tmp2;
  // chi = A_oo psi; chi_o += -1/4 tmp1
  void apply(cplx<R>* chi, const cplx<R>* psi, int isign) const override {
    const Geom& g = *this->g;
    typename LinOp<R>::F t1(g.V * 12), t2(g.V * 12);
    dslash<R>(g, U, t1.data(), psi, isign, 0);
    clover_apply<R>(g, invclov.data(), t2.data(), t1.data(), 0);
    dslash<R>(g, U, t1.data(), t2.data(), isign, 1);
    clover_apply<R>(g, clov.data(), chi, psi, 1);
    axpby<R>(chi, R(-0.25), t1.data(), R(1), this->subset());
  }
};

// EvenOddPrecLogDetLinearOperator::deriv (eoprec_logdet_linop.h:114-157) of the EO clover operator with
// its pieces (eoprec_clover_linop_w.cc:192-260): derivOddOdd / derivEvenEven = clov.deriv(cb 1 / 0),
// derivOddEven / derivEvenOdd = -1/2 D.deriv(cb 1 / 0).  chi, psi full-lattice (odd part used); ds [4][V][3][3].
//   ds  = chi^dag A'_oo psi
//   ds -= chi^dag D'_oe A_ee^-1 D_eo psi                         (tmp2 = A_ee^-1 (-1/2 D_eo psi))
//   ds += chi^dag D_oe A_ee^-1 A'_ee A_ee^-1 D_eo psi           (tmp3 = A_ee^-1 (-1/2 D_eo(msign) chi))
//   ds -= chi^dag D_oe A_ee^-1 D'_eo psi
template <typename R>
void eo_clover_deriv(const EOClover<R>& M, double csw_r, double csw_t, const cplx<R>* chi, const cplx<R>* psi,
                     int isign, cplx<R>* ds) {
  const Geom& g = *M.g;
  const long V = g.V;
  const int msign = -isign;
  std::vector<cplx<R>> t1(V * 12), t2(V * 12), t3(V * 12), d1((size_t)4 * V * 9);
  clover_deriv<R>(g, M.U, csw_r, csw_t, chi, psi, 1, ds);                 // derivOddOddLinOp
  M.evenOddLinOp(t1.data(), psi, isign);
  M.evenEvenInvLinOp(t2.data(), t1.data(), isign);
  dslash_deriv<R>(g, d1.data(), chi, t2.data(), isign, 1);                // derivOddEvenLinOp / (-1/2)
  for (size_t k = 0; k < d1.size(); ++k) ds[k] -= R(-0.5) * d1[k];
  M.evenOddLinOp(t1.data(), chi, msign);
  M.evenEvenInvLinOp(t3.data(), t1.data(), msign);
  clover_deriv<R>(g, M.U, csw_r, csw_t, t3.data(), t2.data(), 0, d1.data());  // derivEvenEvenLinOp
  for (size_t k = 0; k < d1.size(); ++k) ds[k] += d1[k];
  dslash_deriv<R>(g, d1.data(), t3.data(), psi, isign, 0);                // derivEvenOddLinOp / (-1/2)
  for (size_t k = 0; k < d1.size(); ++k) ds[k] -= R(-0.5) * d1[k];
}

// EvenOddPrecWilsonLinOp, eoprec_wilson_linop_w.cc:48-166 and eoprec_wilson_linop_w.h:75-110
template <typename R> struct EOWilson : EOLinOp<R> {
  const cplx<R>* U;
  R fact, invfact;
  EOWilson(const Geom& g_, const cplx<R>* U_, R mass) : U(U_) {
    this->g = &g_;
    fact = mass + R(4);
    invfact = R(1) / fact;
  }
  void scale(cplx<R>* c, const cplx<R>* p, R f, int cb) const {
    Range r{cb * this->g->Vh, (cb + 1) * this->g->Vh};
    axpby<R>(c, f, p, R(0), r);
  }
  void evenEvenLinOp(cplx<R>* c, const cplx<R>* p, int) const override { scale(c, p, fact, 0); }
  void evenEvenInvLinOp(cplx<R>* c, const cplx<R>* p, int) const override { scale(c, p, invfact, 0); }
  void oddOddLinOp(cplx<R>* c, const cplx<R>* p, int) const override { scale(c, p, fact, 1); }
  void evenOddLinOp(cplx<R>* c, const cplx<R>* p, int isign) const override {
    dslash<R>(*this->g, U, c, p, isign, 0);
    scale(c, c, R(-0.5), 0);
  }
  void oddEvenLinOp(cplx<R>* c, const cplx<R>* p, int isign) const override {
    dslash<R>(*this->g, U, c, p, isign, 1);
    scale(c, c, R(-0.5), 1);
  }
  void apply(cplx<R>* chi, const cplx<R>* psi, int isign) const override {
    const Geom& g = *this->g;
    typename LinOp<R>::F t1(g.V * 12), t2(g.V * 12);
    dslash<R>(g, U, t1.data(), psi, isign, 0);
    dslash<R>(g, U, t2.data(), t1.data(), isign, 1);
    const Range o = this->subset();
    for (long k = o.lo * 12; k < o.hi * 12; ++k) chi[k] = fact * psi[k] + (R(-0.25) * invfact) * t2[k];
  }
};

// UnprecWilsonLinOp, unprec_wilson_linop_w.cc:32-77: chi = (Nd+m) psi - 1/2 D psi on all sites
template <typename R> struct UnprecWilson : LinOp<R> {
  const cplx<R>* U;
  R fact;
  UnprecWilson(const Geom& g_, const cplx<R>* U_, R mass) : U(U_) {
    this->g = &g_;
    fact = R(4) + mass;
  }
  Range subset() const override { return Range{0, this->g->V}; }
  void apply(cplx<R>* chi, const cplx<R>* psi, int isign) const override {
    const Geom& g = *this->g;
    typename LinOp<R>::F t(g.V * 12);
    dslash<R>(g, U, t.data(), psi, isign, 0);
    dslash<R>(g, U, t.data(), psi, isign, 1);
    for (long k = 0; k < g.V * 12; ++k) chi[k] = fact * psi[k] + R(-0.5) * t[k];
  }
};

// ------------------------------------------------------------------------------------------------
// Solvers
struct Result {
  int n_count;
  double resid;
};

// InvCG2_a, invcg2.cc:70-232.  CG on A = M^dag M.  Optional trace of |r|^2 per iteration.
template <typename R>
Result invcg2(const LinOp<R>& M, const cplx<R>* chi, cplx<R>* psi, double RsdCG, int MaxCG, std::vector<double>* trace) {
  const Range s = M.subset();
  const long n = M.g->V * 12;
  std::vector<cplx<R>> mp(n), mmp(n), p(n), r(n);
  Result res{0, 0};
  double chi_sq = norm2<R>(chi, s);
  double rsd_sq = RsdCG * RsdCG * chi_sq;
  M.apply(mp.data(), psi, +1);
  M.apply(mmp.data(), mp.data(), -1);
  for (long k = s.lo * 12; k < s.hi * 12; ++k) r[k] = chi[k] - mmp[k];
  double cp = norm2<R>(r.data(), s);
  if (trace) trace->push_back(cp);
  copy<R>(p.data(), r.data(), s);
  if (cp <= rsd_sq) {
    res.n_count = 0;
    res.resid = std::sqrt(cp);
    return res;
  }
  for (int k = 1; k <= MaxCG; ++k) {
    double c = cp;
    M.apply(mp.data(), p.data(), +1);
    double d = norm2<R>(mp.data(), s);
    M.apply(mmp.data(), mp.data(), -1);
    double a = c / d;
    R ar = R(a);
    axpby<R>(r.data(), -ar, mmp.data(), R(1), s);
    cp = norm2<R>(r.data(), s);
    if (trace) trace->push_back(cp);
    axpby<R>(psi, ar, p.data(), R(1), s);
    if (cp <= rsd_sq) {
      res.n_count = k;
      M.apply(mp.data(), psi, +1);
      M.apply(mmp.data(), mp.data(), -1);
      for (long q = s.lo * 12; q < s.hi * 12; ++q) mmp[q] = chi[q] - mmp[q];
      res.resid = std::sqrt(norm2<R>(mmp.data(), s));
      return res;
    }
    double b = cp / c;
    R br = R(b);
    axpby<R>(p.data(), R(1), r.data(), br, s);
  }
  res.n_count = MaxCG;
  res.resid = std::sqrt(cp);
  return res;
}

// MInvCG2_a, minvcg2.cc:72-355: multishift CG for (M^dag M + shifts[s]) psi[s] = chi with one Krylov
// space (Jegerlehner, hep-lat/9708029).  Every psi[s] starts at zero (:118-121); a zero source returns
// at once (:133-153, fuzz 1e-10 in double); the first test uses the smallest shift (:99-104, :247);
// shift s stops updating once c z_s^2 < RsdCG[s]^2 |chi|^2 (:326-339).  Returns the iteration count.
template <typename R>
int minvcg2(const LinOp<R>& M, const cplx<R>* chi, std::vector<std::vector<cplx<R>>>& psi,
            const std::vector<double>& shifts, const std::vector<double>& RsdCG, int MaxCG) {
  const Range sub = M.subset();
  const long n = M.g->V * 12;
  const int ns = (int)shifts.size();
  int isz = 0;
  for (int s = 1; s < ns; ++s)
    if (shifts[s] < shifts[isz]) isz = s;
  psi.assign(ns, std::vector<cplx<R>>(n, cplx<R>(0)));
  const double chi_norm_sq = norm2<R>(chi, sub);
  if (std::sqrt(chi_norm_sq) < 1e-10) return 0;
  std::vector<double> rsd_sq(ns);
  double cp = chi_norm_sq;
  for (int s = 0; s < ns; ++s) rsd_sq[s] = cp * RsdCG[s] * RsdCG[s];
  std::vector<cplx<R>> r(n), p0(n), Mp(n), MMp(n);
  std::vector<std::vector<cplx<R>>> p(ns, std::vector<cplx<R>>(n));
  copy<R>(r.data(), chi, sub);
  copy<R>(p0.data(), chi, sub);
  for (int s = 0; s < ns; ++s) copy<R>(p[s].data(), chi, sub);
  M.apply(Mp.data(), p0.data(), +1);
  double d = norm2<R>(Mp.data(), sub);
  M.apply(MMp.data(), Mp.data(), -1);
  double b = -cp / d;
  axpby<R>(r.data(), R(b), MMp.data(), R(1), sub);  // r += b M^dag M p0
  std::vector<double> bs(ns), z[2] = {std::vector<double>(ns), std::vector<double>(ns)};
  int iz = 1;
  for (int s = 0; s < ns; ++s) {
    z[1 - iz][s] = 1.0;
    z[iz][s] = 1.0 / (1.0 - shifts[s] * b);
    bs[s] = b * z[iz][s];
  }
  for (int s = 0; s < ns; ++s) axpby<R>(psi[s].data(), R(-bs[s]), chi, R(0), sub);  // psi = -bs chi
  double c = norm2<R>(r.data(), sub);
  std::vector<char> conv(ns, 0);
  bool convP = c < rsd_sq[isz];
  int k = 1, n_count = 0;
  for (; k <= MaxCG && !convP; ++k) {
    const double a = c / cp;
    axpby<R>(p0.data(), R(1), r.data(), R(a), sub);  // p0 = r + a p0
    for (int s = 0; s < ns; ++s)
      if (!conv[s]) {
        const double as = a * z[iz][s] * bs[s] / (z[1 - iz][s] * b);
        axpby<R>(p[s].data(), R(z[iz][s]), r.data(), R(as), sub);  // p_s = z_s r + as p_s
      }
    cp = c;
    M.apply(Mp.data(), p0.data(), +1);
    d = norm2<R>(Mp.data(), sub);
    M.apply(MMp.data(), Mp.data(), -1);
    const double bp = b;
    b = -cp / d;
    axpby<R>(r.data(), R(b), MMp.data(), R(1), sub);
    c = norm2<R>(r.data(), sub);
    iz = 1 - iz;
    for (int s = 0; s < ns; ++s)
      if (!conv[s]) {
        const double z0 = z[1 - iz][s], z1 = z[iz][s];
        z[iz][s] = z0 * z1 * bp / (b * a * (z1 - z0) + z1 * bp * (1.0 - shifts[s] * b));
        bs[s] = b * z[iz][s] / z0;
      }
    for (int s = 0; s < ns; ++s)
      if (!conv[s]) axpby<R>(psi[s].data(), R(-bs[s]), p[s].data(), R(1), sub);  // psi_s -= bs p_s
    convP = true;
    for (int s = 0; s < ns; ++s) {
      if (!conv[s]) conv[s] = c * z[iz][s] * z[iz][s] < rsd_sq[s];
      convP = convP && conv[s];
    }
    n_count = k;
  }
  return n_count;
}

// LinOpSysSolverCG::operator(), syssolver_linop_cg.h:57-96: chi_tmp = M^dag chi; CG2; resid = |chi - M psi|
template <typename R>
Result syssolver_cg(const LinOp<R>& A, cplx<R>* psi, const cplx<R>* chi, double RsdCG, int MaxCG, std::vector<double>* trace) {
  const Range s = A.subset();
  const long n = A.g->V * 12;
  std::vector<cplx<R>> chi_tmp(n), tmp(n);
  A.apply(chi_tmp.data(), chi, -1);
  Result res = invcg2<R>(A, chi_tmp.data(), psi, RsdCG, MaxCG, trace);
  A.apply(tmp.data(), psi, +1);
  for (long k = s.lo * 12; k < s.hi * 12; ++k) tmp[k] = chi[k] - tmp[k];
  res.resid = std::sqrt(norm2<R>(tmp.data(), s));
  return res;
}

// InvBiCGStab_a, invbicgstab.cc:10-202.  Breakdowns return n_count = -1 (Chroma calls QDP_abort).
template <typename R>
Result invbicgstab(const LinOp<R>& A, const cplx<R>* chi, cplx<R>* psi, double Rsd, int MaxIt, int isign,
                   std::vector<double>* trace) {
  const Range s = A.subset();
  const long n = A.g->V * 12;
  Result ret{MaxIt, 0};
  double chi_sq = norm2<R>(chi, s);
  double rsd_sq = Rsd * Rsd * chi_sq;
  std::vector<cplx<R>> r(n), r0(n), p(n), v(n), tmp(n), t(n);
  A.apply(r0.data(), psi, isign);
  for (long k = s.lo * 12; k < s.hi * 12; ++k) r[k] = chi[k] - r0[k];
  copy<R>(r0.data(), r.data(), s);
  zero<R>(p.data(), s);
  zero<R>(v.data(), s);
  cd rho, rho_prev = 1.0, alpha = 1.0, omega = 1.0;
  bool conv = false;
  for (int k = 1; k <= MaxIt && !conv; ++k) {
    rho = inner<R>(r0.data(), r.data(), s);
    if (rho == cd(0)) return Result{-1, 0};
    cd beta = (rho / rho_prev) * (alpha / omega);
    cplx<R> omega_r(R(omega.real()), R(omega.imag())), beta_r(R(beta.real()), R(beta.imag()));
    for (long q = s.lo * 12; q < s.hi * 12; ++q) tmp[q] = p[q] - omega_r * v[q];
    for (long q = s.lo * 12; q < s.hi * 12; ++q) p[q] = r[q] + beta_r * tmp[q];
    A.apply(v.data(), p.data(), isign);
    cd ctmp = inner<R>(r0.data(), v.data(), s);
    if (ctmp == cd(0)) return Result{-1, 0};
    alpha = rho / ctmp;
    rho_prev = rho;
    cplx<R> alpha_r(R(alpha.real()), R(alpha.imag()));
    for (long q = s.lo * 12; q < s.hi * 12; ++q) r[q] -= alpha_r * v[q];
    A.apply(t.data(), r.data(), isign);
    double t_norm = norm2<R>(t.data(), s);
    if (t_norm == 0) return Result{-1, 0};
    omega = inner<R>(t.data(), r.data(), s) / t_norm;
    omega_r = cplx<R>(R(omega.real()), R(omega.imag()));
    alpha_r = cplx<R>(R(alpha.real()), R(alpha.imag()));
    for (long q = s.lo * 12; q < s.hi * 12; ++q) tmp[q] = psi[q] + omega_r * r[q];
    for (long q = s.lo * 12; q < s.hi * 12; ++q) psi[q] = tmp[q] + alpha_r * p[q];
    for (long q = s.lo * 12; q < s.hi * 12; ++q) r[q] -= omega_r * t[q];
    double r_norm = norm2<R>(r.data(), s);
    if (trace) trace->push_back(r_norm);
    if (r_norm < rsd_sq) {
      conv = true;
      ret.resid = std::sqrt(r_norm);
      ret.n_count = k;
    }
  }
  return ret;
}

// RelInvCG_a, reliable_cg.cc:10-188: FP32 (AF) inner CG on A^dag A with FP64 (A) reliable updates.
// Non-convergence returns n_count = -1 (Chroma calls QDP_abort, :180-183).
Result relcg(const LinOp<double>& A, const LinOp<float>& AF, const cd* chi, cd* psi, double RsdCG, double Delta,
             int MaxCG, std::vector<double>* trace, int* n_updates) {
  using cf = std::complex<float>;
  const Range s = A.subset();
  const long n = A.g->V * 12;
  Result ret{0, 0};
  bool convP = false;
  std::vector<cf> r(n), x(n), p(n), mp(n), mmp(n);
  std::vector<cd> b(n), r_dble(n), x_dble(n), tmp1(n), tmp2(n);
  copy<double>(b.data(), chi, s);
  double chi_norm = norm2<double>(chi, s);
  double rsd_sq = RsdCG * RsdCG * chi_norm;
  A.apply(tmp1.data(), psi, +1);
  A.apply(tmp2.data(), tmp1.data(), -1);
  for (long k = s.lo * 12; k < s.hi * 12; ++k) b[k] -= tmp2[k];
  zero<float>(x.data(), s);
  for (long k = s.lo * 12; k < s.hi * 12; ++k) r[k] = cf(b[k]);
  double r_sq = norm2<float>(r.data(), s);
  double rNorm = std::sqrt(r_sq), r0Norm = rNorm, maxrx = rNorm, maxrr = rNorm;
  double c = 0, a, d;
  int k, nupd = 0;
  for (k = 0; k < MaxCG && !convP; k++) {
    if (k == 0) {
      copy<float>(p.data(), r.data(), s);
    } else {
      double beta = r_sq / c;
      axpby<float>(p.data(), 1.0f, r.data(), float(beta), s);
    }
    c = r_sq;
    AF.apply(mp.data(), p.data(), +1);
    d = norm2<float>(mp.data(), s);
    AF.apply(mmp.data(), mp.data(), -1);
    a = c / d;
    float ar = float(a);
    axpby<float>(x.data(), ar, p.data(), 1.0f, s);
    axpby<float>(r.data(), -ar, mmp.data(), 1.0f, s);
    r_sq = norm2<float>(r.data(), s);
    if (trace) trace->push_back(r_sq);
    rNorm = std::sqrt(r_sq);
    if (rNorm > maxrx) maxrx = rNorm;
    if (rNorm > maxrr) maxrr = rNorm;
    bool updateX = (rNorm < Delta * r0Norm && r0Norm <= maxrx);
    bool updateR = (rNorm < Delta * maxrr && r0Norm <= maxrr) || updateX;
    if (updateR) {
      nupd++;
      for (long q = s.lo * 12; q < s.hi * 12; ++q) x_dble[q] = cd(x[q]);
      A.apply(tmp1.data(), x_dble.data(), +1);
      A.apply(tmp2.data(), tmp1.data(), -1);
      for (long q = s.lo * 12; q < s.hi * 12; ++q) r_dble[q] = b[q] - tmp2[q];
      for (long q = s.lo * 12; q < s.hi * 12; ++q) r[q] = cf(r_dble[q]);
      r_sq = norm2<double>(r_dble.data(), s);
      rNorm = std::sqrt(r_sq);
      maxrr = rNorm;
      if (updateX) {
        for (long q = s.lo * 12; q < s.hi * 12; ++q) psi[q] += x_dble[q];
        zero<float>(x.data(), s);
        copy<double>(b.data(), r_dble.data(), s);
        r0Norm = rNorm;
        maxrx = rNorm;
      }
    }
    if (r_sq < rsd_sq) {
      for (long q = s.lo * 12; q < s.hi * 12; ++q) psi[q] += cd(x[q]);
      ret.resid = rNorm;
      ret.n_count = k;
      convP = true;
    }
  }
  if (n_updates) *n_updates = nupd;
  if (!convP) ret.n_count = -1;
  return ret;
}

// LinOpSysSolverReliableCGClover::operator(), syssolver_linop_rel_cg_clover.h:104-152:
// chi_tmp = M^dag chi; RelInvCG; resid = |chi - M psi| in FP64.
Result syssolver_relcg(const LinOp<double>& A, const LinOp<float>& AF, cd* psi, const cd* chi, double Rsd, double Delta,
                       int MaxCG, std::vector<double>* trace, int* n_updates) {
  const Range s = A.subset();
  const long n = A.g->V * 12;
  std::vector<cd> chi_tmp(n), tmp(n);
  A.apply(chi_tmp.data(), chi, -1);
  Result res = relcg(A, AF, chi_tmp.data(), psi, Rsd, Delta, MaxCG, trace, n_updates);
  A.apply(tmp.data(), psi, +1);
  for (long k = s.lo * 12; k < s.hi * 12; ++k) tmp[k] = chi[k] - tmp[k];
  res.resid = std::sqrt(norm2<double>(tmp.data(), s));
  return res;
}

// RelInvBiCGStab_a, reliable_bicgstab.cc:14-290 (the single/double instance, :323-335): BiCGStab
// on A (isign) with FP32 (AF) vectors r, r0, x, p, v, t and FP64 reliable updates of r and grouped
// x.  Non-convergence and breakdown return n_count = -1 (Chroma calls QDP_abort, :140-143, :275-278).
Result relbicgstab(const LinOp<double>& A, const LinOp<float>& AF, const cd* chi, cd* psi, double Rsd,
                   double Delta, int MaxIt, int isign, std::vector<double>* trace, int* n_updates) {
  using cf = std::complex<float>;
  const Range s = A.subset();
  const long n = A.g->V * 12;
  Result ret{0, 0};
  std::vector<cd> b(n), tmp(n), r_dble(n), x_dble(n);
  std::vector<cf> r(n), r0(n), x(n), p(n), v(n), t(n);
  const double rsd_sq = Rsd * Rsd * norm2<double>(chi, s);
  A.apply(tmp.data(), psi, isign);
  for (long q = s.lo * 12; q < s.hi * 12; ++q) b[q] = chi[q] - tmp[q];  // xymz_normx (:61-74)
  double r_sq = norm2<double>(b.data(), s);
  for (long q = s.lo * 12; q < s.hi * 12; ++q) r[q] = r0[q] = cf(b[q]);
  double rNorm = std::sqrt(r_sq), r0Norm = rNorm, maxrx = rNorm, maxrr = rNorm;
  cd rho = 1.0, rho_prev = 1.0, alpha = 1.0, omega = 1.0;
  int k, nupd = 0;
  bool convP = false;
  for (k = 0; k < MaxIt && !convP; k++) {
    if (k == 0) {  // (:109-113) <r0|r> = |r|^2 at the start
      rho = r_sq;
      copy<float>(p.data(), r.data(), s);
    } else {       // p = r + beta (p - omega v), beta = (rho/rho_prev)(alpha/omega)  (:115-132)
      const cd beta = (rho / rho_prev) * (alpha / omega);
      const cf beta_r(beta), omega_r(omega);
      for (long q = s.lo * 12; q < s.hi * 12; ++q) p[q] = r[q] + beta_r * (p[q] - omega_r * v[q]);
    }
    AF.apply(v.data(), p.data(), isign);
    const cd ctmp = inner<float>(r0.data(), v.data(), s);
    if (ctmp == cd(0)) return Result{-1, 0};
    alpha = rho / ctmp;
    rho_prev = rho;
    const cf alpha_r(alpha);
    for (long q = s.lo * 12; q < s.hi * 12; ++q) r[q] -= alpha_r * v[q];  // s = r - alpha v
    AF.apply(t.data(), r.data(), isign);
    const double t_norm = norm2<float>(t.data(), s);
    if (t_norm == 0) return Result{-1, 0};
    omega = inner<float>(t.data(), r.data(), s) / t_norm;
    const cf omega_r(omega);
    for (long q = s.lo * 12; q < s.hi * 12; ++q) x[q] = (x[q] + omega_r * r[q]) + alpha_r * p[q];  // xpaypbz
    for (long q = s.lo * 12; q < s.hi * 12; ++q) r[q] -= omega_r * t[q];  // xmay_normx_cdotzx (:199)
    r_sq = norm2<float>(r.data(), s);
    rho = inner<float>(r0.data(), r.data(), s);
    if (trace) trace->push_back(r_sq);
    rNorm = std::sqrt(r_sq);
    if (rNorm > maxrx) maxrx = rNorm;
    if (rNorm > maxrr) maxrr = rNorm;
    const bool updateX = (rNorm < Delta * r0Norm && r0Norm <= maxrx);  // (:213-214)
    const bool updateR = (rNorm < Delta * maxrr && r0Norm <= maxrr) || updateX;
    if (updateR) {  // (:216-256); rho is not recomputed from the new r
      nupd++;
      for (long q = s.lo * 12; q < s.hi * 12; ++q) x_dble[q] = cd(x[q]);
      A.apply(tmp.data(), x_dble.data(), isign);
      for (long q = s.lo * 12; q < s.hi * 12; ++q) r_dble[q] = b[q] - tmp[q];
      r_sq = norm2<double>(r_dble.data(), s);
      for (long q = s.lo * 12; q < s.hi * 12; ++q) r[q] = cf(r_dble[q]);
      rNorm = std::sqrt(r_sq);
      maxrr = rNorm;
      if (updateX) {
        for (long q = s.lo * 12; q < s.hi * 12; ++q) psi[q] += x_dble[q];
        zero<float>(x.data(), s);
        copy<double>(b.data(), r_dble.data(), s);
        r0Norm = rNorm;
        maxrx = rNorm;
      }
    }
    if (r_sq < rsd_sq) {  // (:259-271)
      convP = true;
      for (long q = s.lo * 12; q < s.hi * 12; ++q) psi[q] += cd(x[q]);
      ret.resid = rNorm;
      ret.n_count = k;
    }
  }
  if (n_updates) *n_updates = nupd;
  if (!convP) ret.n_count = -1;
  return ret;
}

// |chi - A psi| (isign = +1) or |chi - A^dag A psi| (mdagm) over the subset, in FP64
double true_resid(const LinOp<double>& A, const cd* psi, const cd* chi, bool mdagm) {
  const Range s = A.subset();
  const long n = A.g->V * 12;
  std::vector<cd> t1(n), t2(n);
  A.apply(t1.data(), psi, +1);
  if (mdagm) {
    A.apply(t2.data(), t1.data(), -1);
    t1.swap(t2);
  }
  for (long q = s.lo * 12; q < s.hi * 12; ++q) t1[q] = chi[q] - t1[q];
  return std::sqrt(norm2<double>(t1.data(), s));
}

}  // namespace orc

// ==================================================================================================
// C-ABI for the Python test harness and the CPU-baseline timer.
using namespace orc;

namespace {
thread_local std::string g_err;

struct OpHandle {
  std::unique_ptr<Geom> g;
  std::vector<cd> U;                      // FP64 links (BC already applied by the caller)
  std::vector<std::complex<float>> Uf;    // FP32 copy for the single-precision operator
  int kind;                               // 0 unprec Wilson, 1 EO Wilson, 2 EO clover
  std::unique_ptr<LinOp<double>> op;
  std::unique_ptr<LinOp<float>> opf;
};

template <typename Fn> int guard(Fn&& fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

int orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

void orc_site_coords(const int L[4], long idx, int x[4]) { Geom(L).coords(idx, x); }

// kind: 0 = UNPRECONDITIONED_WILSON, 1 = WILSON (even-odd), 2 = CLOVER (even-odd)
void* orc_op_create(int kind, const int L[4], const double* U, double mass, double csw_r, double csw_t) {
  OpHandle* h = nullptr;
  int rc = guard([&] {
    auto hh = std::make_unique<OpHandle>();
    hh->g = std::make_unique<Geom>(L);
    const long n = hh->g->V * 36;
    hh->U.assign(reinterpret_cast<const cd*>(U), reinterpret_cast<const cd*>(U) + n);
    hh->Uf.resize(n);
    for (long k = 0; k < n; ++k) hh->Uf[k] = std::complex<float>(hh->U[k]);
    hh->kind = kind;
    if (kind == 0) {
      hh->op = std::make_unique<UnprecWilson<double>>(*hh->g, hh->U.data(), mass);
      hh->opf = std::make_unique<UnprecWilson<float>>(*hh->g, hh->Uf.data(), float(mass));
    } else if (kind == 1) {
      hh->op = std::make_unique<EOWilson<double>>(*hh->g, hh->U.data(), mass);
      hh->opf = std::make_unique<EOWilson<float>>(*hh->g, hh->Uf.data(), float(mass));
    } else if (kind == 2) {
      hh->op = std::make_unique<EOClover<double>>(*hh->g, hh->U.data(), mass, csw_r, csw_t);
      hh->opf = std::make_unique<EOClover<float>>(*hh->g, hh->Uf.data(), float(mass), float(csw_r), float(csw_t));
    } else {
      throw std::runtime_error("unknown operator kind");
    }
    h = hh.release();
  });
  return rc ? nullptr : h;
}

void orc_op_destroy(void* h) { delete static_cast<OpHandle*>(h); }

// chi = M(isign) psi on the operator's subset; full-lattice host arrays
int orc_op_apply(void* h, double* chi, const double* psi, int isign) {
  return guard([&] {
    auto* o = static_cast<OpHandle*>(h);
    o->op->apply(reinterpret_cast<cd*>(chi), reinterpret_cast<const cd*>(psi), isign);
  });
}
int orc_op_apply_f32(void* h, float* chi, const float* psi, int isign) {
  return guard([&] {
    auto* o = static_cast<OpHandle*>(h);
    o->opf->apply(reinterpret_cast<std::complex<float>*>(chi), reinterpret_cast<const std::complex<float>*>(psi), isign);
  });
}

// EO pieces.  which: 0 evenEven, 1 evenEvenInv, 2 evenOdd, 3 oddEven, 4 oddOdd, 5 unprecLinOp
int orc_op_eo_piece(void* h, int which, double* chi, const double* psi, int isign) {
  return guard([&] {
    auto* o = static_cast<OpHandle*>(h);
    auto* eo = dynamic_cast<EOLinOp<double>*>(o->op.get());
    if (!eo) throw std::runtime_error("operator is not even-odd preconditioned");
    auto* c = reinterpret_cast<cd*>(chi);
    auto* p = reinterpret_cast<const cd*>(psi);
    switch (which) {
      case 0: eo->evenEvenLinOp(c, p, isign); break;
      case 1: eo->evenEvenInvLinOp(c, p, isign); break;
      case 2: eo->evenOddLinOp(c, p, isign); break;
      case 3: eo->oddEvenLinOp(c, p, isign); break;
      case 4: eo->oddOddLinOp(c, p, isign); break;
      case 5: eo->unprecLinOp(c, p, isign); break;
      default: throw std::runtime_error("bad piece");
    }
  });
}

// Raw hopping term D on output checkerboard cb (lwldslash_w.h:252-295)
int orc_dslash(const int L[4], const double* U, double* chi, const double* psi, int isign, int cb) {
  return guard([&] {
    Geom g(L);
    dslash<double>(g, reinterpret_cast<const cd*>(U), reinterpret_cast<cd*>(chi), reinterpret_cast<const cd*>(psi),
                   isign, cb);
  });
}
int orc_dslash_f32(const int L[4], const float* U, float* chi, const float* psi, int isign, int cb) {
  return guard([&] {
    Geom g(L);
    using cf = std::complex<float>;
    dslash<float>(g, reinterpret_cast<const cf*>(U), reinterpret_cast<cf*>(chi), reinterpret_cast<const cf*>(psi), isign,
                  cb);
  });
}

int orc_mesfield(const int L[4], const double* U, double* F) {
  return guard([&] {
    Geom g(L);
    mesfield<double>(g, reinterpret_cast<const cd*>(U), reinterpret_cast<cd*>(F));
  });
}

// S1, S2: [V][12][12] complex; out: [16][nmom][L_t] complex
int orc_mesons(const int L[4], const double* S1, const double* S2, const int origin[4], int nmom, const int* moms,
               double* out) {
  return guard([&] {
    Geom g(L);
    mesons(g, reinterpret_cast<const cd*>(S1), reinterpret_cast<const cd*>(S2), origin, nmom, moms,
           reinterpret_cast<cd*>(out));
  });
}

// nstep stout steps; U and out are [4][V][3][3] (may alias)
int orc_stout(const int L[4], const double* U, double* out, double rho, int nstep, int dirs) {
  return guard([&] {
    Geom g(L);
    const size_t n = (size_t)4 * g.V * 9;
    std::vector<cd> a(reinterpret_cast<const cd*>(U), reinterpret_cast<const cd*>(U) + n), b(n);
    for (int k = 0; k < nstep; ++k) {
      stout_step(g, a.data(), b.data(), rho, dirs);
      a.swap(b);
    }
    std::memcpy(out, a.data(), sizeof(cd) * n);
  });
}

// tri: V x 72 doubles (Clov<double>)
int orc_make_clover(const int L[4], const double* U, double csw_r, double csw_t, double mass, double* tri) {
  return guard([&] {
    Geom g(L);
    make_clover<double>(g, reinterpret_cast<const cd*>(U), csw_r, csw_t, mass, reinterpret_cast<Clov<double>*>(tri));
  });
}
int orc_clover_invert(const int L[4], double* tri, int cb, double* logdet, long* nneg) {
  return guard([&] {
    Geom g(L);
    *logdet = clover_invert<double>(g, reinterpret_cast<Clov<double>*>(tri), cb, nneg);
  });
}
int orc_clover_apply(const int L[4], const double* tri, double* chi, const double* psi, int cb) {
  return guard([&] {
    Geom g(L);
    clover_apply<double>(g, reinterpret_cast<const Clov<double>*>(tri), reinterpret_cast<cd*>(chi),
                         reinterpret_cast<const cd*>(psi), cb);
  });
}
// Sum of log|d_i| over the even sites of the operator's inverted clover (logDetEvenEvenLinOp)
int orc_op_logdet_even(void* h, double* out, long* nneg) {
  return guard([&] {
    auto* o = static_cast<OpHandle*>(h);
    auto* c = dynamic_cast<EOClover<double>*>(o->op.get());
    if (!c) throw std::runtime_error("not a clover operator");
    *out = c->logdet_e;
    *nneg = c->nneg;
  });
}

// solver: 0 CG_INVERTER (LinOpSysSolverCG), 1 BICGSTAB_INVERTER, 2 RELIABLE_CG (FP32 inner, FP64 updates),
//         3 RELIABLE_BICGSTAB (FP32 inner, FP64 updates)
// psi: in/out initial guess; trace (optional, may be NULL): |r|^2 per iteration, up to trace_cap entries
int orc_solve(void* h, int solver, double* psi, const double* chi, double rsd, int maxit, double delta, int* n_count,
              double* resid, double* trace, int trace_cap, int* trace_len) {
  return guard([&] {
    auto* o = static_cast<OpHandle*>(h);
    std::vector<double> tr;
    Result r;
    auto* p = reinterpret_cast<cd*>(psi);
    auto* c = reinterpret_cast<const cd*>(chi);
    if (solver == 0) {
      r = syssolver_cg<double>(*o->op, p, c, rsd, maxit, &tr);
    } else if (solver == 1) {
      // LinOpSysSolverBiCGStab: InvBiCGStab(*A, chi, psi, RsdBiCGStab, MaxBiCGStab, PLUS), then the
      // true residual |chi - A psi|
      r = invbicgstab<double>(*o->op, c, p, rsd, maxit, +1, &tr);
      const Range s = o->op->subset();
      std::vector<cd> t(o->g->V * 12);
      o->op->apply(t.data(), p, +1);
      for (long k = s.lo * 12; k < s.hi * 12; ++k) t[k] = c[k] - t[k];
      if (r.n_count >= 0) r.resid = std::sqrt(norm2<double>(t.data(), s));
    } else if (solver == 2) {
      int nupd = 0;
      r = syssolver_relcg(*o->op, *o->opf, p, c, rsd, delta, maxit, &tr, &nupd);
    } else if (solver == 3) {
      // LinOpSysSolverReliableBiCGStabClover (syssolver_linop_rel_bicgstab_clover.h:104-140)
      r = relbicgstab(*o->op, *o->opf, c, p, rsd, delta, maxit, +1, &tr, nullptr);
      if (r.n_count >= 0) r.resid = true_resid(*o->op, p, c, false);
    } else {
      throw std::runtime_error("unknown solver");
    }
    *n_count = r.n_count;
    *resid = r.resid;
    if (trace_len) *trace_len = (int)tr.size();
    if (trace)
      for (int k = 0; k < (int)tr.size() && k < trace_cap; ++k) trace[k] = tr[k];
  });
}

// MdagM system solvers: psi = (A^dag A)^-1 chi.  resid = |chi - A^dag A psi|.
//   0 MdagMSysSolverCG (syssolver_mdagm_cg.h:56-90): InvCG2 on chi directly
//   1 MdagMSysSolverBiCGStab (syssolver_mdagm_bicgstab.h:64-100): Y = A psi; BiCGStab(MINUS) on chi
//     for Y; BiCGStab(PLUS) on Y for psi
//   2 MdagMSysSolverReliableCGClover (syssolver_mdagm_rel_cg_clover.h:104-150): RelInvCG on chi
//   3 MdagMSysSolverReliableBiCGStabClover (syssolver_mdagm_rel_bicgstab_clover.h:104-176): the
//     two-step form of 1 with the reliable BiCGStab
int orc_solve_mdagm(void* h, int solver, double* psi, const double* chi, double rsd, int maxit, double delta,
                    int* n_count, double* resid) {
  return guard([&] {
    auto* o = static_cast<OpHandle*>(h);
    auto* p = reinterpret_cast<cd*>(psi);
    auto* c = reinterpret_cast<const cd*>(chi);
    Result r{0, 0};
    if (solver == 0) {
      r = invcg2<double>(*o->op, c, p, rsd, maxit, nullptr);
    } else if (solver == 2) {
      r = relcg(*o->op, *o->opf, c, p, rsd, delta, maxit, nullptr, nullptr);
    } else if (solver == 1 || solver == 3) {
      std::vector<cd> Y(o->g->V * 12);
      o->op->apply(Y.data(), p, +1);
      Result r1 = solver == 1 ? invbicgstab<double>(*o->op, c, Y.data(), rsd, maxit, -1, nullptr)
                              : relbicgstab(*o->op, *o->opf, c, Y.data(), rsd, delta, maxit, -1, nullptr, nullptr);
      Result r2 = solver == 1 ? invbicgstab<double>(*o->op, Y.data(), p, rsd, maxit, +1, nullptr)
                              : relbicgstab(*o->op, *o->opf, Y.data(), p, rsd, delta, maxit, +1, nullptr, nullptr);
      r.n_count = (r1.n_count < 0 || r2.n_count < 0) ? -1 : r1.n_count + r2.n_count;
    } else {
      throw std::runtime_error("unknown solver");
    }
    *n_count = r.n_count;
    *resid = true_resid(*o->op, p, c, true);
  });
}

// WilsonDslashBase::deriv (lwldslash_base_w.h:132-227): cb = 0 or 1, or -1 for both checkerboards
// (the sum, :110-124).  chi, psi full-lattice fermions; ds_u [4][V][3][3].
// CloverTermBase::deriv on cb (cb = -1: both), derivTrLn of the clover term inverted on cb (the even-even
// block's log det derivative of EvenOddPrecCloverLinOp::derivLogDetEvenEvenLinOp for cb = 0), and the EO
// operator's derivative.  U carries the boundary conditions; ds is [4][V][3][3].
int orc_clover_deriv(const int dims[4], const double* U, double csw_r, double csw_t, const double* chi,
                     const double* psi, int cb, double* ds_u) {
  return guard([&] {
    Geom g(dims);
    auto* ds = reinterpret_cast<cd*>(ds_u);
    auto* u = reinterpret_cast<const cd*>(U);
    auto* c = reinterpret_cast<const cd*>(chi);
    auto* p = reinterpret_cast<const cd*>(psi);
    if (cb >= 0) {
      clover_deriv<double>(g, u, csw_r, csw_t, c, p, cb, ds);
    } else {
      std::vector<cd> t((size_t)4 * g.V * 9);
      clover_deriv<double>(g, u, csw_r, csw_t, c, p, 0, ds);
      clover_deriv<double>(g, u, csw_r, csw_t, c, p, 1, t.data());
      for (size_t k = 0; k < t.size(); ++k) ds[k] += t[k];
    }
  });
}

int orc_clover_deriv_trln(const int dims[4], const double* U, double csw_r, double csw_t, double mass, int cb,
                          double* ds_u) {
  return guard([&] {
    Geom g(dims);
    auto* u = reinterpret_cast<const cd*>(U);
    std::vector<Clov<double>> tri(g.V);
    make_clover<double>(g, u, csw_r, csw_t, mass, tri.data());
    long nneg = 0;
    clover_invert<double>(g, tri.data(), cb, &nneg);
    clover_deriv_trln<double>(g, u, csw_r, csw_t, tri.data(), cb, reinterpret_cast<cd*>(ds_u));
  });
}

int orc_eo_clover_deriv(const int dims[4], const double* U, double mass, double csw_r, double csw_t, const double* chi,
                        const double* psi, int isign, double* ds_u) {
  return guard([&] {
    Geom g(dims);
    EOClover<double> M(g, reinterpret_cast<const cd*>(U), mass, csw_r, csw_t);
    eo_clover_deriv<double>(M, csw_r, csw_t, reinterpret_cast<const cd*>(chi), reinterpret_cast<const cd*>(psi), isign,
                            reinterpret_cast<cd*>(ds_u));
  });
}

int orc_dslash_deriv(const int dims[4], double* ds_u, const double* chi, const double* psi, int isign, int cb) {
  return guard([&] {
    Geom g(dims);
    auto* ds = reinterpret_cast<cd*>(ds_u);
    auto* c = reinterpret_cast<const cd*>(chi);
    auto* p = reinterpret_cast<const cd*>(psi);
    if (cb >= 0) {
      dslash_deriv<double>(g, ds, c, p, isign, cb);
    } else {
      std::vector<cd> t((size_t)4 * g.V * 9);
      dslash_deriv<double>(g, ds, c, p, isign, 0);
      dslash_deriv<double>(g, t.data(), c, p, isign, 1);
      for (size_t k = 0; k < t.size(); ++k) ds[k] += t[k];
    }
  });
}

// MdagMMultiSysSolverCG (multi_syssolver_mdagm_cg.h:58-100) -> MInvCG2: psi holds nshift subset-sized
// fields back to back (full-lattice arrays of the operator); resid[s] = |chi - (A^dag A + shift_s) psi_s|.
int orc_minvcg2(void* h, int nshift, const double* shifts, const double* rsd, int maxit, double* psi,
                const double* chi, int* n_count, double* resid) {
  return guard([&] {
    auto* o = static_cast<OpHandle*>(h);
    const long n = o->g->V * 12;
    std::vector<std::vector<cd>> ps;
    const std::vector<double> sh(shifts, shifts + nshift), rs(rsd, rsd + nshift);
    *n_count = minvcg2<double>(*o->op, reinterpret_cast<const cd*>(chi), ps, sh, rs, maxit);
    const Range sub = o->op->subset();
    std::vector<cd> mp(n), mmp(n);
    for (int s = 0; s < nshift; ++s) {
      std::memcpy(psi + (size_t)s * 2 * n, ps[s].data(), sizeof(cd) * n);
      o->op->apply(mp.data(), ps[s].data(), +1);
      o->op->apply(mmp.data(), mp.data(), -1);
      for (long q = sub.lo * 12; q < sub.hi * 12; ++q)
        mmp[q] = reinterpret_cast<const cd*>(chi)[q] - (mmp[q] + shifts[s] * ps[s][q]);
      resid[s] = std::sqrt(norm2<double>(mmp.data(), sub));
    }
  });
}

// PrecFermActQprop::operator(), eoprec_fermact_qprop.cc:41-80, for an even-odd operator with the CG
// inverter (solver 0), reliable CG (solver 2) or reliable BiCGStab (solver 3).  psi, chi are full-lattice.  resid = |M_unprec psi - chi|.
int orc_qprop(void* h, int solver, double* psi, const double* chi, double rsd, int maxit, double delta, int* n_count,
              double* resid) {
  return guard([&] {
    auto* o = static_cast<OpHandle*>(h);
    auto* eo = dynamic_cast<EOLinOp<double>*>(o->op.get());
    if (!eo) throw std::runtime_error("operator is not even-odd preconditioned");
    const Geom& g = *o->g;
    const long n = g.V * 12;
    auto* p = reinterpret_cast<cd*>(psi);
    auto* c = reinterpret_cast<const cd*>(chi);
    std::vector<cd> chi_tmp(n), t1(n), t2(n);
    eo->evenEvenInvLinOp(t1.data(), c, +1);
    eo->oddEvenLinOp(t2.data(), t1.data(), +1);
    for (long k = g.Vh * 12; k < n; ++k) chi_tmp[k] = c[k] - t2[k];
    Result r;
    if (solver == 0)
      r = syssolver_cg<double>(*eo, p, chi_tmp.data(), rsd, maxit, nullptr);
    else if (solver == 2)
      r = syssolver_relcg(*eo, *o->opf, p, chi_tmp.data(), rsd, delta, maxit, nullptr, nullptr);
    else if (solver == 3)
      r = relbicgstab(*eo, *o->opf, chi_tmp.data(), p, rsd, delta, maxit, +1, nullptr, nullptr);
    else
      throw std::runtime_error("qprop: unsupported solver");
    eo->evenOddLinOp(t1.data(), p, +1);
    for (long k = 0; k < g.Vh * 12; ++k) t2[k] = c[k] - t1[k];
    eo->evenEvenInvLinOp(p, t2.data(), +1);
    eo->unprecLinOp(t1.data(), p, +1);
    for (long k = 0; k < n; ++k) t1[k] -= c[k];
    *n_count = r.n_count;
    *resid = std::sqrt(norm2<double>(t1.data(), Range{0, g.V}));
  });
}

// quarkProp4_a (quarkprop4_w.cc:70-117) with a point source at site src_site, QUARK_SPIN_TYPE_FULL,
// followed by the prop_corr sanity observable of inline_propagator_w.cc:383-395
// (prop_corr[t] = sum over the time slice t of localNorm2 of the propagator).
// prop: optional V x 12(sink spin-colour) x 12(source spin-colour) complex output, may be NULL.
int orc_quarkprop4_point(void* h, int solver, long src_site, double rsd, int maxit, double delta, int* n_counts,
                         double* resids, double* prop_corr, double* prop) {
  return guard([&] {
    auto* o = static_cast<OpHandle*>(h);
    const Geom& g = *o->g;
    const long n = g.V * 12;
    std::vector<double> corr(g.L[3], 0.0);
    int idx = 0;
    for (int color_source = 0; color_source < 3; ++color_source)
      for (int spin_source = 0; spin_source < 4; ++spin_source, ++idx) {
        std::vector<cd> psi(n, cd(0)), chi(n, cd(0));
        chi[src_site * 12 + spin_source * 3 + color_source] = 1.0;
        double nrm = std::sqrt(norm2<double>(chi.data(), Range{0, g.V}));
        double fact = 1.0 / nrm;
        for (auto& z : chi) z *= fact;
        int nc;
        double rs;
        int rc = orc_qprop(h, solver, reinterpret_cast<double*>(psi.data()), reinterpret_cast<const double*>(chi.data()),
                           rsd, maxit, delta, &nc, &rs);
        if (rc) throw std::runtime_error(g_err);
        n_counts[idx] = nc;
        resids[idx] = rs;
        for (auto& z : psi) z /= fact;
        for (long s = 0; s < g.V; ++s) {
          int x[4];
          g.coords(s, x);
          double acc = 0;
          for (int k = 0; k < 12; ++k) acc += std::norm(psi[s * 12 + k]);
          corr[x[3]] += acc;
          if (prop)
            for (int k = 0; k < 12; ++k) {
              long off = ((s * 12 + k) * 12 + (spin_source * 3 + color_source)) * 2;
              prop[off] = psi[s * 12 + k].real();
              prop[off + 1] = psi[s * 12 + k].imag();
            }
        }
      }
    for (int t = 0; t < g.L[3]; ++t) prop_corr[t] = corr[t];
  });
}

// Timers for the CPU baseline: average seconds per call over n calls.
// what: 0 hopping term D (cb 1), 1 EO operator M, 2 one CG2 iteration equivalent (M, M^dag and 3 BLAS passes)
int orc_time(void* h, int what, int n, double* sec_per_call) {
  return guard([&] {
    auto* o = static_cast<OpHandle*>(h);
    const Geom& g = *o->g;
    std::vector<cd> a(g.V * 12), b(g.V * 12), c(g.V * 12);
    for (long k = 0; k < g.V * 12; ++k) a[k] = cd(std::sin(0.1 * k), std::cos(0.3 * k));
    double t0 = omp_get_wtime();
    for (int it = 0; it < n; ++it) {
      if (what == 0)
        dslash<double>(g, o->U.data(), b.data(), a.data(), +1, 1);
      else if (what == 1)
        o->op->apply(b.data(), a.data(), +1);
      else {
        const Range s = o->op->subset();
        o->op->apply(b.data(), a.data(), +1);
        double d = norm2<double>(b.data(), s);
        o->op->apply(c.data(), b.data(), -1);
        axpby<double>(a.data(), cd(-1e-3 / (1 + d)), c.data(), cd(1), s);
        (void)norm2<double>(a.data(), s);
        axpby<double>(b.data(), cd(1e-3), a.data(), cd(1), s);
      }
    }
    *sec_per_call = (omp_get_wtime() - t0) / n;
  });
}

}  // extern "C"
