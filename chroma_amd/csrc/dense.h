// Small dense complex linear algebra on the host for the Krylov-basis solvers (matrices of order
// NKrylov + NDefl, a few tens): LU solve with partial pivoting, thin Householder QR, and the
// eigen-decomposition of a general complex matrix (Hessenberg reduction, single-shift QR with
// Wilkinson shifts, eigenvectors by inverse iteration).  Plays the role of the LAPACK calls of
// syssolver_linop_fgmres_dr.cc (zgetrf/zgetrs :364-377, zgeev :386, zgeqrf/zunmqr/zungqr :540-595).
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace cmx {
namespace dense {

using cd = std::complex<double>;

// row-major rows x cols
struct Mat {
  int rows = 0, cols = 0;
  std::vector<cd> a;
  Mat() = default;
  Mat(int r, int c) : rows(r), cols(c), a((size_t)r * c, cd(0)) {}
  cd& operator()(int r, int c) { return a[(size_t)r * cols + c]; }
  const cd& operator()(int r, int c) const { return a[(size_t)r * cols + c]; }
};

inline Mat adjoint(const Mat& A) {
  Mat B(A.cols, A.rows);
  for (int r = 0; r < A.rows; ++r)
    for (int c = 0; c < A.cols; ++c) B(c, r) = std::conj(A(r, c));
  return B;
}

inline Mat matmul(const Mat& A, const Mat& B) {
  if (A.cols != B.rows) throw std::runtime_error("dense::matmul: shape mismatch");
  Mat C(A.rows, B.cols);
  for (int r = 0; r < A.rows; ++r)
    for (int k = 0; k < A.cols; ++k) {
      const cd x = A(r, k);
      if (x == cd(0)) continue;
      for (int c = 0; c < B.cols; ++c) C(r, c) += x * B(k, c);
    }
  return C;
}

// Solves A x = b (A square) by Gaussian elimination with partial pivoting.
inline std::vector<cd> lu_solve(Mat A, std::vector<cd> b) {
  const int n = A.rows;
  if (A.cols != n || (int)b.size() != n) throw std::runtime_error("dense::lu_solve: shape mismatch");
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int r = k + 1; r < n; ++r)
      if (std::abs(A(r, k)) > std::abs(A(p, k))) p = r;
    if (A(p, k) == cd(0)) throw std::runtime_error("dense::lu_solve: singular matrix");
    if (p != k) {
      for (int c = 0; c < n; ++c) std::swap(A(k, c), A(p, c));
      std::swap(b[k], b[p]);
    }
    for (int r = k + 1; r < n; ++r) {
      const cd f = A(r, k) / A(k, k);
      if (f == cd(0)) continue;
      for (int c = k; c < n; ++c) A(r, c) -= f * A(k, c);
      b[r] -= f * b[k];
    }
  }
  std::vector<cd> x(n);
  for (int r = n - 1; r >= 0; --r) {
    cd s = b[r];
    for (int c = r + 1; c < n; ++c) s -= A(r, c) * x[c];
    x[r] = s / A(r, r);
  }
  return x;
}

// Thin QR of A (m x n, m >= n) by Householder reflections: A = Q R, Q m x n with orthonormal
// columns, R n x n upper triangular.  If `full`, Q is m x m.
inline void qr(const Mat& A, Mat& Q, Mat& R, bool full = false) {
  // Householder QR with LAPACK's reflector convention (zgeqrf / zlarfg / zungqr, which the reference
  // calls): H_k = I - tau_k v_k v_k^H with v_k[k] = 1, R's diagonal real, beta = -sign(Re alpha) |x|.
  // The phases of Q's columns, the completion columns of the full form included, are then the
  // reference's, and so is the real part of the incremental residual g that FGMRES-DR tests (:172).
  const int m = A.rows, n = A.cols;
  if (m < n) throw std::runtime_error("dense::qr: rows < cols");
  Mat W = A;
  std::vector<std::vector<cd>> vs;
  std::vector<cd> taus;
  for (int k = 0; k < n; ++k) {
    std::vector<cd> v(m, cd(0));
    v[k] = 1.0;
    const cd alpha = W(k, k);
    double xn2 = 0;
    for (int r = k + 1; r < m; ++r) xn2 += std::norm(W(r, k));
    cd tau = 0;
    if (!(xn2 == 0 && alpha.imag() == 0)) {  // zlarfg
      const double beta = -std::copysign(std::sqrt(std::norm(alpha) + xn2), alpha.real());
      tau = cd((beta - alpha.real()) / beta, -alpha.imag() / beta);
      const cd scal = 1.0 / (alpha - beta);
      for (int r = k + 1; r < m; ++r) v[r] = W(r, k) * scal;
      W(k, k) = beta;
      for (int r = k + 1; r < m; ++r) W(r, k) = 0;
    }
    // W(k:, k+1:) <- H_k^H W(k:, k+1:)   (zgeqr2 applies conj(tau))
    for (int c = k + 1; c < n; ++c) {
      cd s = 0;
      for (int r = k; r < m; ++r) s += std::conj(v[r]) * W(r, c);
      s *= std::conj(tau);
      for (int r = k; r < m; ++r) W(r, c) -= v[r] * s;
    }
    vs.push_back(v);
    taus.push_back(tau);
  }
  R = Mat(n, n);
  for (int r = 0; r < n; ++r)
    for (int c = r; c < n; ++c) R(r, c) = W(r, c);
  // Q = H_0 H_1 ... H_{n-1} applied to the first qc columns of the identity (zungqr)
  const int qc = full ? m : n;
  Q = Mat(m, qc);
  for (int c = 0; c < qc; ++c) Q(c, c) = 1.0;
  for (int k = n - 1; k >= 0; --k) {
    const std::vector<cd>& v = vs[k];
    for (int c = 0; c < qc; ++c) {
      cd s = 0;
      for (int r = k; r < m; ++r) s += std::conj(v[r]) * Q(r, c);
      s *= taus[k];
      for (int r = k; r < m; ++r) Q(r, c) -= v[r] * s;
    }
  }
}

// Eigenvalues and right eigenvectors (columns of `vecs`, unit 2-norm) of a general complex matrix.
inline void eig(const Mat& A0, std::vector<cd>& vals, Mat& vecs) {
  const int n = A0.rows;
  if (A0.cols != n) throw std::runtime_error("dense::eig: not square");
  Mat H = A0;
  // Householder reduction to upper Hessenberg form
  for (int k = 0; k + 2 < n + 0 && k < n - 2; ++k) {
    double nrm = 0;
    for (int r = k + 1; r < n; ++r) nrm += std::norm(H(r, k));
    nrm = std::sqrt(nrm);
    if (nrm == 0) continue;
    std::vector<cd> v(n, cd(0));
    const cd x0 = H(k + 1, k);
    const cd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cd(1);
    v[k + 1] = x0 + ph * nrm;
    for (int r = k + 2; r < n; ++r) v[r] = H(r, k);
    double vn = 0;
    for (int r = k + 1; r < n; ++r) vn += std::norm(v[r]);
    vn = std::sqrt(vn);
    for (int r = k + 1; r < n; ++r) v[r] /= vn;
    for (int c = 0; c < n; ++c) {  // H <- P H
      cd s = 0;
      for (int r = k + 1; r < n; ++r) s += std::conj(v[r]) * H(r, c);
      for (int r = k + 1; r < n; ++r) H(r, c) -= 2.0 * v[r] * s;
    }
    for (int r = 0; r < n; ++r) {  // H <- H P
      cd s = 0;
      for (int c = k + 1; c < n; ++c) s += H(r, c) * v[c];
      for (int c = k + 1; c < n; ++c) H(r, c) -= 2.0 * s * std::conj(v[c]);
    }
    for (int r = k + 2; r < n; ++r) H(r, k) = 0;
  }
  // shifted QR iteration on the Hessenberg matrix (eigenvalues only)
  vals.assign(n, cd(0));
  int hi = n - 1, iter = 0;
  const double eps = 1e-15;
  while (hi >= 0) {
    if (hi == 0) {
      vals[0] = H(0, 0);
      break;
    }
    // find the start of the active unreduced block
    int lo = hi;
    while (lo > 0) {
      const double s = std::abs(H(lo, lo)) + std::abs(H(lo - 1, lo - 1));
      if (std::abs(H(lo, lo - 1)) <= eps * (s > 0 ? s : 1.0)) {
        H(lo, lo - 1) = 0;
        break;
      }
      --lo;
    }
    if (lo == hi) {
      vals[hi] = H(hi, hi);
      --hi;
      iter = 0;
      continue;
    }
    if (++iter > 200 * n) {
      int bad = 0;
      double mx = 0;
      for (const cd& z : A0.a) {
        if (!std::isfinite(z.real()) || !std::isfinite(z.imag())) ++bad;
        else mx = std::max(mx, std::abs(z));
      }
      throw std::runtime_error("dense::eig: QR iteration did not converge (order " + std::to_string(n) +
                               ", active block " + std::to_string(lo) + ".." + std::to_string(hi) +
                               ", non-finite entries " + std::to_string(bad) + ", max |a| " + std::to_string(mx) + ")");
    }
    // Wilkinson shift from the trailing 2x2 block (exceptional shift every 11 iterations)
    const cd a = H(hi - 1, hi - 1), b = H(hi - 1, hi), c = H(hi, hi - 1), d = H(hi, hi);
    cd mu;
    if (iter % 11 == 0) {
      mu = d + std::abs(c);
    } else {
      const cd tr = a + d, det = a * d - b * c;
      const cd disc = std::sqrt(tr * tr - 4.0 * det);
      const cd l1 = 0.5 * (tr + disc), l2 = 0.5 * (tr - disc);
      mu = std::abs(l1 - d) < std::abs(l2 - d) ? l1 : l2;
    }
    // one QR step on the block [lo, hi] with Givens rotations
    for (int r = lo; r <= hi; ++r) H(r, r) -= mu;
    std::vector<cd> cs(hi - lo), sn(hi - lo);
    for (int k = lo; k < hi; ++k) {
      const cd f = H(k, k), g = H(k + 1, k);
      const double t = std::sqrt(std::norm(f) + std::norm(g));
      cd cc = 1, ss = 0;
      if (t > 0) {
        cc = f / t;
        ss = g / t;
      }
      cs[k - lo] = cc;
      sn[k - lo] = ss;
      for (int col = k; col < n; ++col) {  // rows k, k+1 <- G^H [rows]
        const cd x = H(k, col), y = H(k + 1, col);
        H(k, col) = std::conj(cc) * x + std::conj(ss) * y;
        H(k + 1, col) = -ss * x + cc * y;
      }
    }
    for (int k = lo; k < hi; ++k) {  // columns k, k+1 <- [cols] G
      const cd cc = cs[k - lo], ss = sn[k - lo];
      for (int row = 0; row <= std::min(k + 2, hi); ++row) {
        const cd x = H(row, k), y = H(row, k + 1);
        H(row, k) = x * cc + y * ss;
        H(row, k + 1) = -x * std::conj(ss) + y * std::conj(cc);
      }
    }
    for (int r = lo; r <= hi; ++r) H(r, r) += mu;
  }
  // eigenvectors of A0 by inverse iteration with a slightly perturbed shift
  double anorm = 0;
  for (const cd& z : A0.a) anorm = std::max(anorm, std::abs(z));
  if (anorm == 0) anorm = 1;
  vecs = Mat(n, n);
  for (int e = 0; e < n; ++e) {
    Mat B = A0;
    const cd shift = vals[e] + cd(1e-10 * anorm, 1e-10 * anorm);
    for (int r = 0; r < n; ++r) B(r, r) -= shift;
    std::vector<cd> x(n);
    for (int r = 0; r < n; ++r) x[r] = cd(1.0 + 0.1 * r, 0.05 * (r % 3));
    for (int it = 0; it < 3; ++it) {
      x = lu_solve(B, x);
      double nr = 0;
      for (const cd& z : x) nr += std::norm(z);
      nr = std::sqrt(nr);
      for (cd& z : x) z /= nr;
    }
    for (int r = 0; r < n; ++r) vecs(r, e) = x[r];
  }
}

}  // namespace dense
}  // namespace cmx
