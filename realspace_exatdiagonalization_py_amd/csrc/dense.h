// Small dense complex linear algebra on the host (n <= a few hundred):
// Hermitian eigensolver (Householder reduction to a real tridiagonal matrix,
// then implicit-shift QL with eigenvector accumulation), Cholesky, and
// triangular inversion.  Used for the Rayleigh-Ritz and block-orthogonalisation
// steps of the block Lanczos solver, whose sizes are the Krylov basis size.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace edx {
namespace dense {

using cplx = std::complex<double>;

// Row-major n x n matrix helper.
struct Mat {
  int n = 0, m = 0;
  std::vector<cplx> a;
  Mat() = default;
  Mat(int rows, int cols) : n(rows), m(cols), a((size_t)rows * cols, cplx(0.0, 0.0)) {}
  cplx& operator()(int i, int j) { return a[(size_t)i * m + j]; }
  const cplx& operator()(int i, int j) const { return a[(size_t)i * m + j]; }
};

// Implicit QL on the symmetric tridiagonal (d, e), e[i] coupling i and i+1,
// rotating the columns of the row-major real matrix z (n x n).
inline void tridiag_ql(int n, std::vector<double>& d, std::vector<double>& e, std::vector<double>& z) {
  const double eps = std::numeric_limits<double>::epsilon();
  if (n > 0) e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    for (int iter = 0;; ++iter) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m == l) break;
      if (iter > 60) throw std::runtime_error("tridiagonal QL did not converge");
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      bool split = false;
      for (int i = m - 1; i >= l; --i) {
        double f = s * e[i];
        const double b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          split = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        for (int k = 0; k < n; ++k) {
          double* zk = &z[(size_t)k * n];
          f = zk[i + 1];
          zk[i + 1] = s * zk[i] + c * f;
          zk[i] = c * zk[i] - s * f;
        }
      }
      if (split) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
}

// Eigen-decomposition of the Hermitian matrix given by the lower triangle of
// A (the upper triangle is ignored; with lower=false the upper triangle is
// used instead).  Eigenvalues ascending in w, eigenvectors as the columns of
// Z (row-major n x n).
inline void herm_eig(const Mat& A0, std::vector<double>& w, Mat& Z, bool lower = true) {
  const int n = A0.n;
  Mat A(n, n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      if (i == j) A(i, j) = cplx(A0(i, i).real(), 0.0);
      else if ((i > j) == lower) A(i, j) = A0(i, j);
      else A(i, j) = std::conj(A0(j, i));
    }
  Mat Q(n, n);
  for (int i = 0; i < n; ++i) Q(i, i) = 1.0;
  std::vector<cplx> v(n), p(n), q(n), qv(n);
  for (int k = 0; k + 2 < n; ++k) {
    double xn2 = 0.0;
    for (int i = k + 1; i < n; ++i) xn2 += std::norm(A(i, k));
    const double xn = std::sqrt(xn2);
    if (xn == 0.0) continue;
    const cplx x0 = A(k + 1, k);
    const cplx ph = std::abs(x0) > 0.0 ? x0 / std::abs(x0) : cplx(1.0, 0.0);
    const cplx alpha = -ph * xn;
    std::fill(v.begin(), v.end(), cplx(0.0, 0.0));
    for (int i = k + 1; i < n; ++i) v[i] = A(i, k);
    v[k + 1] -= alpha;
    double vn2 = 0.0;
    for (int i = k + 1; i < n; ++i) vn2 += std::norm(v[i]);
    const double vn = std::sqrt(vn2);
    if (vn == 0.0) continue;
    for (int i = k + 1; i < n; ++i) v[i] /= vn;
    // trailing block B = A[k+1.., k+1..]: p = B v, K = v^H p, q = p - K v,
    // B <- B - 2 (v q^H + q v^H)
    for (int i = k + 1; i < n; ++i) {
      cplx s(0.0, 0.0);
      for (int j = k + 1; j < n; ++j) s += A(i, j) * v[j];
      p[i] = s;
    }
    cplx K(0.0, 0.0);
    for (int i = k + 1; i < n; ++i) K += std::conj(v[i]) * p[i];
    for (int i = k + 1; i < n; ++i) q[i] = p[i] - K.real() * v[i];
    for (int i = k + 1; i < n; ++i)
      for (int j = k + 1; j < n; ++j)
        A(i, j) -= 2.0 * (v[i] * std::conj(q[j]) + q[i] * std::conj(v[j]));
    A(k + 1, k) = alpha;
    A(k, k + 1) = std::conj(alpha);
    for (int i = k + 2; i < n; ++i) A(i, k) = A(k, i) = 0.0;
    // Q <- Q (I - 2 v v^H)
    for (int r = 0; r < n; ++r) {
      cplx s(0.0, 0.0);
      for (int j = k + 1; j < n; ++j) s += Q(r, j) * v[j];
      qv[r] = s;
    }
    for (int r = 0; r < n; ++r)
      for (int j = k + 1; j < n; ++j) Q(r, j) -= 2.0 * qv[r] * std::conj(v[j]);
  }
  std::vector<double> d(n), e(n, 0.0);
  std::vector<cplx> phi(n, cplx(1.0, 0.0));
  for (int i = 0; i < n; ++i) d[i] = A(i, i).real();
  for (int i = 0; i + 1 < n; ++i) {
    const cplx ei = A(i + 1, i);
    const double ae = std::abs(ei);
    e[i] = ae;
    phi[i + 1] = ae > 0.0 ? phi[i] * (ei / ae) : phi[i];
  }
  std::vector<double> z((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
  tridiag_ql(n, d, e, z);
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
  w.resize(n);
  Z = Mat(n, n);
  for (int c = 0; c < n; ++c) {
    const int src = order[c];
    w[c] = d[src];
    for (int r = 0; r < n; ++r) {
      cplx s(0.0, 0.0);
      for (int i = 0; i < n; ++i) s += Q(r, i) * phi[i] * z[(size_t)i * n + src];
      Z(r, c) = s;
    }
  }
}

// Upper Cholesky factor R with A = R^H R.  Returns false when a pivot is not
// above tiny * (largest diagonal entry).
inline bool cholesky_upper(const Mat& A, Mat& R, double tiny) {
  const int n = A.n;
  R = Mat(n, n);
  double dmax = 0.0;
  for (int i = 0; i < n; ++i) dmax = std::max(dmax, A(i, i).real());
  for (int j = 0; j < n; ++j) {
    double s = A(j, j).real();
    for (int k = 0; k < j; ++k) s -= std::norm(R(k, j));
    if (!(s > tiny * dmax) || dmax <= 0.0) return false;
    const double rjj = std::sqrt(s);
    R(j, j) = rjj;
    for (int i = j + 1; i < n; ++i) {
      cplx t = A(j, i);
      for (int k = 0; k < j; ++k) t -= std::conj(R(k, j)) * R(k, i);
      R(j, i) = t / rjj;
    }
  }
  return true;
}

// Inverse of an upper-triangular matrix.
inline Mat upper_inverse(const Mat& R) {
  const int n = R.n;
  Mat X(n, n);
  for (int j = 0; j < n; ++j) {
    X(j, j) = 1.0 / R(j, j);
    for (int i = j - 1; i >= 0; --i) {
      cplx s(0.0, 0.0);
      for (int k = i + 1; k <= j; ++k) s += R(i, k) * X(k, j);
      X(i, j) = -s / R(i, i);
    }
  }
  return X;
}

}  // namespace dense
}  // namespace edx
