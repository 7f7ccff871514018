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

// Selected parts of the eigen-decomposition of the Hermitian matrix given by
// one triangle of A0 (the lower one, or the upper one with lower=false).
// Eigenvalues ascending in w.  With ncol > 0, Zc (n x ncol) receives the
// eigenvectors of the ncol lowest eigenvalues.  With row0 in [0, n), Zr
// ((n - row0) x n) receives rows row0 .. n-1 of every eigenvector (column c
// belongs to w[c]); this is what residual norms of a Krylov basis need, and
// without Zc it is obtained at O(n^2 (n - row0)) cost instead of O(n^3).
//
// Method: Householder reduction to a complex tridiagonal matrix (reflectors
// kept, not multiplied out), a diagonal unitary scaling to a real symmetric
// tridiagonal one, implicit-shift QL on that with the plane rotations applied
// only to what is requested (all of z, stored transposed so a rotation touches
// two contiguous rows, or just the requested rows of Q D), then the reflectors
// applied to the requested eigenvector columns.  Real and imaginary parts are
// kept in separate arrays.
void herm_eig_sel(const Mat& A0, bool lower, std::vector<double>& w, int ncol, Mat* Zc, int row0, Mat* Zr);

// Eigen-decomposition of the Hermitian matrix given by the lower triangle of
// A (the upper triangle with lower=false): eigenvalues ascending in w,
// eigenvectors as the columns of Z (row-major n x n).
inline void herm_eig(const Mat& A0, std::vector<double>& w, Mat& Z, bool lower = true) {
  if (A0.n == 0) {
    w.clear();
    Z = Mat();
    return;
  }
  herm_eig_sel(A0, lower, w, A0.n, &Z, -1, nullptr);
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
