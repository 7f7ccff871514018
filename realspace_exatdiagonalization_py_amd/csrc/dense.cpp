// Hermitian eigensolver of dense.h, compiled by the host C++ compiler (this
// file holds no device code).
#include "dense.h"

namespace edx {
namespace dense {

void herm_eig_sel(const Mat& A0, bool lower, std::vector<double>& w, int ncol, Mat* Zc, int row0, Mat* Zr) {
  const int n = A0.n;
  const size_t nn = (size_t)n * n;
  std::vector<double> ar(nn), ai(nn);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      cplx x;
      if (i == j) x = cplx(A0(i, i).real(), 0.0);
      else if ((i > j) == lower) x = A0(i, j);
      else x = std::conj(A0(j, i));
      ar[(size_t)i * n + j] = x.real();
      ai[(size_t)i * n + j] = x.imag();
    }
  // reflectors H_k = I - 2 v_k v_k^H, v_k zero in entries <= k; A0 = Q T Q^H
  // with Q = H_0 H_1 ... and T tridiagonal
  const int nref = std::max(n - 2, 0);
  std::vector<double> vr((size_t)nref * n, 0.0), vi((size_t)nref * n, 0.0);
  std::vector<char> has(nref, 0);
  std::vector<double> pr(n), pim(n), qr(n), qi(n);
  for (int k = 0; k + 2 < n; ++k) {
    double xn2 = 0.0;
    for (int i = k + 1; i < n; ++i) xn2 += ar[(size_t)i * n + k] * ar[(size_t)i * n + k] + ai[(size_t)i * n + k] * ai[(size_t)i * n + k];
    const double xn = std::sqrt(xn2);
    if (xn == 0.0) continue;
    const double x0r = ar[(size_t)(k + 1) * n + k], x0i = ai[(size_t)(k + 1) * n + k];
    const double ax0 = std::hypot(x0r, x0i);
    const double phr = ax0 > 0.0 ? x0r / ax0 : 1.0, phi_ = ax0 > 0.0 ? x0i / ax0 : 0.0;
    const double alr = -phr * xn, ali = -phi_ * xn;
    double* v_r = &vr[(size_t)k * n];
    double* v_i = &vi[(size_t)k * n];
    for (int i = k + 1; i < n; ++i) {
      v_r[i] = ar[(size_t)i * n + k];
      v_i[i] = ai[(size_t)i * n + k];
    }
    v_r[k + 1] -= alr;
    v_i[k + 1] -= ali;
    double vn2 = 0.0;
    for (int i = k + 1; i < n; ++i) vn2 += v_r[i] * v_r[i] + v_i[i] * v_i[i];
    const double vn = std::sqrt(vn2);
    if (vn == 0.0) {
      std::fill(v_r, v_r + n, 0.0);
      std::fill(v_i, v_i + n, 0.0);
      continue;
    }
    const double inv = 1.0 / vn;
    for (int i = k + 1; i < n; ++i) {
      v_r[i] *= inv;
      v_i[i] *= inv;
    }
    has[k] = 1;
    // trailing block B = A[k+1.., k+1..] (lower triangle kept): p = B v,
    // K = v^H p (real), q = p - K v, B <- B - 2 (v q^H + q v^H)
    for (int i = k + 1; i < n; ++i) pr[i] = pim[i] = 0.0;
    for (int i = k + 1; i < n; ++i) {
      const double* br = &ar[(size_t)i * n];
      const double* bi = &ai[(size_t)i * n];
      const double vir = v_r[i], vii = v_i[i];
      double sr = br[i] * vir, si = br[i] * vii;  // diagonal (real)
      for (int j = k + 1; j < i; ++j) {
        // p_i += B_ij v_j ; p_j += conj(B_ij) v_i
        sr += br[j] * v_r[j] - bi[j] * v_i[j];
        si += br[j] * v_i[j] + bi[j] * v_r[j];
        pr[j] += br[j] * vir + bi[j] * vii;
        pim[j] += br[j] * vii - bi[j] * vir;
      }
      pr[i] += sr;
      pim[i] += si;
    }
    double K = 0.0;
    for (int i = k + 1; i < n; ++i) K += v_r[i] * pr[i] + v_i[i] * pim[i];
    for (int i = k + 1; i < n; ++i) {
      qr[i] = pr[i] - K * v_r[i];
      qi[i] = pim[i] - K * v_i[i];
    }
    for (int i = k + 1; i < n; ++i) {
      double* br = &ar[(size_t)i * n];
      double* bi = &ai[(size_t)i * n];
      const double a1r = 2.0 * v_r[i], a1i = 2.0 * v_i[i], a2r = 2.0 * qr[i], a2i = 2.0 * qi[i];
      for (int j = k + 1; j <= i; ++j) {
        // v_i conj(q_j) + q_i conj(v_j), times 2
        br[j] -= a1r * qr[j] + a1i * qi[j] + a2r * v_r[j] + a2i * v_i[j];
        bi[j] -= a1i * qr[j] - a1r * qi[j] + a2i * v_r[j] - a2r * v_i[j];
      }
    }
    ar[(size_t)(k + 1) * n + k] = alr;
    ai[(size_t)(k + 1) * n + k] = ali;
    for (int i = k + 2; i < n; ++i) ar[(size_t)i * n + k] = ai[(size_t)i * n + k] = 0.0;
  }
  // T = D T_real D^H with D = diag(phi), |phi| = 1
  std::vector<double> d(n), e(n, 0.0), phr(n, 1.0), phim(n, 0.0);
  for (int i = 0; i < n; ++i) d[i] = ar[(size_t)i * n + i];
  for (int i = 0; i + 1 < n; ++i) {
    const double er = ar[(size_t)(i + 1) * n + i], ei = ai[(size_t)(i + 1) * n + i];
    const double ae = std::hypot(er, ei);
    e[i] = ae;
    if (ae > 0.0) {
      const double ur = er / ae, ui = ei / ae;
      phr[i + 1] = phr[i] * ur - phim[i] * ui;
      phim[i + 1] = phr[i] * ui + phim[i] * ur;
    } else {
      phr[i + 1] = phr[i];
      phim[i + 1] = phim[i];
    }
  }
  // row r of Q D: u = e_r^T H_0 H_1 ..., then scaled by phi
  auto q_row = [&](int r, double* ur, double* ui) {
    std::fill(ur, ur + n, 0.0);
    std::fill(ui, ui + n, 0.0);
    ur[r] = 1.0;
    for (int k = 0; k < nref; ++k) {
      if (!has[k]) continue;
      const double* v_r = &vr[(size_t)k * n];
      const double* v_i = &vi[(size_t)k * n];
      double sr = 0.0, si = 0.0;  // u v
      for (int j = k + 1; j < n; ++j) {
        sr += ur[j] * v_r[j] - ui[j] * v_i[j];
        si += ur[j] * v_i[j] + ui[j] * v_r[j];
      }
      for (int j = k + 1; j < n; ++j) {  // u -= 2 (u v) v^H
        ur[j] -= 2.0 * (sr * v_r[j] + si * v_i[j]);
        ui[j] -= 2.0 * (si * v_r[j] - sr * v_i[j]);
      }
    }
    for (int j = 0; j < n; ++j) {
      const double xr = ur[j] * phr[j] - ui[j] * phim[j];
      const double xi = ur[j] * phim[j] + ui[j] * phr[j];
      ur[j] = xr;
      ui[j] = xi;
    }
  };
  const bool full = ncol > 0 && Zc;
  const bool rows = row0 >= 0 && row0 < n && Zr;
  const int m = rows ? n - row0 : 0;
  // rotation targets, transposed: zt[c][i] = z[i][c]; wt[c][r] = (Q D z)[row0 + r][c]
  std::vector<double> zt, wtr, wti;
  if (full) {
    zt.assign(nn, 0.0);
    for (int i = 0; i < n; ++i) zt[(size_t)i * n + i] = 1.0;
  } else if (rows) {
    wtr.assign((size_t)n * m, 0.0);
    wti.assign((size_t)n * m, 0.0);
    std::vector<double> ur(n), ui(n);
    for (int r = 0; r < m; ++r) {
      q_row(row0 + r, ur.data(), ui.data());
      for (int c = 0; c < n; ++c) {
        wtr[(size_t)c * m + r] = ur[c];
        wti[(size_t)c * m + r] = ui[c];
      }
    }
  }
  // implicit-shift QL on (d, e), e[i] coupling i and i+1
  {
    const double eps = std::numeric_limits<double>::epsilon();
    if (n > 0) e[n - 1] = 0.0;
    auto rotate = [&](int i, double s, double c) {
      if (full) {
        double* z0 = &zt[(size_t)i * n];
        double* z1 = &zt[(size_t)(i + 1) * n];
        for (int k = 0; k < n; ++k) {
          const double f = z1[k];
          z1[k] = s * z0[k] + c * f;
          z0[k] = c * z0[k] - s * f;
        }
      } else if (rows) {
        double* w0r = &wtr[(size_t)i * m];
        double* w1r = &wtr[(size_t)(i + 1) * m];
        double* w0i = &wti[(size_t)i * m];
        double* w1i = &wti[(size_t)(i + 1) * m];
        for (int k = 0; k < m; ++k) {
          const double fr = w1r[k], fi = w1i[k];
          w1r[k] = s * w0r[k] + c * fr;
          w1i[k] = s * w0i[k] + c * fi;
          w0r[k] = c * w0r[k] - s * fr;
          w0i[k] = c * w0i[k] - s * fi;
        }
      }
    };
    for (int l = 0; l < n; ++l) {
      for (int iter = 0;; ++iter) {
        int mm = l;
        for (; mm < n - 1; ++mm) {
          const double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
          if (std::fabs(e[mm]) <= eps * dd) break;
        }
        if (mm == l) break;
        if (iter > 60) throw std::runtime_error("tridiagonal QL did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[mm] - d[l] + e[l] / (g + std::copysign(r, g));
        double s = 1.0, c = 1.0, p = 0.0;
        bool split = false;
        for (int i = mm - 1; i >= l; --i) {
          const double f = s * e[i];
          const double b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[mm] = 0.0;
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
          rotate(i, s, c);
        }
        if (split) continue;
        d[l] -= p;
        e[l] = g;
        e[mm] = 0.0;
      }
    }
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
  w.resize(n);
  for (int c = 0; c < n; ++c) w[c] = d[order[c]];
  if (full) {
    const int nc = std::min(ncol, n);
    *Zc = Mat(n, nc);
    std::vector<double> yr(n), yi(n);
    for (int c = 0; c < nc; ++c) {
      const double* z = &zt[(size_t)order[c] * n];
      for (int i = 0; i < n; ++i) {
        yr[i] = phr[i] * z[i];
        yi[i] = phim[i] * z[i];
      }
      for (int k = nref - 1; k >= 0; --k) {  // y <- H_k y
        if (!has[k]) continue;
        const double* v_r = &vr[(size_t)k * n];
        const double* v_i = &vi[(size_t)k * n];
        double sr = 0.0, si = 0.0;  // v^H y
        for (int j = k + 1; j < n; ++j) {
          sr += v_r[j] * yr[j] + v_i[j] * yi[j];
          si += v_r[j] * yi[j] - v_i[j] * yr[j];
        }
        for (int j = k + 1; j < n; ++j) {
          yr[j] -= 2.0 * (v_r[j] * sr - v_i[j] * si);
          yi[j] -= 2.0 * (v_r[j] * si + v_i[j] * sr);
        }
      }
      for (int i = 0; i < n; ++i) (*Zc)(i, c) = cplx(yr[i], yi[i]);
    }
    if (rows) {
      *Zr = Mat(n - row0, n);
      std::vector<double> ur(n), ui(n);
      for (int r = 0; r < n - row0; ++r) {
        q_row(row0 + r, ur.data(), ui.data());
        for (int c = 0; c < n; ++c) {
          const double* z = &zt[(size_t)order[c] * n];
          double sr = 0.0, si = 0.0;
          for (int i = 0; i < n; ++i) {
            sr += ur[i] * z[i];
            si += ui[i] * z[i];
          }
          (*Zr)(r, c) = cplx(sr, si);
        }
      }
    }
  } else if (rows) {
    *Zr = Mat(m, n);
    for (int c = 0; c < n; ++c)
      for (int r = 0; r < m; ++r)
        (*Zr)(r, c) = cplx(wtr[(size_t)order[c] * m + r], wti[(size_t)order[c] * m + r]);
  }
}

}  // namespace dense
}  // namespace edx
