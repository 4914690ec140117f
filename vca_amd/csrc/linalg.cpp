// linalg.cpp -- small host eigensolvers of the solver core.
//
// The reference calls LAPACK through SciFortran for these (eigh on the
// Lanczos tridiagonal, VCA_GF_NORMAL.f90:1012; eigh on dense small sectors,
// VCA_DIAG.f90:194).  This build carries its own, from the textbook
// methods: implicit-shift QL for the symmetric tridiagonal problem
// (Bowdler, Martin, Reinsch & Wilkinson, Numer. Math. 11 (1968) 293), and
// Householder reduction of a Hermitian matrix to real tridiagonal form
// followed by the same QL.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "vca_internal.hpp"

namespace vca {

// sqrt(a^2 + b^2) without hypot's overflow guards: the tridiagonal
// matrices here hold Hamiltonian-scale numbers (far from 1e+-150).
static inline double pyth(double a, double b) { return std::sqrt(a * a + b * b); }

// QL with implicit Wilkinson-type shifts on d[0..n), off-diagonal f[0..n)
// with f[i] coupling i and i+1 (f[n-1] = 0).  If z != null it is updated
// with the rotations (column-major, ldz rows, nz rows to update).
static void ql_implicit(int n, double *d, double *f, double *z, int nz, int ldz) {
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(f[m]) <= eps * dd) break;
      }
      if (m != l) {
        if (++iter > 60) break;  // no convergence: leave as is
        double g = (d[l + 1] - d[l]) / (2.0 * f[l]);
        double r = pyth(g, 1.0);
        g = d[m] - d[l] + f[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        bool deflated = false;
        for (i = m - 1; i >= l; --i) {
          double ff = s * f[i];
          double b = c * f[i];
          r = pyth(ff, g);
          f[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            f[m] = 0.0;
            deflated = true;
            break;
          }
          s = ff / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          if (z) {
            double *zi = z + (size_t)i * ldz, *zi1 = z + (size_t)(i + 1) * ldz;
            for (int k = 0; k < nz; ++k) {
              double t = zi1[k];
              zi1[k] = s * zi[k] + c * t;
              zi[k] = c * zi[k] - s * t;
            }
          }
        }
        if (deflated) continue;
        d[l] -= p;
        f[l] = g;
        f[m] = 0.0;
      }
    } while (m != l);
  }
}

void tridiag_eig(int n, const double *d, const double *e, double *w, double *z) {
  std::vector<double> dd(d, d + n), ff(n, 0.0);
  for (int i = 0; i + 1 < n; ++i) ff[i] = e[i + 1];
  std::vector<double> zz;
  if (z) {
    zz.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) zz[(size_t)i * n + i] = 1.0;
  }
  ql_implicit(n, dd.data(), ff.data(), z ? zz.data() : nullptr, n, n);
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return dd[a] < dd[b]; });
  for (int i = 0; i < n; ++i) {
    w[i] = dd[idx[i]];
    if (z)
      for (int k = 0; k < n; ++k) z[(size_t)i * n + k] = zz[(size_t)idx[i] * n + k];
  }
}

void tridiag_eig_first(int n, const double *d, const double *e, double *w, double *z0) {
  std::vector<double> dd(d, d + n), ff(n, 0.0);
  for (int i = 0; i + 1 < n; ++i) ff[i] = e[i + 1];
  // row 0 of the identity, carried through the same rotations: O(n) per
  // sweep instead of O(n^2) for the whole eigenvector matrix
  std::vector<double> zr(n, 0.0);
  zr[0] = 1.0;
  ql_implicit(n, dd.data(), ff.data(), zr.data(), 1, 1);
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return dd[a] < dd[b]; });
  for (int i = 0; i < n; ++i) {
    w[i] = dd[idx[i]];
    z0[i] = zr[idx[i]];
  }
}

double tridiag_lowest(int m, const double *d, const double *e) {
  std::vector<double> w(m);
  tridiag_eig(m, d, e, w.data(), nullptr);
  return w[0];
}

void herm_eig(int n, zc *a, double *w) {
  // a: column-major, a[i + j*n] = A(i, j).  Q accumulates the reflectors.
  auto A = [&](int i, int j) -> zc & { return a[i + (size_t)j * n]; };
  std::vector<zc> q((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) q[i + (size_t)i * n] = 1.0;
  std::vector<zc> v(n), wv(n), u(n);
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;  // length of the column below the diagonal
    double xnorm = 0;
    for (int i = 0; i < m; ++i) xnorm += std::norm(A(k + 1 + i, k));
    xnorm = std::sqrt(xnorm);
    if (xnorm == 0.0) continue;
    zc x0 = A(k + 1, k);
    zc phase = std::abs(x0) > 0 ? x0 / std::abs(x0) : zc(1.0);
    zc alpha = -phase * xnorm;
    for (int i = 0; i < m; ++i) v[i] = A(k + 1 + i, k);
    v[0] -= alpha;
    double vn = 0;
    for (int i = 0; i < m; ++i) vn += std::norm(v[i]);
    vn = std::sqrt(vn);
    if (vn == 0.0) continue;
    for (int i = 0; i < m; ++i) v[i] /= vn;
    // w = A_sub v, c = v^H w (real), u = 2 (w - c v)
    zc c = 0;
    for (int i = 0; i < m; ++i) {
      zc s = 0;
      for (int j = 0; j < m; ++j) s += A(k + 1 + i, k + 1 + j) * v[j];
      wv[i] = s;
    }
    for (int i = 0; i < m; ++i) c += std::conj(v[i]) * wv[i];
    for (int i = 0; i < m; ++i) u[i] = 2.0 * (wv[i] - c.real() * v[i]);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i)
        A(k + 1 + i, k + 1 + j) -= v[i] * std::conj(u[j]) + u[i] * std::conj(v[j]);
    A(k + 1, k) = alpha;
    A(k, k + 1) = std::conj(alpha);
    for (int i = 1; i < m; ++i) {
      A(k + 1 + i, k) = 0;
      A(k, k + 1 + i) = 0;
    }
    // Q <- Q P on columns k+1..n-1
    for (int r = 0; r < n; ++r) {
      zc s = 0;
      for (int i = 0; i < m; ++i) s += q[r + (size_t)(k + 1 + i) * n] * v[i];
      for (int i = 0; i < m; ++i) q[r + (size_t)(k + 1 + i) * n] -= 2.0 * s * std::conj(v[i]);
    }
  }
  // Real tridiagonal by a diagonal phase transform D.
  std::vector<double> d(n), f(n, 0.0);
  std::vector<zc> ph(n);
  ph[0] = 1.0;
  for (int i = 0; i < n; ++i) d[i] = A(i, i).real();
  for (int i = 0; i + 1 < n; ++i) {
    zc t = A(i + 1, i);
    double at = std::abs(t);
    f[i] = at;
    ph[i + 1] = at > 0 ? ph[i] * t / at : ph[i];
  }
  std::vector<double> z((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
  ql_implicit(n, d.data(), f.data(), z.data(), n, n);
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  // eigenvectors: Q D Z
  std::vector<zc> out((size_t)n * n);
  for (int jj = 0; jj < n; ++jj) {
    const double *zc_ = z.data() + (size_t)idx[jj] * n;
    for (int r = 0; r < n; ++r) {
      zc s = 0;
      for (int i = 0; i < n; ++i) s += q[r + (size_t)i * n] * ph[i] * zc_[i];
      out[r + (size_t)jj * n] = s;
    }
    w[jj] = d[idx[jj]];
  }
  std::copy(out.begin(), out.end(), a);
}

}  // namespace vca
