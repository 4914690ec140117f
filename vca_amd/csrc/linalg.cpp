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

// The n-step Lanczos tridiagonal of the sum of two spectral measures, each
// given by an n-step chain: chain i has start norm ni and coefficients
// (ai[0..n), bi[1..n)), i.e. the n-node Gauss quadrature of its measure
// ni^2 <u|delta(x - H)|u>.  The Lanczos coefficients alpha_0..alpha_{n-1},
// beta_1..beta_{n-1} of a measure depend only on its moments 0..2n-1, and an
// n-node Gauss quadrature reproduces those exactly, so the union of the two
// quadratures (2n nodes) has the coefficients of the sum measure itself:
// with a real H and real a, b, the measure of a + i b (or a - i b) is the
// sum of those of a and b (the cross terms cancel, (z - H)^-1 being
// symmetric), so the chain of the start vector a - i b is the chain of this
// sum.
//
// The tridiagonal of the 2n-node measure is rebuilt from its nodes and
// weights by the Rutishauser-Kahan-Pal-Walker update (W. B. Gragg and W. J.
// Harrod, Numer. Math. 44 (1984) 317; W. Gautschi, Orthogonal Polynomials:
// Computation and Approximation, OUP 2004, sec. 2.2.3): the nodes enter one
// at a time, each folded into the Jacobi matrix of the previous ones by a
// sweep of plane rotations, O(m^2) for m nodes and stable (the same
// coefficients as Lanczos with full reorthogonalisation, to rounding).  An
// exhausted measure (a coefficient beta below 1e-13 of the node range) ends
// the chain as the device Lanczos does on an exact breakdown: the remaining
// alpha and beta are 0.  Returns the start norm sqrt(n1^2 + n2^2); al and be
// receive n values each (be[0] = 0).  lanczos_of_quadratures takes the two
// quadratures themselves (nodes x, first eigenvector components z: weights
// z^2, as tridiag_eig_first returns them); lanczos_of_sum the chains.
double lanczos_of_quadratures(int n, const double *x1, const double *z1, double n1,
                              const double *x2, const double *z2, double n2, double *al,
                              double *be) {
  for (int k = 0; k < n; ++k) {
    al[k] = 0.0;
    be[k] = 0.0;
  }
  if (!(n1 * n1 + n2 * n2 > 0.0)) return 0.0;  // both empty
  // the nodes of the non-empty quadratures, with their weights
  std::vector<double> x, q;
  x.reserve(2 * n);
  q.reserve(2 * n);
  for (int c = 0; c < 2; ++c) {
    const double nc = c ? n2 : n1;
    const double *xc = c ? x2 : x1, *zc = c ? z2 : z1;
    if (!(nc > 0.0)) continue;
    for (int i = 0; i < n; ++i) {
      x.push_back(xc[i]);
      q.push_back(zc[i] * zc[i] * nc * nc);
    }
  }
  const int m = (int)x.size();
  const double wt = n1 * n1 + n2 * n2;
  double xr = 0.0, qs = 0.0;
  for (int i = 0; i < m; ++i) {
    qs += q[i];
    xr = std::max(xr, std::fabs(x[i]));
  }
  // RKPW: d = diagonal, g = squared off-diagonals (g[0]: the total weight)
  std::vector<double> d(m, 0.0), g(m, 0.0);
  d[0] = x[0];
  g[0] = q[0] / qs;
  for (int j = 1; j < m; ++j) {
    double wn = q[j] / qs, c = 1.0, sgm = 0.0, t = 0.0;
    const double lam = x[j];
    for (int k = 0; k <= j; ++k) {
      const double rho = g[k] + wn;
      const double gk = c * rho;
      const double sgm_prev = sgm;
      if (rho <= 0.0) {
        c = 1.0;
        sgm = 0.0;
      } else {
        c = g[k] / rho;
        sgm = wn / rho;
      }
      const double tk = sgm * (d[k] - lam) - c * t;
      d[k] -= tk - t;
      t = tk;
      wn = sgm <= 0.0 ? sgm_prev * g[k] : t * t / sgm;
      g[k] = gk;
    }
  }
  for (int k = 0; k < n && k < m; ++k) {
    if (k > 0) {
      const double b = std::sqrt(std::max(g[k], 0.0));
      if (!(b > 1e-13 * xr)) break;  // exhausted: the rest stays 0
      be[k] = b;
    }
    al[k] = d[k];
  }
  return std::sqrt(wt);
}

double lanczos_of_sum(int n, const double *a1, const double *b1, double n1, const double *a2,
                      const double *b2, double n2, double *al, double *be) {
  std::vector<double> x1(n), z1(n), x2(n), z2(n);
  tridiag_eig_first(n, a1, b1, x1.data(), z1.data());
  tridiag_eig_first(n, a2, b2, x2.data(), z2.data());
  return lanczos_of_quadratures(n, x1.data(), z1.data(), n1, x2.data(), z2.data(), n2, al, be);
}

double tridiag_lowest(int m, const double *d, const double *e) {
  std::vector<double> w(m);
  tridiag_eig(m, d, e, w.data(), nullptr);
  return w[0];
}

// Householder reduction of the Hermitian (T = complex) or real symmetric
// (T = double) matrix a to real tridiagonal form, QL on it, and the
// eigenvectors Q D Z.  The loops run down columns (column-major storage),
// so the O(n^3) parts stream contiguous memory.
static inline double cnj(double x) { return x; }
static inline zc cnj(const zc &x) { return std::conj(x); }
static inline double sqnorm(double x) { return x * x; }
static inline double sqnorm(const zc &x) { return std::norm(x); }
static inline double re(double x) { return x; }
static inline double re(const zc &x) { return x.real(); }

template <typename T>
static void herm_eig_t(int n, T *a, double *w, zc *vec_out) {
  auto A = [&](int i, int j) -> T & { return a[i + (size_t)j * n]; };
  std::vector<T> q((size_t)n * n, T(0));
  for (int i = 0; i < n; ++i) q[i + (size_t)i * n] = T(1);
  std::vector<T> v(n), wv(n), u(n), sr(n);
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;  // length of the column below the diagonal
    T *colk = &A(k + 1, k);
    double xnorm = 0;
    for (int i = 0; i < m; ++i) xnorm += sqnorm(colk[i]);
    xnorm = std::sqrt(xnorm);
    if (xnorm == 0.0) continue;
    const T x0 = colk[0];
    const double ax0 = std::abs(x0);
    const T phase = ax0 > 0 ? x0 / ax0 : T(1);
    const T alpha = -phase * xnorm;
    for (int i = 0; i < m; ++i) v[i] = colk[i];
    v[0] -= alpha;
    double vn = 0;
    for (int i = 0; i < m; ++i) vn += sqnorm(v[i]);
    vn = std::sqrt(vn);
    if (vn == 0.0) continue;
    for (int i = 0; i < m; ++i) v[i] /= vn;
    // w = A_sub v (by columns), c = v^H w (real), u = 2 (w - c v)
    std::fill(wv.begin(), wv.begin() + m, T(0));
    for (int j = 0; j < m; ++j) {
      const T vj = v[j];
      const T *cj = &A(k + 1, k + 1 + j);
      for (int i = 0; i < m; ++i) wv[i] += cj[i] * vj;
    }
    T c = 0;
    for (int i = 0; i < m; ++i) c += cnj(v[i]) * wv[i];
    const double cr = re(c);
    for (int i = 0; i < m; ++i) u[i] = 2.0 * (wv[i] - cr * v[i]);
    for (int j = 0; j < m; ++j) {
      T *cj = &A(k + 1, k + 1 + j);
      const T uj = cnj(u[j]), vj = cnj(v[j]);
      for (int i = 0; i < m; ++i) cj[i] -= v[i] * uj + u[i] * vj;
    }
    colk[0] = alpha;
    A(k, k + 1) = cnj(alpha);
    for (int i = 1; i < m; ++i) {
      colk[i] = 0;
      A(k, k + 1 + i) = 0;
    }
    // Q <- Q P on columns k+1..n-1: s = Q_sub v, Q_sub -= 2 s v^H
    std::fill(sr.begin(), sr.end(), T(0));
    for (int i = 0; i < m; ++i) {
      const T vi = v[i];
      const T *qc = &q[(size_t)(k + 1 + i) * n];
      for (int r = 0; r < n; ++r) sr[r] += qc[r] * vi;
    }
    for (int i = 0; i < m; ++i) {
      const T ci = 2.0 * cnj(v[i]);
      T *qc = &q[(size_t)(k + 1 + i) * n];
      for (int r = 0; r < n; ++r) qc[r] -= sr[r] * ci;
    }
  }
  // Real tridiagonal by a diagonal phase transform D.
  std::vector<double> d(n), f(n, 0.0);
  std::vector<T> ph(n);
  ph[0] = T(1);
  for (int i = 0; i < n; ++i) d[i] = re(A(i, i));
  for (int i = 0; i + 1 < n; ++i) {
    const T t = A(i + 1, i);
    const double at = std::abs(t);
    f[i] = at;
    ph[i + 1] = at > 0 ? ph[i] * t / at : ph[i];
  }
  std::vector<double> z((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
  ql_implicit(n, d.data(), f.data(), z.data(), n, n);
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  // eigenvectors: Q D Z, column by column
  std::vector<T> col(n);
  for (int jj = 0; jj < n; ++jj) {
    const double *zj = z.data() + (size_t)idx[jj] * n;
    std::fill(col.begin(), col.end(), T(0));
    for (int i = 0; i < n; ++i) {
      const T ci = ph[i] * zj[i];
      const T *qc = &q[(size_t)i * n];
      for (int r = 0; r < n; ++r) col[r] += qc[r] * ci;
    }
    for (int r = 0; r < n; ++r) vec_out[r + (size_t)jj * n] = zc(col[r]);
    w[jj] = d[idx[jj]];
  }
}

void herm_eig(int n, zc *a, double *w) {
  // a real symmetric matrix (real Hamiltonians) runs in real arithmetic
  bool real = true;
  for (size_t i = 0; i < (size_t)n * n && real; ++i) real = a[i].imag() == 0.0;
  if (real) {
    std::vector<double> ar((size_t)n * n);
    for (size_t i = 0; i < ar.size(); ++i) ar[i] = a[i].real();
    herm_eig_t<double>(n, ar.data(), w, a);
  } else {
    std::vector<zc> ac(a, a + (size_t)n * n);
    herm_eig_t<zc>(n, ac.data(), w, a);
  }
}

}  // namespace vca
