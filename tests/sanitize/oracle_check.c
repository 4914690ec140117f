/* Driver of the oracle's C restatement under ASan/UBSan: Fock maps and the
 * binary search over every popcount of 12 bits, the stored-path hop tables,
 * one H x v (whole and as DW slices, which must agree) and an excitation
 * vector on a 3-site cluster with 3 bath levels (Nspin = 1), and the same
 * on a 2-site, 2-orbital cluster with Nspin = 2. */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex zc;
typedef struct {
  int nlat, norb, nspin, nlat_bath, norb_bath, hfmode;
  double uloc[5], ust, jh, xmu;
  const zc *hloc, *bath_h, *bath_v;
} orc_model;

int orc_build_map(int ns, int n, int32_t *map);
int orc_binary_search(const int32_t *a, int n, int32_t value);
int64_t orc_hxv(const orc_model *m, int nup, int ndw, const zc *v, zc *hv);
int64_t orc_hxv_slices(const orc_model *m, int nup, int ndw, const zc *v, zc *hv, int nslices,
                       int first, int count, int nthreads);
int64_t orc_spin_hops(const orc_model *m, int spin, int npart, int32_t *rows, int32_t *cols,
                      zc *amps);
int orc_diag(const orc_model *m, int nup, int ndw, zc *diag);
int orc_excite(const orc_model *m, int nup, int ndw, const zc *psi, int spin, int pos,
               int dagger, double coef_re, double coef_im, zc *out);

static int failures = 0;
static void expect(int ok, const char *what) {
  printf("%s %s\n", ok ? "ok  " : "FAIL", what);
  if (!ok) ++failures;
}

static int64_t binom(int n, int k) {
  int64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

static void exercise(const orc_model *m, int ns, int nup, int ndw, const char *name) {
  const int64_t du = binom(ns, nup), dd = binom(ns, ndw), dim = du * dd;
  zc *v = malloc(sizeof(zc) * dim), *hv = malloc(sizeof(zc) * dim),
     *hs = calloc(dim, sizeof(zc)), *dg = malloc(sizeof(zc) * dim);
  for (int64_t i = 0; i < dim; ++i) v[i] = sin(0.37 * i) + I * cos(0.11 * i);
  char what[128];
  snprintf(what, sizeof what, "%s: H x v", name);
  expect(orc_hxv(m, nup, ndw, v, hv) == 0, what);
  snprintf(what, sizeof what, "%s: DW slices equal one application", name);
  expect(orc_hxv_slices(m, nup, ndw, v, hs, 5, 0, 5, 1) == 0, what);
  double err = 0, nrm = 0;
  for (int64_t i = 0; i < dim; ++i) {
    err = fmax(err, cabs(hs[i] - hv[i]));
    nrm = fmax(nrm, cabs(hv[i]));
  }
  expect(err <= 1e-12 * nrm, what);
  snprintf(what, sizeof what, "%s: diagonal", name);
  expect(orc_diag(m, nup, ndw, dg) == 0, what);
  for (int spin = 1; spin <= 2; ++spin) {
    const int np = spin == 1 ? nup : ndw;
    const int64_t nnz = orc_spin_hops(m, spin, np, NULL, NULL, NULL);
    int32_t *r = malloc(sizeof(int32_t) * (nnz + 1)), *c = malloc(sizeof(int32_t) * (nnz + 1));
    zc *a = malloc(sizeof(zc) * (nnz + 1));
    snprintf(what, sizeof what, "%s: spin %d hop table", name, spin);
    expect(orc_spin_hops(m, spin, np, r, c, a) == nnz, what);
    free(r);
    free(c);
    free(a);
  }
  if (nup < ns) {
    const int64_t dj = binom(ns, nup + 1) * dd;
    zc *out = calloc(dj, sizeof(zc));
    snprintf(what, sizeof what, "%s: excitation c+_1", name);
    expect(orc_excite(m, nup, ndw, v, 1, 1, 1, 1.0, 0.0, out) == 0, what);
    free(out);
  }
  free(v);
  free(hv);
  free(hs);
  free(dg);
}

int main(void) {
  for (int n = 0; n <= 12; ++n) {
    const int d = orc_build_map(12, n, NULL);
    int32_t *map = malloc(sizeof(int32_t) * (d + 1));
    orc_build_map(12, n, map);
    int ok = d == binom(12, n);
    for (int i = 0; i < d; ++i) ok = ok && orc_binary_search(map, d, map[i]) == i + 1;
    ok = ok && orc_binary_search(map, d, -1) == 0;
    char what[64];
    snprintf(what, sizeof what, "map and binary search, 12 bits, n = %d", n);
    expect(ok, what);
    free(map);
  }
  {  /* 3 sites + 3 bath levels, Nspin = 1 */
    zc hloc[9] = {0}, bh[9] = {0}, bv[9] = {0};
    for (int i = 0; i < 2; ++i) hloc[i + 3 * (i + 1)] = hloc[(i + 1) + 3 * i] = -1.0;
    for (int i = 0; i < 3; ++i) {
      bh[i + 3 * i] = 0.4 * (i - 1);
      for (int j = 0; j < 3; ++j) bv[i + 3 * j] = 0.3 + 0.1 * i - 0.05 * j;
    }
    bh[1] = bh[3] = 0.15;
    orc_model m = {3, 1, 1, 3, 1, 1, {3.0, 0, 0, 0, 0}, 0, 0, 0.1, hloc, bh, bv};
    exercise(&m, 6, 3, 2, "bath model");
  }
  {  /* 2 sites x 2 orbitals, Nspin = 2, Ust, Jh, complex hopping */
    zc hloc[2 * 2 * 2 * 2 * 2 * 2] = {0};
#define H6(il, jl, s, t, io, jo) hloc[(il) + 2 * ((jl) + 2 * ((s) + 2 * ((t) + 2 * ((io) + 2 * (jo)))))]
    for (int s = 0; s < 2; ++s)
      for (int io = 0; io < 2; ++io)
        for (int jo = 0; jo < 2; ++jo) {
          H6(0, 1, s, s, io, jo) = -0.8 + 0.2 * I * (io - jo);
          H6(1, 0, s, s, jo, io) = -0.8 - 0.2 * I * (io - jo);
          if (io == jo) H6(0, 0, s, s, io, io) = H6(1, 1, s, s, io, io) = 0.1 * (s ? -1 : 1);
        }
    zc one = 0;
    orc_model m = {2, 2, 2, 0, 0, 1, {3.0, 2.5, 0, 0, 0}, 1.2, 0.3, 0.0, hloc, &one, &one};
    exercise(&m, 4, 2, 2, "two-orbital model");
  }
  printf("%s\n", failures ? "FAILED" : "ALL OK");
  return failures ? 1 : 0;
}
