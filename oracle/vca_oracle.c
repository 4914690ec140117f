/*
 * vca_oracle.c -- CPU restatement of the reference VCA Hamiltonian path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP product
 * path is checked against, and the scalar CPU baseline that bench.py times
 * beside the GPU.  Nothing in the product library links or calls it.
 *
 * It restates, in plain C with the reference's own loop structure, the
 * Fock-space bookkeeping and the Hamiltonian matrix elements of QcmPlab/VCA:
 *   - sector maps            VCA_SETUP.f90:660-695   (build_sector)
 *   - binary search          VCA_SETUP.f90:946-963   (binary_search)
 *   - c / c^+ with signs     VCA_SETUP.f90:724-750   (c, cdg)
 *   - state layout           VCA_SETUP.f90:481-518   (i = iup + (idw-1)*DimUp)
 *   - diagonal elements      VCA_HAMILTONIAN/stored/H_cluster_local.f90:1-99
 *   - up-spin hops           VCA_HAMILTONIAN/stored/H_cluster_up.f90:1-89
 *   - down-spin hops         VCA_HAMILTONIAN/stored/H_cluster_dw.f90:1-87
 *   - bath extraction        VCA_HAMILTONIAN/diag_hybr_bath.f90:1-18
 *   - excitation vectors     VCA_GF_NORMAL.f90:184-198, 255-269, 624-660
 * The bath and hybridisation terms follow the stored path, which accumulates
 * them at the correct target index (the direct path of the reference writes
 * them to a stale index).  The H x v sweeps follow the direct path's loop
 * orders: local sweep, up sweep with jdw outer / jup inner, down sweep with
 * jup outer / jdw inner (VCA_HAMILTONIAN/direct/HxV_{local,up,dw}.f90).
 *
 * All arrays with several indices are Fortran column-major, 1-based in the
 * comments and 0-based in the code.
 */
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zc;

typedef struct orc_model {
  int32_t nlat, norb, nspin, nlat_bath, norb_bath;
  int32_t hfmode;
  double uloc[5];
  double ust, jh, xmu;
  const zc *hloc;   /* [nlat,nlat,nspin,nspin,norb,norb]                 */
  const zc *bath_h; /* [nlat_bath,nlat_bath,nspin,nspin,norb_bath,norb_bath]
                       (already shifted as vca_set_bath does)           */
  const zc *bath_v; /* [nlat,nlat_bath,nspin,nspin,norb,norb_bath]       */
} orc_model;

/* ---------------------------------------------------------------- indices */
static inline int ns_of(const orc_model *m) {
  return m->nlat * m->norb + m->nlat_bath * m->norb_bath; /* VCA_SETUP.f90:82 */
}
/* VCA_SETUP.f90:705-710, 1-based orbital position */
static inline int imp_state_index(const orc_model *m, int ilat, int iorb) {
  return iorb + (ilat - 1) * m->norb;
}
/* VCA_SETUP.f90:341-349, 1-based bath position */
static inline int bath_stride(const orc_model *m, int ilat_b, int iorb_b) {
  return m->nlat * m->norb + (ilat_b - 1) * m->norb_bath + iorb_b;
}
/* element (i1..i6) of a rank-6 column-major array with extents d1..d6 */
static inline int64_t idx6(int i1, int i2, int i3, int i4, int i5, int i6,
                           int d1, int d2, int d3, int d4, int d5) {
  return (int64_t)(i1 - 1) +
         (int64_t)d1 * ((i2 - 1) + (int64_t)d2 * ((i3 - 1) + (int64_t)d3 *
         ((i4 - 1) + (int64_t)d4 * ((i5 - 1) + (int64_t)d5 * (i6 - 1)))));
}
static inline zc hloc_at(const orc_model *m, int il, int jl, int is, int js,
                         int io, int jo) {
  return m->hloc[idx6(il, jl, is, js, io, jo, m->nlat, m->nlat, m->nspin,
                      m->nspin, m->norb)];
}
static inline zc bathh_at(const orc_model *m, int il, int jl, int is, int js,
                          int io, int jo) {
  return m->bath_h[idx6(il, jl, is, js, io, jo, m->nlat_bath, m->nlat_bath,
                        m->nspin, m->nspin, m->norb_bath)];
}
static inline zc bathv_at(const orc_model *m, int il, int jl, int is, int js,
                          int io, int jo) {
  return m->bath_v[idx6(il, jl, is, js, io, jo, m->nlat, m->nlat_bath,
                        m->nspin, m->nspin, m->norb)];
}

/* ------------------------------------------------------------ Fock space */
/* VCA_SETUP.f90:679-684: ascending integers in [0,2^ns) of popcount n */
int orc_build_map(int ns, int n, int32_t *map) {
  int dim = 0;
  for (int32_t s = 0; s < (1 << ns); ++s)
    if (__builtin_popcount((unsigned)s) == n) {
      if (map) map[dim] = s;
      ++dim;
    }
  return dim;
}

/* VCA_SETUP.f90:946-963: recursive halving with mid = size/2+1 (1-based);
 * returns the 1-based position of value, 0 when absent. */
static int bsearch_rec(const int32_t *a, int n, int32_t value) {
  if (n == 0) return 0;
  int mid = n / 2 + 1;
  if (a[mid - 1] > value) return bsearch_rec(a, mid - 1, value);
  if (a[mid - 1] < value) {
    int r = bsearch_rec(a + mid, n - mid, value);
    return r != 0 ? mid + r : 0;
  }
  return mid;
}
int orc_binary_search(const int32_t *a, int n, int32_t value) {
  return bsearch_rec(a, n, value);
}

/* VCA_SETUP.f90:724-737: annihilate at 1-based pos; returns -1 on invalid */
int orc_c(int pos, int32_t in, int32_t *out, double *sgn) {
  if (!((in >> (pos - 1)) & 1)) return -1;
  double f = 1.0;
  for (int l = 1; l <= pos - 1; ++l)
    if ((in >> (l - 1)) & 1) f = -f;
  *sgn = f;
  *out = in & ~(1 << (pos - 1));
  return 0;
}
/* VCA_SETUP.f90:739-750 */
int orc_cdg(int pos, int32_t in, int32_t *out, double *sgn) {
  if ((in >> (pos - 1)) & 1) return -1;
  double f = 1.0;
  for (int l = 1; l <= pos - 1; ++l)
    if ((in >> (l - 1)) & 1) f = -f;
  *sgn = f;
  *out = in | (1 << (pos - 1));
  return 0;
}

static inline int bit1(int32_t s, int pos) { return (s >> (pos - 1)) & 1; }

/* ------------------------------------------------- diagonal matrix element */
/* H_cluster_local.f90:11-90 for one (mup, mdw) pair */
static zc diag_element(const orc_model *m, int32_t mup, int32_t mdw) {
  const int nlat = m->nlat, norb = m->norb, ns_dw = m->nspin;
  zc htmp = 0;
  for (int il = 1; il <= nlat; ++il)
    for (int io = 1; io <= norb; ++io) {
      double nu = bit1(mup, imp_state_index(m, il, io));
      double nd = bit1(mdw, imp_state_index(m, il, io));
      htmp += hloc_at(m, il, il, 1, 1, io, io) * nu;
      htmp += hloc_at(m, il, il, ns_dw, ns_dw, io, io) * nd;
      htmp -= m->xmu * (nu + nd);
    }
  for (int il = 1; il <= nlat; ++il)
    for (int io = 1; io <= norb; ++io) {
      double nu = bit1(mup, imp_state_index(m, il, io));
      double nd = bit1(mdw, imp_state_index(m, il, io));
      htmp += m->uloc[io - 1] * nu * nd;
    }
  if (norb > 1) {
    for (int il = 1; il <= nlat; ++il)
      for (int io = 1; io <= norb; ++io)
        for (int jo = io + 1; jo <= norb; ++jo) {
          double nui = bit1(mup, imp_state_index(m, il, io));
          double ndi = bit1(mdw, imp_state_index(m, il, io));
          double nuj = bit1(mup, imp_state_index(m, il, jo));
          double ndj = bit1(mdw, imp_state_index(m, il, jo));
          htmp += m->ust * (nui * ndj + nuj * ndi);
        }
    for (int il = 1; il <= nlat; ++il)
      for (int io = 1; io <= norb; ++io)
        for (int jo = io + 1; jo <= norb; ++jo) {
          double nui = bit1(mup, imp_state_index(m, il, io));
          double ndi = bit1(mdw, imp_state_index(m, il, io));
          double nuj = bit1(mup, imp_state_index(m, il, jo));
          double ndj = bit1(mdw, imp_state_index(m, il, jo));
          htmp += (m->ust - m->jh) * (nui * nuj + ndi * ndj);
        }
  }
  if (m->hfmode) {
    for (int il = 1; il <= nlat; ++il)
      for (int io = 1; io <= norb; ++io) {
        double nu = bit1(mup, imp_state_index(m, il, io));
        double nd = bit1(mdw, imp_state_index(m, il, io));
        htmp += -0.5 * m->uloc[io - 1] * (nu + nd) + 0.25 * m->uloc[io - 1];
      }
    if (norb > 1)
      for (int il = 1; il <= nlat; ++il)
        for (int io = 1; io <= norb; ++io)
          for (int jo = io + 1; jo <= norb; ++jo) {
            double s = bit1(mup, imp_state_index(m, il, io)) +
                       bit1(mdw, imp_state_index(m, il, io)) +
                       bit1(mup, imp_state_index(m, il, jo)) +
                       bit1(mdw, imp_state_index(m, il, jo));
            htmp += -0.5 * m->ust * s + 0.25 * m->ust;
            htmp += -0.5 * (m->ust - m->jh) * s + 0.25 * (m->ust - m->jh);
          }
  }
  if (m->nlat_bath > 0 && m->norb_bath > 0) {
    for (int ib = 1; ib <= m->nlat_bath; ++ib)
      for (int ob = 1; ob <= m->norb_bath; ++ob) {
        int a = bath_stride(m, ib, ob);
        /* bath_diag(b,s,o) = bath_h(b,b,s,s,o,o): diag_hybr_bath.f90:10 */
        htmp += bathh_at(m, ib, ib, 1, 1, ob, ob) * bit1(mup, a);
        htmp += bathh_at(m, ib, ib, m->nspin, m->nspin, ob, ob) * bit1(mdw, a);
        htmp -= m->xmu * (double)(bit1(mup, a) + bit1(mdw, a));
      }
  }
  return htmp;
}

/* One-spin off-diagonal sweep for input string mj.  Calls emit(target
 * string, amplitude) for every matrix element H(target, mj) in the order of
 * H_cluster_up.f90:108-186 (ispin=1) or H_cluster_dw.f90:8-84 (ispin=nspin). */
typedef void (*emit_fn)(void *ctx, int32_t target, zc amp);

static void spin_hops(const orc_model *m, int spin_index, int32_t mj,
                      emit_fn emit, void *ctx) {
  const int nlat = m->nlat, norb = m->norb;
  int32_t k1, k2;
  double sg1, sg2;
  for (int il = 1; il <= nlat; ++il)
    for (int jl = 1; jl <= nlat; ++jl)
      for (int io = 1; io <= norb; ++io)
        for (int jo = 1; jo <= norb; ++jo) {
          int is = imp_state_index(m, il, io), js = imp_state_index(m, jl, jo);
          zc h = hloc_at(m, il, jl, spin_index, spin_index, io, jo);
          if (h != 0 && bit1(mj, js) == 1 && bit1(mj, is) == 0) {
            orc_c(js, mj, &k1, &sg1);
            orc_cdg(is, k1, &k2, &sg2);
            emit(ctx, k2, h * sg1 * sg2);
          }
        }
  if (m->nlat_bath > 0 && m->norb_bath > 0) {
    for (int il = 1; il <= m->nlat_bath; ++il)
      for (int jl = 1; jl <= m->nlat_bath; ++jl)
        for (int io = 1; io <= m->norb_bath; ++io)
          for (int jo = 1; jo <= m->norb_bath; ++jo) {
            int ia = bath_stride(m, il, io), ib = bath_stride(m, jl, jo);
            zc h = bathh_at(m, il, jl, spin_index, spin_index, io, jo);
            if (ia != ib && h != 0 && bit1(mj, ib) == 1 && bit1(mj, ia) == 0) {
              orc_c(ib, mj, &k1, &sg1);
              orc_cdg(ia, k1, &k2, &sg2);
              emit(ctx, k2, h * sg1 * sg2);
            }
          }
    for (int il = 1; il <= nlat; ++il)
      for (int io = 1; io <= norb; ++io)
        for (int jl = 1; jl <= m->nlat_bath; ++jl)
          for (int jo = 1; jo <= m->norb_bath; ++jo) {
            int ia = bath_stride(m, jl, jo), is = imp_state_index(m, il, io);
            /* diag_hybr(il,jl,s,io,jo) = bath_v(il,jl,s,s,io,jo) */
            zc v = bathv_at(m, il, jl, spin_index, spin_index, io, jo);
            if (v != 0 && bit1(mj, ia) == 1 && bit1(mj, is) == 0) {
              orc_c(ia, mj, &k1, &sg1);
              orc_cdg(is, k1, &k2, &sg2);
              emit(ctx, k2, v * sg1 * sg2);
            }
            if (v != 0 && bit1(mj, ia) == 0 && bit1(mj, is) == 1) {
              orc_c(is, mj, &k1, &sg1);
              orc_cdg(ia, k1, &k2, &sg2);
              emit(ctx, k2, conj(v) * sg1 * sg2);
            }
          }
  }
}

/* ------------------------------------------------------ sector and H x v */
typedef struct {
  int ns, nup, ndw, dimup, dimdw;
  int32_t *mapup, *mapdw;
} orc_sector;

static int sector_init(const orc_model *m, int nup, int ndw, orc_sector *s) {
  s->ns = ns_of(m);
  s->nup = nup;
  s->ndw = ndw;
  s->dimup = orc_build_map(s->ns, nup, NULL);
  s->dimdw = orc_build_map(s->ns, ndw, NULL);
  s->mapup = (int32_t *)malloc(sizeof(int32_t) * (size_t)s->dimup);
  s->mapdw = (int32_t *)malloc(sizeof(int32_t) * (size_t)s->dimdw);
  if (!s->mapup || !s->mapdw) return -1;
  orc_build_map(s->ns, nup, s->mapup);
  orc_build_map(s->ns, ndw, s->mapdw);
  return 0;
}
static void sector_free(orc_sector *s) {
  free(s->mapup);
  free(s->mapdw);
}

typedef struct {
  const orc_sector *s;
  const zc *v;
  zc *hv;
  int64_t jstate; /* input index j (0-based)   */
  int fixed;      /* the other spin's index     */
  int up;         /* 1 = up sweep, 0 = down     */
  int64_t bad;    /* count of targets not found */
} hxv_ctx;

static void emit_hxv(void *p, int32_t target, zc amp) {
  hxv_ctx *c = (hxv_ctx *)p;
  if (c->up) {
    int iup = orc_binary_search(c->s->mapup, c->s->dimup, target);
    if (iup == 0) { c->bad++; return; }
    int64_t i = (int64_t)(iup - 1) + (int64_t)c->fixed * c->s->dimup;
    c->hv[i] += amp * c->v[c->jstate];
  } else {
    int idw = orc_binary_search(c->s->mapdw, c->s->dimdw, target);
    if (idw == 0) { c->bad++; return; }
    int64_t i = (int64_t)c->fixed + (int64_t)(idw - 1) * c->s->dimup;
    c->hv[i] += amp * c->v[c->jstate];
  }
}

/* Hv = H v on sector (nup, ndw); the down sweep covers only the range
 * jdw in [dw_begin, dw_end) of input columns and the local/up sweeps the
 * same range of output columns -- with the full range this is one H x v,
 * with slices it is one worker's share of the reference DW decomposition
 * (VCA_HAMILTONIAN.f90:77-90).  Returns the number of hop targets absent
 * from the map (0 for a consistent model). */
static int64_t hxv_range(const orc_model *m, const orc_sector *s, const zc *v,
                         zc *hv, int dw_begin, int dw_end) {
  const int dimup = s->dimup;
  hxv_ctx c = {s, v, hv, 0, 0, 1, 0};
  /* local part: HxV_local.f90 loop over states */
  for (int idw = dw_begin; idw < dw_end; ++idw)
    for (int iup = 0; iup < dimup; ++iup) {
      int64_t i = iup + (int64_t)idw * dimup;
      hv[i] += diag_element(m, s->mapup[iup], s->mapdw[idw]) * v[i];
    }
  /* up part: jdw outer, jup inner (HxV_up.f90:1-2) */
  c.up = 1;
  for (int jdw = dw_begin; jdw < dw_end; ++jdw)
    for (int jup = 0; jup < dimup; ++jup) {
      c.jstate = jup + (int64_t)jdw * dimup;
      c.fixed = jdw;
      spin_hops(m, 1, s->mapup[jup], emit_hxv, &c);
    }
  return c.bad;
}

static int64_t hxv_down(const orc_model *m, const orc_sector *s, const zc *v,
                        zc *hv, int up_begin, int up_end) {
  const int dimup = s->dimup, dimdw = s->dimdw;
  hxv_ctx c = {s, v, hv, 0, 0, 0, 0};
  /* down part: jup outer, jdw inner (HxV_dw.f90:1-2) */
  for (int jup = up_begin; jup < up_end; ++jup)
    for (int jdw = 0; jdw < dimdw; ++jdw) {
      c.jstate = jup + (int64_t)jdw * dimup;
      c.fixed = jup;
      spin_hops(m, m->nspin, s->mapdw[jdw], emit_hxv, &c);
    }
  return c.bad;
}

int64_t orc_hxv(const orc_model *m, int nup, int ndw, const zc *v, zc *hv) {
  orc_sector s;
  if (sector_init(m, nup, ndw, &s)) return -1;
  int64_t dim = (int64_t)s.dimup * s.dimdw;
  memset(hv, 0, sizeof(zc) * (size_t)dim);
  int64_t bad = hxv_range(m, &s, v, hv, 0, s.dimdw);
  bad += hxv_down(m, &s, v, hv, 0, s.dimup);
  sector_free(&s);
  return bad;
}

/* Diagonal of H on the sector, in state order. */
int orc_diag(const orc_model *m, int nup, int ndw, zc *diag) {
  orc_sector s;
  if (sector_init(m, nup, ndw, &s)) return -1;
  for (int idw = 0; idw < s.dimdw; ++idw)
    for (int iup = 0; iup < s.dimup; ++iup)
      diag[iup + (int64_t)idw * s.dimup] =
          diag_element(m, s.mapup[iup], s.mapdw[idw]);
  sector_free(&s);
  return 0;
}

/* One-spin hop matrix in coordinate form, in the stored path's insertion
 * order: for each input string index j (0-based), each element (i, j, amp).
 * spin = 1 (up) or 2 (down, uses the nspin block).  Returns the count; when
 * rows/cols/amps are NULL only counts.  Target index -1 marks a target
 * absent from the map. */
typedef struct {
  const int32_t *map;
  int dim;
  int32_t *rows, *cols;
  zc *amps;
  int64_t n;
  int32_t j;
} coo_ctx;
static void emit_coo(void *p, int32_t target, zc amp) {
  coo_ctx *c = (coo_ctx *)p;
  if (c->rows) {
    c->rows[c->n] = orc_binary_search(c->map, c->dim, target) - 1;
    c->cols[c->n] = c->j;
    c->amps[c->n] = amp;
  }
  c->n++;
}
int64_t orc_spin_hops(const orc_model *m, int spin, int npart, int32_t *rows,
                      int32_t *cols, zc *amps) {
  int ns = ns_of(m);
  int dim = orc_build_map(ns, npart, NULL);
  int32_t *map = (int32_t *)malloc(sizeof(int32_t) * (size_t)dim);
  if (!map) return -1;
  orc_build_map(ns, npart, map);
  coo_ctx c = {map, dim, rows, cols, amps, 0, 0};
  for (int j = 0; j < dim; ++j) {
    c.j = j;
    spin_hops(m, spin == 1 ? 1 : m->nspin, map[j], emit_coo, &c);
  }
  free(map);
  return c.n;
}

/* ---------------------------------------------- multi-core DW-slice H x v */
/* The reference distributes a sector over MPI ranks by DW slices
 * (VCA_HAMILTONIAN.f90:77-90); the down sweep then runs on a transposed
 * vector.  Here each worker owns a contiguous range of output columns for
 * the local+up sweeps and a contiguous range of jup rows for the down sweep
 * (the same work split after the transpose), on shared memory.  Two workers
 * never write the same element: the local/up sweeps only write columns of
 * their own slice, the down sweep only rows of its own slice. */
#ifdef _OPENMP
#include <omp.h>
#endif
int64_t orc_hxv_parallel(const orc_model *m, int nup, int ndw, const zc *v,
                         zc *hv, int nworkers) {
  orc_sector s;
  if (sector_init(m, nup, ndw, &s)) return -1;
  int64_t dim = (int64_t)s.dimup * s.dimdw;
  memset(hv, 0, sizeof(zc) * (size_t)dim);
  int64_t bad = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nworkers) reduction(+ : bad)
  {
    int w = omp_get_thread_num(), nw = omp_get_num_threads();
    int b0 = (int)((int64_t)s.dimdw * w / nw), b1 = (int)((int64_t)s.dimdw * (w + 1) / nw);
    bad += hxv_range(m, &s, v, hv, b0, b1);
#pragma omp barrier
    int u0 = (int)((int64_t)s.dimup * w / nw), u1 = (int)((int64_t)s.dimup * (w + 1) / nw);
    bad += hxv_down(m, &s, v, hv, u0, u1);
  }
#else
  (void)nworkers;
  bad += hxv_range(m, &s, v, hv, 0, s.dimdw);
  bad += hxv_down(m, &s, v, hv, 0, s.dimup);
#endif
  sector_free(&s);
  return bad;
}

/* Slices [first, first + count) of an nslices-way split of one H x v, as
 * the reference's DW decomposition distributes it (VCA_HAMILTONIAN.f90:77-90):
 * slice k applies the local and up parts to the output columns of its DW
 * slice and the down part to the rows of its UP slice.  The slices run on
 * OpenMP threads (one slice per thread at a time), local and up parts of
 * every slice first, then the down parts, so that no two threads write one
 * element at once.  hv must hold zeros where the slices write.  A bounded
 * sample of a sector too large to apply whole on one core: count/nslices
 * of the work of one application.  Returns the number of absent targets. */
int64_t orc_hxv_slices(const orc_model *m, int nup, int ndw, const zc *v, zc *hv,
                       int nslices, int first, int count, int nthreads) {
  orc_sector s;
  if (nslices < 1 || first < 0 || count < 1 || first + count > nslices) return -1;
  if (sector_init(m, nup, ndw, &s)) return -1;
  int64_t bad = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static, 1) reduction(+ : bad)
#endif
  for (int k = first; k < first + count; ++k)
    bad += hxv_range(m, &s, v, hv, (int)((int64_t)s.dimdw * k / nslices),
                     (int)((int64_t)s.dimdw * (k + 1) / nslices));
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static, 1) reduction(+ : bad)
#endif
  for (int k = first; k < first + count; ++k)
    bad += hxv_down(m, &s, v, hv, (int)((int64_t)s.dimup * k / nslices),
                    (int)((int64_t)s.dimup * (k + 1) / nslices));
  (void)nthreads;
  sector_free(&s);
  return bad;
}

/* ------------------------------------------------------ excitation vectors */
/* out(j) += coef * sgn * psi(i) for every state i of sector (nup, ndw) on
 * which the operator acts, following VCA_GF_NORMAL.f90:184-198 (dagger=1,
 * cdg) and :255-269 (dagger=0, c).  spin = 1 or 2 selects the string the
 * operator acts on; pos is the 1-based orbital position.  The target sector
 * is (nup+-1, ndw) or (nup, ndw+-1).  Returns -1 on error. */
int orc_excite(const orc_model *m, int nup, int ndw, const zc *psi, int spin,
               int pos, int dagger, double coef_re, double coef_im, zc *out) {
  const zc coef = coef_re + I * coef_im;
  orc_sector si, sj;
  int jnup = nup, jndw = ndw, d = dagger ? 1 : -1;
  if (spin == 1) jnup += d; else jndw += d;
  int ns = ns_of(m);
  if (jnup < 0 || jnup > ns || jndw < 0 || jndw > ns) return -1;
  if (sector_init(m, nup, ndw, &si)) return -1;
  if (sector_init(m, jnup, jndw, &sj)) return -1;
  for (int64_t i = 0; i < (int64_t)si.dimup * si.dimdw; ++i) {
    int iup = (int)(i % si.dimup), idw = (int)(i / si.dimup);
    int32_t str = spin == 1 ? si.mapup[iup] : si.mapdw[idw];
    int32_t r;
    double sgn;
    if (dagger) {
      if (bit1(str, pos) != 0) continue;
      orc_cdg(pos, str, &r, &sgn);
    } else {
      if (bit1(str, pos) != 1) continue;
      orc_c(pos, str, &r, &sgn);
    }
    int64_t j;
    if (spin == 1) {
      int jup = orc_binary_search(sj.mapup, sj.dimup, r) - 1;
      j = jup + (int64_t)idw * sj.dimup;
    } else {
      int jdw = orc_binary_search(sj.mapdw, sj.dimdw, r) - 1;
      j = iup + (int64_t)jdw * sj.dimup;
    }
    out[j] += coef * sgn * psi[i];
  }
  sector_free(&si);
  sector_free(&sj);
  return 0;
}
