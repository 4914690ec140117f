// rows.hpp -- H x v one row per lane, for launches that cover many sectors.
//
// The matrix elements are those of the sector kernels (kernels.hip; the
// reference's HxV_local / HxV_up / HxV_dw, VCA_HAMILTONIAN/direct/): the
// row (iup, idw) gets the diagonal E_up + E_dw + cross(up, dw), the up-spin
// hops as gathers inside its column and the down-spin hops from the
// neighbouring columns, with the one-spin ELL tables of the sector
// (column | signed amplitude index << 21, k-major).  No LDS staging of
// columns, so any sector shape runs, and a workgroup's rows may belong to
// any sector: the kernels of the chain set (chains.hip) and of the batched
// eigensolver (eigs.hip) take one table set per work item.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "vca_internal.hpp"

namespace vca {

template <typename R> struct RC2;
template <> struct RC2<double> { using T = double2; };
template <> struct RC2<float> { using T = float2; };
template <typename R> using rcplx = typename RC2<R>::T;

// One sector's tables, for the precision of the launch.
struct RowTab {
  const RowMeta *meta_up;  // energy, string, hop count per up row
  const double *e_dw;
  const uint32_t *map_dw;
  const int32_t *cnt_dw;
  const uint32_t *ent_up, *ent_dw;  // k-major ELL [maxnnz][dim]
  const void *amp_up, *amp_dw;      // signed amplitudes, complex of the precision
  int namp_up, namp_dw;             // entries of the amplitude tables
  int dimup, dimdw;
};

inline RowTab row_tab(const Sector &js) {
  RowTab t{};
  const bool f64 = js.precision == VCA_FP64;
  t.meta_up = js.up.meta;
  t.e_dw = js.dw.energy;
  t.map_dw = js.dw.map;
  t.cnt_dw = js.dw.cnt;
  t.ent_up = js.up.ent;
  t.ent_dw = js.dw.ent;
  t.amp_up = f64 ? (const void *)js.up.amp64 : (const void *)js.up.amp32;
  t.amp_dw = f64 ? (const void *)js.dw.amp64 : (const void *)js.dw.amp32;
  t.namp_up = 2 * (js.up.nterms > 1 ? js.up.nterms : 1);
  t.namp_dw = 2 * (js.dw.nterms > 1 ? js.dw.nterms : 1);
  t.dimup = js.up.dim;
  t.dimdw = js.dw.dim;
  return t;
}

namespace rows {

__device__ inline double cross_e(const CrossTerm &c, uint32_t su, uint32_t sd) {
  if (c.norb == 1) return c.wx[0][0] * (double)__builtin_popcount(su & sd & c.site_mask);
  double e = 0;
  for (int io = 0; io < c.norb; ++io)
    for (int jo = 0; jo < c.norb; ++jo)
      e += c.wx[io][jo] * (double)__builtin_popcount((su >> io) & (sd >> jo) & c.site_mask);
  return e;
}

// Diagonal of row (iup, idw); mu receives the up row's record.
__device__ inline double diag(const RowTab &t, const CrossTerm &c, int iup, int idw,
                              RowMeta &mu) {
  mu = t.meta_up[iup];
  return mu.e + t.e_dw[idw] + cross_e(c, mu.s, t.map_dw[idw]);
}

// The amplitude tables into LDS (up spin at s_amu, down spin at s_amd =
// s_amu + namp_up); the caller's next barrier publishes them.
template <typename R, int BS>
__device__ inline void stage_amps(const RowTab &t, rcplx<R> *s_amu, rcplx<R> *s_amd) {
  using C = rcplx<R>;
  for (int i = threadIdx.x; i < t.namp_up; i += BS) s_amu[i] = static_cast<const C *>(t.amp_up)[i];
  for (int i = threadIdx.x; i < t.namp_dw; i += BS) s_amd[i] = static_cast<const C *>(t.amp_dw)[i];
}

template <typename R, bool REAL_AMP>
__device__ inline void cmac(R &ar, R &ai, const rcplx<R> &h, const rcplx<R> &y) {
  if (REAL_AMP) {
    ar = fma(h.x, y.x, ar);
    ai = fma(h.x, y.y, ai);
  } else {
    ar = fma(h.x, y.x, ar);
    ar = fma(-h.y, y.y, ar);
    ai = fma(h.x, y.y, ai);
    ai = fma(h.y, y.x, ai);
  }
}

template <typename R>
__device__ inline void mac_on(R &acc, const rcplx<R> &h, R y, bool on) {
  acc = fma(on ? h.x : (R)0, y, acc);
}

template <typename R, bool REAL_AMP>
__device__ inline void mac_on(R &ar, R &ai, const rcplx<R> &h, const rcplx<R> &y, bool on) {
  rcplx<R> hh = h;
  if (!on) hh.x = hh.y = (R)0;
  cmac<R, REAL_AMP>(ar, ai, hh, y);
}

// Hops of one row, kHop at a time: the chunk's table entries, then its
// gathered elements and amplitudes, are issued before any is summed (the
// row's loads overlap instead of waiting one by one).  Slots past the row's
// count are not read from the table (the ELL holds nothing there; a row
// may have no hops at all): they take entry 0 -- element 0 of the column
// or row 0's neighbour, amplitude 0, all in bounds -- weighted 0.
// E = R (real vectors) or rcplx<R>; vin is the sector vector, col its
// column idw; ar (and ai) accumulate unscaled.
constexpr int kHop = 4;

template <typename R, bool REAL_AMP, typename E>
__device__ inline void hops(const RowTab &t, const E *vin, const E *col, int iup, int idw, int cu,
                            int cd, const rcplx<R> *s_amu, const rcplx<R> *s_amd, R &ar, R &ai) {
  const int nmax = max(cu, cd);
  for (int q0 = 0; q0 < nmax; q0 += kHop) {
    uint32_t eu[kHop], ed[kHop];
#pragma unroll
    for (int j = 0; j < kHop; ++j) {
      const int q = q0 + j;
      eu[j] = q < cu ? t.ent_up[(int64_t)q * t.dimup + iup] : 0u;
      ed[j] = q < cd ? t.ent_dw[(int64_t)q * t.dimdw + idw] : 0u;
    }
    E yu[kHop], yd[kHop];
    rcplx<R> hu[kHop], hd[kHop];
#pragma unroll
    for (int j = 0; j < kHop; ++j) {
      yu[j] = col[eu[j] & kColMask];
      yd[j] = vin[(int64_t)(ed[j] & kColMask) * t.dimup + iup];
      hu[j] = s_amu[eu[j] >> kColBits];
      hd[j] = s_amd[ed[j] >> kColBits];
    }
#pragma unroll
    for (int j = 0; j < kHop; ++j) {
      if constexpr (std::is_same<E, R>::value) {
        mac_on<R>(ar, hu[j], yu[j], q0 + j < cu);
        mac_on<R>(ar, hd[j], yd[j], q0 + j < cd);
      } else {
        mac_on<R, REAL_AMP>(ar, ai, hu[j], yu[j], q0 + j < cu);
        mac_on<R, REAL_AMP>(ar, ai, hd[j], yd[j], q0 + j < cd);
      }
    }
  }
}

// The same sum for a row of a column staged in LDS (s_col, the column idw
// of vin): the up-spin gathers read LDS; the down-spin hop list is the
// column's, the same for every lane (uniform loads), each hop a coalesced
// load of the neighbour column.
template <typename R, bool REAL_AMP, typename E>
__device__ inline void hops_col(const RowTab &t, const E *vin, const E *s_col, int iup, int idw,
                                int cu, int cd, const rcplx<R> *s_amu, const rcplx<R> *s_amd,
                                R &ar, R &ai) {
  for (int q0 = 0; q0 < cu; q0 += kHop) {
    uint32_t eu[kHop];
#pragma unroll
    for (int j = 0; j < kHop; ++j) {
      const int q = q0 + j;
      eu[j] = q < cu ? t.ent_up[(int64_t)q * t.dimup + iup] : 0u;
    }
    E yu[kHop];
    rcplx<R> hu[kHop];
#pragma unroll
    for (int j = 0; j < kHop; ++j) {
      yu[j] = s_col[eu[j] & kColMask];
      hu[j] = s_amu[eu[j] >> kColBits];
    }
#pragma unroll
    for (int j = 0; j < kHop; ++j) {
      if constexpr (std::is_same<E, R>::value) mac_on<R>(ar, hu[j], yu[j], q0 + j < cu);
      else mac_on<R, REAL_AMP>(ar, ai, hu[j], yu[j], q0 + j < cu);
    }
  }
  for (int q0 = 0; q0 < cd; q0 += kHop) {
    uint32_t ed[kHop];
#pragma unroll
    for (int j = 0; j < kHop; ++j) {
      const int q = q0 + j;
      ed[j] = q < cd ? t.ent_dw[(int64_t)q * t.dimdw + idw] : 0u;
    }
    E yd[kHop];
    rcplx<R> hd[kHop];
#pragma unroll
    for (int j = 0; j < kHop; ++j) {
      yd[j] = vin[(int64_t)(ed[j] & kColMask) * t.dimup + iup];
      hd[j] = s_amd[ed[j] >> kColBits];
    }
#pragma unroll
    for (int j = 0; j < kHop; ++j) {
      if constexpr (std::is_same<E, R>::value) mac_on<R>(ar, hd[j], yd[j], q0 + j < cd);
      else mac_on<R, REAL_AMP>(ar, ai, hd[j], yd[j], q0 + j < cd);
    }
  }
}

}  // namespace rows
}  // namespace vca
