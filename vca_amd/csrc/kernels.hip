// kernels.hip -- CDNA4 (gfx950) kernels of the VCA solver core.
//
// H x v (reference: directMatVec_main, VCA_HAMILTONIAN_DIRECT_HxV.f90:42-78
// with HxV_local/HxV_up/HxV_dw, matrix elements of the stored path
// VCA_HAMILTONIAN/stored/H_cluster_{local,up,dw}.f90):
//   Hv(iup,idw) = [E_up(iup) + E_dw(idw) + cross(up,dw)] v(iup,idw)
//               + sum_q H_up(iup, jup_q) v(jup_q, idw)      (column gather)
//               + sum_q H_dw(idw, jdw_q) v(iup, jdw_q)      (column stream)
// One workgroup owns one output column idw (or several short ones).  The
// input column v(:, idw) is staged in LDS when it fits, so the up-spin
// gathers hit LDS; the down-spin terms read whole neighbouring columns
// v(:, jdw) with 64-lane coalesced loads, the hop list of the column being
// wavefront-uniform.  Fermion signs and target ranks come from the tables
// built by k_build_spin (popcount parity and combinatorial rank).
//
// Lanczos step (SciFortran sp_lanc_tridiag / sp_lanc_eigh as called at
// VCA_GF_NORMAL.f90:218-223 and VCA_DIAG.f90:172-184): vectors are kept
// unnormalised, u_k = nrm_k v_k; k_hxv computes H v_k = H u_k / nrm_k and
// alpha_k = <v_k|H v_k> with a deterministic grid reduction finished by the
// last workgroup; k_lanczos_update forms u_{k+1} = Hv - alpha v_k -
// beta_k v_{k-1} and beta_{k+1} = |u_{k+1}|.  Step index, alpha, beta and
// norms live in device memory, so a run of steps needs no host sync.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <type_traits>

#include "vca_internal.hpp"

namespace vca {

template <typename R> struct C2;
template <> struct C2<double> { using T = double2; };
template <> struct C2<float> { using T = float2; };
template <typename R> using cplx = typename C2<R>::T;

constexpr int kBlock = 256;
// dynamic LDS per workgroup for the staged column and the amplitude tables
constexpr size_t kLdsBudget = 60 * 1024;

// ------------------------------------------------------------ reductions
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum, result valid in thread 0.  blockDim.x must be kBlock.
__device__ inline double block_sum(double v) {
  __shared__ double s_part[kBlock / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) s_part[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) r += s_part[w];
  return r;
}

// Grid-wide sums are two-stage: every workgroup stores its partial with a
// plain store, and a kernel that runs after it (same stream) sums the
// partials in block order.  No device-scope fence inside a kernel: on
// MI355X the XCD L2s are not coherent with each other, so such a fence
// writes the L2 back, and one per workgroup stalls large grids.
__device__ inline void block_partial(double v, double *partials) {
  const double b = block_sum(v);
  if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

// Sum of nblk partials in block order (deterministic), valid in thread 0
// of a kBlock-lane workgroup.
__device__ inline double sum_partials(const double *p, int nblk) {
  double s = 0;
  for (int i = threadIdx.x; i < nblk; i += kBlock) s += p[i];
  return block_sum(s);
}

// ------------------------------------------------------------ tables ----
__global__ void k_build_spin(int ns, int npart, int dim, int nterms, const int2 *ab,
                             const uint32_t *binom, const double *eps,
                             const double *pair, double e_add, uint32_t *map,
                             double *energy, int32_t *cnt, uint32_t *ent) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  const uint32_t s = unrank_string(ns, npart, (uint32_t)i, binom);
  map[i] = s;
  energy[i] = spin_energy(s, ns, eps, pair) + e_add;
  cnt[i] = row_hops(s, nterms, ab, binom, ent + i, dim);
}

__device__ __host__ constexpr int align16(int n) { return (n + 15) & ~15; }

// LDS layout (byte offsets) of the staged one-column kernel k_hxv_buf: the
// namp signed gather amplitudes and a zero at 0, the staged column of dimup
// elements, then the column's down-spin hop list (maxnd amplitudes and
// offsets).  Every region starts on a 16-byte boundary.
struct BufLds {
  int col, dw_amp, dw_soff, total;
};
__device__ __host__ inline BufLds buf_lds(int namp, int dimup, int maxnd, int esz) {
  BufLds l;
  l.col = align16((namp + 1) * esz);
  l.dw_amp = l.col + align16(dimup * esz);
  l.dw_soff = l.dw_amp + align16(maxnd * esz);
  l.total = l.dw_soff + align16(maxnd * 4);
  return l;
}

// LDS layout of the signed-column form: the zero element at 0, the column
// scaled by +t, the column scaled by -t.
struct SgnLds {
  int colp, coln, total;
};
__device__ __host__ inline SgnLds sgn_lds(int dimup, int esz) {
  SgnLds l;
  l.colp = align16(esz);
  l.coln = l.colp + align16(dimup * esz);
  l.total = l.coln + align16(dimup * esz);
  return l;
}

// One-plane hop entries (element at o_col + f rsz) readdressed for the
// interleaved column pair of k_hxv_plane2 (element f at o_col + 2 f rsz):
// the element address ya becomes 2 ya - o_col, the amplitude bits stay.
__global__ void k_pair_addr(const uint32_t *ent4p, uint32_t *ent4q, int64_t n, uint32_t col) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t e = ent4p[i];
  ent4q[i] = (2 * (e & kLdsAddrMask) - col) | (e & ~kLdsAddrMask);
}

// Row records and 4-hop packing of the ELL table for the buffer-load H x v,
// each hop as the two LDS byte addresses its gather reads (element of the
// staged column, signed amplitude) for elements of esz bytes and namp
// amplitudes; padding entries read row 0 and the zero amplitude.  lds_form
// 0 (addresses would not fit the entry): entries left 0, never read;
// lds_form 2: signed-column entries (sgn_lds; every term's amplitude is the
// same real t, so amplitude 2k is +t and 2k + 1 is -t) as 16-bit indices of
// esz-byte elements, 4 per 8 bytes; padding entries index the zero element.
__global__ void k_pack_rows(int dim, int maxnnz, const uint32_t *map, const double *energy,
                            const int32_t *cnt, const uint32_t *ent, RowMeta *meta,
                            uint32_t *ent4, int esz, int namp, int lds_form) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  RowMeta m;
  m.e = energy[i];
  m.s = map[i];
  m.cnt = cnt[i];
  meta[i] = m;
  const uint32_t o_col = (uint32_t)buf_lds(namp, 0, 0, esz).col;
  const uint32_t pad = o_col | ((uint32_t)(namp * esz) << kLdsAddrBits);
  const int nc = (maxnnz + 3) / 4;
  for (int c = 0; c < nc; ++c)
    for (int t = 0; t < 4; ++t) {
      const int q = 4 * c + t;
      uint32_t w = 0u;
      if (lds_form == 2) {  // 16-bit element indices, 4 to 8 bytes
        uint32_t idx = 0u;
        if (q < m.cnt) {
          const SgnLds sl = sgn_lds(dim, esz);
          const uint32_t e = ent[(int64_t)q * dim + i];
          idx = (uint32_t)((((e >> kColBits) & 1u) ? sl.coln : sl.colp) / esz) + (e & kColMask);
        }
        reinterpret_cast<uint16_t *>(ent4)[((int64_t)c * dim + i) * 4 + t] = (uint16_t)idx;
        continue;
      } else if (lds_form) {
        w = pad;
        if (q < m.cnt) {
          const uint32_t e = ent[(int64_t)q * dim + i];
          w = (o_col + (e & kColMask) * (uint32_t)esz) |
              (((e >> kColBits) * (uint32_t)esz) << kLdsAddrBits);
        }
      }
      ent4[((int64_t)c * dim + i) * 4 + t] = w;
    }
}

// ------------------------------------------------------------ H x v -----
struct HxvArgs {
  const void *v0, *v1;  // input (Lanczos: u_k in buffer k&1)
  void *hv;
  int dimup, dimdw, cpb;
  const uint32_t *map_up, *map_dw;
  const double *e_up, *e_dw;
  const int32_t *cnt_up, *cnt_dw;
  const uint32_t *ent_up, *ent_dw;
  const void *amp_up, *amp_dw;
  int namp_up, namp_dw;
  CrossTerm cross;
  LanczosDev *lz;  // null: plain product; else one per chain (blockIdx.y)
  double *partials;
  int64_t vstride;  // elements between the vectors of consecutive chains
  int dw_maxnnz;    // capacity of the staged down-spin hop list
  const RowMeta *meta_up;
  const uint32_t *ent4_up;
  int up_nc;        // 4-hop chunks per row
  // chain-affine batches (one-column kernels): a 1-D grid of ncb x nb8
  // workgroups, nb8 = nchains rounded up to a multiple of 8
  int xcd_chains, ncb, nchains;
  // one-column kernels: the workgroup of column c of chain b stores its
  // partial sum at pbase + b * pstride + c; the passes of the transposed
  // form (skip_dw, skip_diag) share one row of partials per chain
  int64_t pbase, pstride;
  int skip_dw, skip_diag;
  int nt_store;  // Hv written with nontemporal stores (batches beyond the MALL)
};

__device__ inline double cross_energy(const CrossTerm &c, uint32_t su, uint32_t sd) {
  if (c.norb == 1) return c.wx[0][0] * (double)__builtin_popcount(su & sd & c.site_mask);
  double e = 0;
  for (int io = 0; io < c.norb; ++io)
    for (int jo = 0; jo < c.norb; ++jo)
      e += c.wx[io][jo] * (double)__builtin_popcount((su >> io) & (sd >> jo) & c.site_mask);
  return e;
}

// Element of type T at LDS byte address addr.  Used by k_hxv_buf, which has
// no static LDS, so that its dynamic LDS (smem) starts at address 0 and the
// hop tables' LDS addresses are read with no base added (launch_buf checks
// the kernel's static LDS size once).
template <typename R> struct Vec2;
template <> struct Vec2<double> { typedef double T __attribute__((ext_vector_type(2))); };
template <> struct Vec2<float> { typedef float T __attribute__((ext_vector_type(2))); };

template <typename R>
__device__ inline cplx<R> lds_load(uint32_t addr) {
  using V = typename Vec2<R>::T;
  const V v = *(const __attribute__((address_space(3))) V *)(size_t)addr;
  cplx<R> c;
  c.x = v.x;
  c.y = v.y;
  return c;
}

// Real part only (real amplitudes).
template <typename R>
__device__ inline R lds_load_re(uint32_t addr) {
  return *(const __attribute__((address_space(3))) R *)(size_t)addr;
}

template <typename R, bool REAL_AMP>
__device__ inline void cmac(R &ar, R &ai, const cplx<R> &h, const cplx<R> &y) {
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

// Copies n elements into LDS with each lane's loads of a chunk of U
// elements issued before any of its stores (indices clamped to n - 1, so
// the loads are unconditional and need no branch): the HBM latency of a
// staged column is paid once per chunk instead of once per element.
template <int BS, typename T>
constexpr int stage_unroll() {
  return sizeof(T) >= 16 ? 8 : 16;
}

template <int BS, typename T, typename Ld, typename St>
__device__ inline void stage_copy(int n, Ld ld, St st) {
  constexpr int U = stage_unroll<BS, T>();
  for (int f0 = 0; f0 < n; f0 += U * BS) {
    T x[U];
#pragma unroll
    for (int q = 0; q < U; ++q) x[q] = ld(min(f0 + (int)threadIdx.x + q * BS, n - 1));
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int f = f0 + (int)threadIdx.x + q * BS;
      if (f < n) st(f, x[q]);
    }
  }
}

template <typename R, bool REAL_AMP, bool STAGE, bool ONECOL>
__global__ __launch_bounds__(kBlock) void k_hxv(const HxvArgs a) {
  using C = cplx<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  C *s_amp_up = reinterpret_cast<C *>(smem);
  C *s_amp_dw = s_amp_up + a.namp_up;
  C *s_col = s_amp_dw + a.namp_dw;

  const C *g_amp_up = static_cast<const C *>(a.amp_up);
  const C *g_amp_dw = static_cast<const C *>(a.amp_dw);
  for (int t = threadIdx.x; t < a.namp_up; t += kBlock) s_amp_up[t] = g_amp_up[t];
  for (int t = threadIdx.x; t < a.namp_dw; t += kBlock) s_amp_dw[t] = g_amp_dw[t];

  const int64_t chain_off = (int64_t)blockIdx.y * a.vstride;
  LanczosDev *lz = a.lz ? a.lz + blockIdx.y : nullptr;
  int k = 0;
  bool done = false;
  R scale = (R)1;
  const C *vin = static_cast<const C *>(a.v0) + chain_off;
  if (lz) {
    k = lz->k;
    done = lz->done != 0;
    if (k & 1) vin = static_cast<const C *>(a.v1) + chain_off;
    if (!done) scale = (R)(1.0 / lz->nrm[k]);
  }
  C *hv = static_cast<C *>(a.hv) + chain_off;
  const int dimup = a.dimup;
  const int64_t col0 = (int64_t)blockIdx.x * a.cpb;
  const int ncol = (int)min((int64_t)a.cpb, (int64_t)a.dimdw - col0);
  const int nel = ncol * dimup;
  const C *vcol = vin + col0 * dimup;
  double dot = 0.0;

  if (!done) {
    if (STAGE)
      stage_copy<kBlock, C>(nel, [&](int f) { return vcol[f]; },
                            [&](int f, const C &x) { s_col[f] = x; });
    __syncthreads();
    for (int f = threadIdx.x; f < nel; f += kBlock) {
      int c, iup;
      if (ONECOL) {
        c = 0;
        iup = f;
      } else {
        c = f / dimup;
        iup = f - c * dimup;
      }
      const int idw = (int)col0 + c;
      const C x = STAGE ? s_col[f] : vcol[f];
      const double ed = a.e_up[iup] + a.e_dw[idw] +
                        cross_energy(a.cross, a.map_up[iup], a.map_dw[idw]);
      R ar = (R)ed * x.x, ai = (R)ed * x.y;
      // up-spin hops: gather inside the column
      const int nu = a.cnt_up[iup];
      const C *colp = STAGE ? (s_col + c * dimup) : (vcol + (int64_t)c * dimup);
      for (int q = 0; q < nu; ++q) {
        const uint32_t e = a.ent_up[(int64_t)q * dimup + iup];
        const C y = colp[e & kColMask];
        cmac<R, REAL_AMP>(ar, ai, s_amp_up[e >> kColBits], y);
      }
      // down-spin hops: whole neighbouring columns, coalesced in iup
      const int nd = a.cnt_dw[idw];
      for (int q = 0; q < nd; ++q) {
        const uint32_t e = a.ent_dw[(int64_t)q * a.dimdw + idw];
        const C y = vin[(int64_t)(e & kColMask) * dimup + iup];
        cmac<R, REAL_AMP>(ar, ai, s_amp_dw[e >> kColBits], y);
      }
      C out;
      out.x = ar * scale;
      out.y = ai * scale;
      hv[col0 * dimup + f] = out;
      if (lz) dot += (double)(x.x * scale) * (double)out.x + (double)(x.y * scale) * (double)out.y;
    }
  }
  if (lz) block_partial(dot, a.partials + (int64_t)blockIdx.y * gridDim.x);
}

// Block sum for a block of BS threads, result valid in thread 0.
template <int BS>
__device__ inline double block_sum_t(double v) {
  __shared__ double s_part[BS / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) s_part[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < BS / 64; ++w) r += s_part[w];
  return r;
}

template <int BS>
__device__ inline void block_partial_t(double v, double *partials, int64_t slot) {
  const double b = block_sum_t<BS>(v);
  if (threadIdx.x == 0) partials[slot] = b;
}

// The same with the per-wave sums in caller-provided LDS (BS/64 doubles,
// free once every wave has arrived): a kernel with no static LDS has its
// dynamic LDS at address 0, so LDS addresses held in tables need no base.
template <int BS>
__device__ inline void block_partial_lds(double v, double *partials, int64_t slot,
                                         double *s_part) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) s_part[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0;
    for (int w = 0; w < BS / 64; ++w) r += s_part[w];
    partials[slot] = r;
  }
}

// Workgroup b -> output column.  Workgroups are dealt round-robin to the 8
// XCDs, so column ranges are assigned per XCD: each XCD's L2 then holds
// the neighbour columns of one contiguous range of columns.
__device__ inline int xcd_column(int b, int ncols) {
  const int x = b & 7, pos = b >> 3;
  const int q = ncols >> 3, r = ncols & 7;
  return x * q + min(x, r) + pos;
}

// Workgroup -> (output column, chain, slot of its partial sum) of the
// one-column kernels.  Default grid (columns, chains): the columns of each
// chain are spread over the 8 XCDs by xcd_column.  Chain-affine batches
// (xcd_chains): workgroup b runs on XCD b & 7 and takes, in order, the
// columns of the chains (b & 7) + 8q, so each chain's vector is read
// through one XCD's L2 with its columns in rank order, instead of through
// all eight L2s at once.  Workgroups past the last chain have nothing to do.
struct WgPos {
  int idw, chain;
  int64_t slot;
  bool valid;
};
__device__ inline WgPos wg_pos(const HxvArgs &a) {
  WgPos w;
  if (!a.xcd_chains) {
    w.idw = xcd_column(blockIdx.x, a.dimdw);
    w.chain = blockIdx.y;
    w.slot = a.pbase + (int64_t)blockIdx.y * a.pstride + blockIdx.x;
    w.valid = true;
  } else {
    const int x = blockIdx.x & 7;
    const int p = blockIdx.x >> 3;
    const int q = p / a.ncb;
    w.idw = p - q * a.ncb;
    w.chain = x + 8 * q;
    w.slot = a.pbase + (int64_t)w.chain * a.pstride + w.idw;
    w.valid = w.chain < a.nchains;
  }
  return w;
}


// One workgroup of BS lanes per output column idw (DimUp >= BS/2):
//  * the down-spin hop list of the column (<= maxnnz entries) is staged in
//    LDS as (column offset, signed amplitude);
//  * the input column v(:, idw) is staged in LDS when it fits (STAGE), so
//    the diagonal and the up-spin gathers read LDS;
//  * each lane handles KR rows at once, so the down-spin loop issues KR
//    independent coalesced loads of v(:, jdw) per hop.
// BS (lanes) and KR (rows per lane) are picked per sector shape by the
// autotuner (hxv_tune).

template <typename R, bool REAL_AMP, bool STAGE, int BS, int KR, int U>
__global__ __launch_bounds__(BS) void k_hxv_col(const HxvArgs a) {
  using C = cplx<R>;
  // U hops per batch of independent loads: KR*U loads in flight per lane
  extern __shared__ __align__(16) unsigned char smem[];
  const int maxnd = a.dw_maxnnz;
  // every region starts on a 16-byte boundary (128-bit LDS accesses); the
  // regions are byte offsets from smem so the compiler keeps LDS accesses
  const int o_dw_amp = align16(a.namp_up * (int)sizeof(C));
  const int o_dw_off = o_dw_amp + align16(maxnd * (int)sizeof(C));
  const int o_col = o_dw_off + align16(maxnd * (int)sizeof(int64_t));
  C *s_amp_up = reinterpret_cast<C *>(smem);
  C *s_dw_amp = reinterpret_cast<C *>(smem + o_dw_amp);
  int64_t *s_dw_off = reinterpret_cast<int64_t *>(smem + o_dw_off);
  C *s_col = reinterpret_cast<C *>(smem + o_col);

  const int dimup = a.dimup;
  const WgPos wp = wg_pos(a);
  if (!wp.valid) return;
  const int idw = wp.idw;
  const int64_t chain_off = (int64_t)wp.chain * a.vstride;
  LanczosDev *lz = a.lz ? a.lz + wp.chain : nullptr;
  int k = 0;
  bool done = false;
  R scale = (R)1;
  const C *vin = static_cast<const C *>(a.v0) + chain_off;
  if (lz) {
    k = lz->k;
    done = lz->done != 0;
    if (k & 1) vin = static_cast<const C *>(a.v1) + chain_off;
    if (!done) scale = (R)(1.0 / lz->nrm[k]);
  }
  C *hv = static_cast<C *>(a.hv) + chain_off + (int64_t)idw * dimup;
  const C *vcol = vin + (int64_t)idw * dimup;
  double dot = 0.0;

  if (!done) {
    // one staging phase: amplitudes, the column's down-spin hop list and
    // (STAGE) the input column, then one barrier
    const C *g_amp_up = static_cast<const C *>(a.amp_up);
    const C *g_amp_dw = static_cast<const C *>(a.amp_dw);
    const int nd = a.skip_dw ? 0 : a.cnt_dw[idw];  // same address in every lane
    for (int t = threadIdx.x; t < a.namp_up; t += BS) s_amp_up[t] = g_amp_up[t];
    for (int q = threadIdx.x; q < nd; q += BS) {
      const uint32_t e = a.ent_dw[(int64_t)q * a.dimdw + idw];
      s_dw_off[q] = (int64_t)(e & kColMask) * dimup;
      s_dw_amp[q] = g_amp_dw[e >> kColBits];
    }
    if (STAGE)
      stage_copy<BS, C>(dimup, [&](int f) { return vcol[f]; },
                        [&](int f, const C &x) { s_col[f] = x; });
    __syncthreads();
    const double edw = a.skip_diag ? 0.0 : a.e_dw[idw];
    const uint32_t sdw = a.skip_diag ? 0u : a.map_dw[idw];
    for (int base = threadIdx.x; base < dimup; base += BS * KR) {
      int row[KR];
      R ar[KR], ai[KR];
      C x[KR];
      int nu[KR];
#pragma unroll
      for (int j = 0; j < KR; ++j) {
        row[j] = min(base + j * BS, dimup - 1);  // clamped: loads stay in bounds
        nu[j] = (base + j * BS < dimup) ? a.cnt_up[row[j]] : 0;
        x[j] = STAGE ? s_col[row[j]] : vcol[row[j]];
        const double ed =
            a.skip_diag ? 0.0
                        : a.e_up[row[j]] + edw + cross_energy(a.cross, a.map_up[row[j]], sdw);
        ar[j] = (R)ed * x[j].x;
        ai[j] = (R)ed * x[j].y;
      }
      // up-spin hops, U hop entries of every row fetched before use
      int nu_max = 0;
#pragma unroll
      for (int j = 0; j < KR; ++j) nu_max = max(nu_max, nu[j]);
      for (int q0 = 0; q0 < nu_max; q0 += U) {
        uint32_t e[U][KR];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int j = 0; j < KR; ++j) {
            const int q = min(q0 + u, max(nu[j] - 1, 0));
            e[u][j] = a.ent_up[(int64_t)q * dimup + row[j]];
          }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int j = 0; j < KR; ++j)
            if (q0 + u < nu[j]) {
              const uint32_t ee = e[u][j];
              const C y = STAGE ? s_col[ee & kColMask] : vcol[ee & kColMask];
              cmac<R, REAL_AMP>(ar[j], ai[j], s_amp_up[ee >> kColBits], y);
            }
      }
      // down-spin hops: U hops x KR rows of independent coalesced loads
      for (int q0 = 0; q0 < nd; q0 += U) {
        C y[U][KR];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int64_t off = s_dw_off[min(q0 + u, nd - 1)];
#pragma unroll
          for (int j = 0; j < KR; ++j) y[u][j] = vin[off + row[j]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (q0 + u < nd) {
            const C h = s_dw_amp[q0 + u];
#pragma unroll
            for (int j = 0; j < KR; ++j) cmac<R, REAL_AMP>(ar[j], ai[j], h, y[u][j]);
          }
      }
#pragma unroll
      for (int j = 0; j < KR; ++j) {
        if (base + j * BS < dimup) {
          C out;
          out.x = ar[j] * scale;
          out.y = ai[j] * scale;
          hv[row[j]] = out;
          if (lz)
            dot += (double)(x[j].x * scale) * (double)out.x +
                   (double)(x[j].y * scale) * (double)out.y;
        }
      }
    }
  }
  if (lz) block_partial_t<BS>(dot, a.partials, wp.slot);
}

// Buffer-load form of the one-column kernel.  The same arithmetic as
// k_hxv_col, with fewer and cheaper instructions:
//  * neighbour-column loads are raw buffer loads with the row offset in a
//    VGPR fixed per lane and the column offset in an SGPR (read first
//    lane of the staged hop list), so no per-load address arithmetic;
//  * each row's energy, string and hop count come in one 16-byte record;
//  * up-spin hop entries come 4 per 16-byte load.
// The input column is always staged in LDS (the caller checks it fits).
template <typename R, bool REAL_AMP, int BS, int KR>
__global__ __launch_bounds__(BS) void k_hxv_buf(const HxvArgs a) {
  using C = cplx<R>;
  constexpr int ESZ = sizeof(C);
  constexpr int U = KR >= 4 ? 4 : 8;
  extern __shared__ __align__(16) unsigned char smem[];
  const int maxnd = a.dw_maxnnz;
  const BufLds lds = buf_lds(a.namp_up, a.dimup, maxnd, ESZ);
  C *s_amp_up = reinterpret_cast<C *>(smem);
  C *s_col = reinterpret_cast<C *>(smem + lds.col);
  C *s_dw_amp = reinterpret_cast<C *>(smem + lds.dw_amp);
  uint32_t *s_dw_soff = reinterpret_cast<uint32_t *>(smem + lds.dw_soff);

  const int dimup = a.dimup;
  const WgPos wp = wg_pos(a);
  if (!wp.valid) return;
  const int idw = wp.idw;
  const int64_t chain_off = (int64_t)wp.chain * a.vstride;
  LanczosDev *lz = a.lz ? a.lz + wp.chain : nullptr;
  int k = 0;
  bool done = false;
  R scale = (R)1;
  const C *vin = static_cast<const C *>(a.v0) + chain_off;
  if (lz) {
    k = lz->k;
    done = lz->done != 0;
    if (k & 1) vin = static_cast<const C *>(a.v1) + chain_off;
    if (!done) scale = (R)(1.0 / lz->nrm[k]);
  }
  C *hv = static_cast<C *>(a.hv) + chain_off + (int64_t)idw * dimup;
  const C *vcol = vin + (int64_t)idw * dimup;
  double dot = 0.0;

  if (!done) {
    const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<C *>(vin), (short)0, (int)(uint32_t)(a.vstride * ESZ), 0x00020000);
    const C *g_amp_up = static_cast<const C *>(a.amp_up);
    const C *g_amp_dw = static_cast<const C *>(a.amp_dw);
    const int nd = a.skip_dw ? 0 : a.cnt_dw[idw];
    for (int t = threadIdx.x; t < a.namp_up; t += BS) s_amp_up[t] = g_amp_up[t];
    if (threadIdx.x == 0) s_amp_up[a.namp_up] = C{(R)0, (R)0};  // padding hops
    for (int q = threadIdx.x; q < nd; q += BS) {
      const uint32_t e = a.ent_dw[(int64_t)q * a.dimdw + idw];
      s_dw_soff[q] = (e & kColMask) * (uint32_t)dimup * (uint32_t)ESZ;
      s_dw_amp[q] = g_amp_dw[e >> kColBits];
    }
    stage_copy<BS, C>(dimup, [&](int f) { return vcol[f]; },
                      [&](int f, const C &x) { s_col[f] = x; });
    __syncthreads();
    const double edw = a.skip_diag ? 0.0 : a.e_dw[idw];
    const uint32_t sdw = a.skip_diag ? 0u : a.map_dw[idw];
    // a wavefront runs a row block while any of its lanes has a row there
    // (its other lanes take a clamped row with no hops and store nothing),
    // so the wavefront-wide reductions below see all 64 lanes
    for (int base0 = 0; base0 < dimup; base0 += BS * KR) {
      if (base0 + (int)(threadIdx.x & ~63u) >= dimup) break;
      const int base = base0 + (int)threadIdx.x;
      int row[KR];
      R ar[KR], ai[KR];
      C x[KR];
      int nu[KR];
#pragma unroll
      for (int j = 0; j < KR; ++j) {
        row[j] = min(base + j * BS, dimup - 1);
        const RowMeta m = a.meta_up[row[j]];
        nu[j] = (base + j * BS < dimup) ? m.cnt : 0;
        x[j] = s_col[row[j]];
        const double ed = a.skip_diag ? 0.0 : m.e + edw + cross_energy(a.cross, m.s, sdw);
        ar[j] = (R)ed * x[j].x;
        ai[j] = (R)ed * x[j].y;
      }
      // up-spin hops: 4 entries per 16-byte load, each entry the LDS byte
      // addresses of its element and signed amplitude, so a gather costs a
      // mask, a shift and two LDS reads.  The chunk count is the maximum
      // over the wavefront (uniform loop); entries past a row's hops are
      // padding that read row 0 and the zero amplitude.  The next chunk's
      // entries are loaded while this one is summed.
      int nc_w = 0;
#pragma unroll
      for (int j = 0; j < KR; ++j) nc_w = max(nc_w, (nu[j] + 3) >> 2);
      for (int off = 32; off > 0; off >>= 1) nc_w = max(nc_w, __shfl_xor(nc_w, off, 64));
      nc_w = __builtin_amdgcn_readfirstlane(nc_w);
      // The entries (L2 hits: every workgroup reads the same table) come
      // through a ring of PF chunks in flight, so a chunk's load has PF - 1
      // chunks of gathers to hide its latency behind.  Every ring load is
      // issued, the chunk index clamped to the table's last chunk, so the
      // count of loads in flight is the same on every path and the waits
      // stay partial.  Depth 4 costs ~10 VGPRs over depth 2: taken only by
      // the 1024-lane shape (picked where a large staged column holds one
      // workgroup per CU anyway); at 256/512 lanes 64 VGPRs keep 8
      // wavefronts per SIMD (the 3x4 sector's 924 workgroups in one round).
      constexpr int PF = (BS == 1024 && KR == 1) ? 4 : 2;
      const int last = a.up_nc - 1;  // >= 0 whenever nc_w > 0
      const uint4 *ent4 = reinterpret_cast<const uint4 *>(a.ent4_up);
      if (nc_w > 0) {
        uint4 e4[PF][KR];
#pragma unroll
        for (int p = 0; p < PF; ++p)
#pragma unroll
          for (int j = 0; j < KR; ++j) e4[p][j] = ent4[(int64_t)min(p, last) * dimup + row[j]];
        for (int c0 = 0; c0 < nc_w; c0 += PF) {
#pragma unroll
          for (int p = 0; p < PF; ++p) {
            const int c = c0 + p;
            if (c < nc_w) {
#pragma unroll
              for (int j = 0; j < KR; ++j) {
                const uint32_t ee[4] = {e4[p][j].x, e4[p][j].y, e4[p][j].z, e4[p][j].w};
                C y[4], h[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                  y[t] = lds_load<R>(ee[t] & kLdsAddrMask);
                  if constexpr (REAL_AMP) {
                    h[t].x = lds_load_re<R>(ee[t] >> kLdsAddrBits);
                    h[t].y = (R)0;
                  } else {
                    h[t] = lds_load<R>(ee[t] >> kLdsAddrBits);
                  }
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) cmac<R, REAL_AMP>(ar[j], ai[j], h[t], y[t]);
              }
            }
            const int cn = min(c + PF, last);
#pragma unroll
            for (int j = 0; j < KR; ++j) e4[p][j] = ent4[(int64_t)cn * dimup + row[j]];
          }
        }
      }
      // down-spin hops: buffer loads, column offset in an SGPR.  complex64:
      // two stages of U/2 hops, one stage's loads issued before the other
      // stage is summed, so the loads never wait on the sums (16.5 against
      // 17.5 us on the 3x4 sector).  complex128: one batch of U hops at a
      // time; two stages there cost more registers than they hide latency
      // (22.8 against 19.9 us).
      if constexpr (ESZ == 16) {
        for (int q0 = 0; q0 < nd; q0 += U) {
          C y[U][KR];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const uint32_t so = __builtin_amdgcn_readfirstlane(s_dw_soff[min(q0 + u, nd - 1)]);
#pragma unroll
            for (int j = 0; j < KR; ++j)
              y[u][j] = __builtin_bit_cast(
                  C, __builtin_amdgcn_raw_buffer_load_b128(rv, row[j] * ESZ, so, 0));
          }
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (q0 + u < nd) {
              const C h = s_dw_amp[q0 + u];
#pragma unroll
              for (int j = 0; j < KR; ++j) cmac<R, REAL_AMP>(ar[j], ai[j], h, y[u][j]);
            }
        }
      } else {
        constexpr int UH = U / 2;
        C ya[UH][KR], yb[UH][KR];
        auto load_stage = [&](C(&y)[UH][KR], int q0) {
#pragma unroll
          for (int u = 0; u < UH; ++u) {
            const uint32_t so = __builtin_amdgcn_readfirstlane(s_dw_soff[min(q0 + u, nd - 1)]);
#pragma unroll
            for (int j = 0; j < KR; ++j)
              y[u][j] = __builtin_bit_cast(
                  C, __builtin_amdgcn_raw_buffer_load_b64(rv, row[j] * ESZ, so, 0));
          }
        };
        auto sum_stage = [&](const C(&y)[UH][KR], int q0) {
#pragma unroll
          for (int u = 0; u < UH; ++u)
            if (q0 + u < nd) {
              const C h = s_dw_amp[q0 + u];
#pragma unroll
              for (int j = 0; j < KR; ++j) cmac<R, REAL_AMP>(ar[j], ai[j], h, y[u][j]);
            }
        };
        if (nd > 0) load_stage(ya, 0);
        for (int q0 = 0; q0 < nd; q0 += 2 * UH) {
          if (q0 + UH < nd) load_stage(yb, q0 + UH);
          sum_stage(ya, q0);
          if (q0 + 2 * UH < nd) load_stage(ya, q0 + 2 * UH);
          if (q0 + UH < nd) sum_stage(yb, q0 + UH);
        }
      }
#pragma unroll
      for (int j = 0; j < KR; ++j) {
        if (base + j * BS < dimup) {
          C out;
          out.x = ar[j] * scale;
          out.y = ai[j] * scale;
          hv[row[j]] = out;
          if (lz)
            dot += (double)(x[j].x * scale) * (double)out.x +
                   (double)(x[j].y * scale) * (double)out.y;
        }
      }
    }
  }
  // the per-wave sums reuse the start of the LDS (at least 128 B + a column)
  if (lz) block_partial_lds<BS>(dot, a.partials, wp.slot, reinterpret_cast<double *>(smem));
}

// Gather pass of the transposed form (skip_dw) on one real plane at a time,
// for real amplitudes: the real and imaginary parts of v transform
// independently, so NPL = 2 planes of a complex vector are staged in LDS one
// after the other (a 12,870-row complex128 column, 206 KB, does not fit the
// LDS; one plane, 103 KB, does) and NPL = 1 is a real vector.  One workgroup
// per column, BS lanes, one row per lane per block; LDS: the namp real
// amplitudes and a zero at 0, then the staged plane, addressed by the
// one-plane hop table (DevSpin::ent4p).  Each plane's results go straight
// to that plane of the output (in place is safe: a plane is staged before
// it is overwritten, and the other plane is not touched meanwhile).
template <typename R, int NPL, int BS>
__global__ __launch_bounds__(BS) void k_hxv_plane(const HxvArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int PF = BS == 1024 ? 4 : 2;
  constexpr int RSZ = sizeof(R);
  R *s_amp = reinterpret_cast<R *>(smem);
  R *s_col = reinterpret_cast<R *>(smem + buf_lds(a.namp_up, 0, 0, RSZ).col);
  const int dimup = a.dimup;
  const WgPos wp = wg_pos(a);
  if (!wp.valid) return;
  const int idw = wp.idw;
  const int64_t chain_off = (int64_t)wp.chain * a.vstride * NPL;
  LanczosDev *lz = a.lz ? a.lz + wp.chain : nullptr;
  bool done = false;
  R scale = (R)1;
  const R *vin = static_cast<const R *>(a.v0) + chain_off;
  if (lz) {
    const int k = lz->k;
    done = lz->done != 0;
    if (k & 1) vin = static_cast<const R *>(a.v1) + chain_off;
    if (!done) scale = (R)(1.0 / lz->nrm[k]);
  }
  const R *vcol = vin + (int64_t)idw * dimup * NPL;
  R *hcol = static_cast<R *>(a.hv) + chain_off + (int64_t)idw * dimup * NPL;
  double dot = 0.0;

  if (!done) {
    const cplx<R> *g_amp = static_cast<const cplx<R> *>(a.amp_up);
    for (int t = threadIdx.x; t < a.namp_up; t += BS) s_amp[t] = g_amp[t].x;
    if (threadIdx.x == 0) s_amp[a.namp_up] = (R)0;  // padding hops
    const double edw = a.skip_diag ? 0.0 : a.e_dw[idw];
    const uint32_t sdw = a.skip_diag ? 0u : a.map_dw[idw];
    const uint4 *ent4 = reinterpret_cast<const uint4 *>(a.ent4_up);
    const int last = a.up_nc - 1;
    // two planes of a column of at most kPlaneRows rows per lane: each
    // complex element is loaded once, its imaginary part held in registers
    // for the second plane (no second strided pass over the column)
    constexpr int kPlaneRows = 13;
    const bool hold = NPL == 2 && dimup <= kPlaneRows * BS;
    R xim[kPlaneRows];
    for (int pl = 0; pl < NPL; ++pl) {
      __syncthreads();  // the previous plane's gathers are done
      if (hold && pl == 0) {
        const cplx<R> *vc = reinterpret_cast<const cplx<R> *>(vcol);
#pragma unroll
        for (int q = 0; q < kPlaneRows; ++q) {
          const int f = (int)threadIdx.x + q * BS;
          if (f < dimup) {
            const cplx<R> x = vc[f];
            s_col[f] = x.x;
            xim[q] = x.y;
          }
        }
      } else if (hold) {
#pragma unroll
        for (int q = 0; q < kPlaneRows; ++q) {
          const int f = (int)threadIdx.x + q * BS;
          if (f < dimup) s_col[f] = xim[q];
        }
      } else {
        stage_copy<BS, R>(dimup, [&](int f) { return vcol[(int64_t)f * NPL + pl]; },
                          [&](int f, const R &x) { s_col[f] = x; });
      }
      __syncthreads();
      for (int base0 = 0; base0 < dimup; base0 += BS) {
        if (base0 + (int)(threadIdx.x & ~63u) >= dimup) break;
        const int base = base0 + (int)threadIdx.x;
        const int row = min(base, dimup - 1);
        const RowMeta m = a.meta_up[row];
        const int nu = base < dimup ? m.cnt : 0;
        const R x = s_col[row];
        const double ed = a.skip_diag ? 0.0 : m.e + edw + cross_energy(a.cross, m.s, sdw);
        R acc = (R)ed * x;
        int nc_w = (nu + 3) >> 2;
        for (int off = 32; off > 0; off >>= 1) nc_w = max(nc_w, __shfl_xor(nc_w, off, 64));
        nc_w = __builtin_amdgcn_readfirstlane(nc_w);
        if (nc_w > 0) {
          uint4 e4[PF];
#pragma unroll
          for (int p = 0; p < PF; ++p) e4[p] = ent4[(int64_t)min(p, last) * dimup + row];
          for (int c0 = 0; c0 < nc_w; c0 += PF) {
#pragma unroll
            for (int p = 0; p < PF; ++p) {
              if (c0 + p < nc_w) {
                const uint32_t ee[4] = {e4[p].x, e4[p].y, e4[p].z, e4[p].w};
                R y[4], h[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                  y[t] = lds_load_re<R>(ee[t] & kLdsAddrMask);
                  h[t] = lds_load_re<R>(ee[t] >> kLdsAddrBits);
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) acc = fma(h[t], y[t], acc);
              }
              e4[p] = ent4[(int64_t)min(c0 + p + PF, last) * dimup + row];
            }
          }
        }
        if (base < dimup) {
          const R out = acc * scale;
          hcol[(int64_t)row * NPL + pl] = out;
          if (lz) dot += (double)(x * scale) * (double)out;
        }
      }
    }
  }
  if (lz) block_partial_lds<BS>(dot, a.partials, wp.slot, reinterpret_cast<double *>(smem));
}

// Overlapped form of k_hxv_buf for short columns (DimUp <= NR * BS, each
// lane owning rows tid and tid + BS): the down-spin hops need nothing from
// LDS -- their hop list is read with uniform scalar loads, one neighbour
// column offset per hop in an SGPR -- so they run before the LDS barrier,
// while the lane's share of the input column (NR elements, held in
// registers) and the amplitude table are still in flight to LDS.  The
// up-spin gathers and the diagonal follow the barrier.  On a latency-bound
// launch (one round of workgroups) the column staging no longer adds to
// the critical path.
//
// SGN: the signed-column form (one real hopping amplitude t; the table is
// DevSpin::ent4s): the column is staged twice, scaled by +t and by -t, and
// each up-spin hop is one LDS load and two adds -- no amplitude gather.
template <typename R, bool REAL_AMP, int BS, int NR, bool SGN = false>
__global__ __launch_bounds__(BS) void k_hxv_ov(const HxvArgs a) {
  using C = cplx<R>;
  constexpr int ESZ = sizeof(C);
  constexpr int U = NR == 1 ? 8 : 4;
  // entry ring depth: one chunk for two complex128 rows, so that the
  // gathers fit 64 VGPRs (8 wavefronts per SIMD: the 3x4 sector's 924
  // workgroups of 512 lanes in one round)
  constexpr int PF = (NR == 2 && ESZ == 16 && !SGN) ? 1 : 2;
  static_assert(!SGN || REAL_AMP, "the signed-column form needs a real amplitude");
  extern __shared__ __align__(16) unsigned char smem[];
  const BufLds lds = buf_lds(a.namp_up, a.dimup, 0, ESZ);
  const SgnLds sl = sgn_lds(a.dimup, ESZ);
  C *s_amp_up = reinterpret_cast<C *>(smem);
  C *s_col = reinterpret_cast<C *>(smem + (SGN ? sl.colp : lds.col));
  C *s_coln = reinterpret_cast<C *>(smem + sl.coln);

  const int dimup = a.dimup;
  const WgPos wp = wg_pos(a);
  if (!wp.valid) return;
  const int idw = wp.idw;
  const int64_t chain_off = (int64_t)wp.chain * a.vstride;
  LanczosDev *lz = a.lz ? a.lz + wp.chain : nullptr;
  bool done = false;
  R scale = (R)1;
  const C *vin = static_cast<const C *>(a.v0) + chain_off;
  if (lz) {
    const int k = lz->k;
    done = lz->done != 0;
    if (k & 1) vin = static_cast<const C *>(a.v1) + chain_off;
    if (!done) scale = (R)(1.0 / lz->nrm[k]);
  }
  C *hv = static_cast<C *>(a.hv) + chain_off + (int64_t)idw * dimup;
  const C *vcol = vin + (int64_t)idw * dimup;
  double dot = 0.0;

  if (!done) {
    const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<C *>(vin), (short)0, (int)(uint32_t)(a.vstride * ESZ), 0x00020000);
    const C *g_amp_up = static_cast<const C *>(a.amp_up);
    const C *g_amp_dw = static_cast<const C *>(a.amp_dw);
    // the lane's rows, its share of the column and its row records, all
    // requested before the down-spin loads
    int row[NR], nu[NR];
    bool in[NR];
    C x[NR];
    RowMeta m[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const int r = (int)threadIdx.x + j * BS;
      in[j] = r < dimup;
      row[j] = min(r, dimup - 1);
      x[j] = vcol[row[j]];
      m[j] = a.meta_up[row[j]];
    }
    C amp = C{(R)0, (R)0};
    if (SGN) amp = g_amp_up[0];  // +t
    else if ((int)threadIdx.x < a.namp_up) amp = g_amp_up[threadIdx.x];
    R ar[NR], ai[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) ar[j] = ai[j] = (R)0;
    // down-spin hops: whole neighbour columns, offsets and amplitudes from
    // uniform loads of the column's hop list
    const int nd = a.skip_dw ? 0 : a.cnt_dw[idw];
    const uint32_t *entd = a.ent_dw + idw;
    // signed-column complex64: two stages of U/2 neighbour columns, one
    // stage's loads issued before the other is summed (as many loads in
    // flight, no gaps; in complex128 the second stage would not fit 64
    // VGPRs)
    constexpr bool kStaged = SGN && ESZ == 8;
    if constexpr (kStaged) {
      constexpr int UH = U / 2;
      C ya[UH][NR], yb[UH][NR], ha[UH], hb[UH];
      auto load_stage = [&](C(&y)[UH][NR], C(&h)[UH], int q0) {
#pragma unroll
        for (int u = 0; u < UH; ++u) {
          const uint32_t e = entd[(int64_t)min(q0 + u, nd - 1) * a.dimdw];
          const uint32_t so = (e & kColMask) * (uint32_t)dimup * (uint32_t)ESZ;
          h[u] = g_amp_dw[e >> kColBits];
#pragma unroll
          for (int j = 0; j < NR; ++j) {
            if constexpr (ESZ == 16)
              y[u][j] = __builtin_bit_cast(
                  C, __builtin_amdgcn_raw_buffer_load_b128(rv, row[j] * ESZ, so, 0));
            else
              y[u][j] = __builtin_bit_cast(
                  C, __builtin_amdgcn_raw_buffer_load_b64(rv, row[j] * ESZ, so, 0));
          }
        }
      };
      auto sum_stage = [&](const C(&y)[UH][NR], const C(&h)[UH], int q0) {
#pragma unroll
        for (int u = 0; u < UH; ++u)
          if (q0 + u < nd) {
#pragma unroll
            for (int j = 0; j < NR; ++j) cmac<R, REAL_AMP>(ar[j], ai[j], h[u], y[u][j]);
          }
      };
      if (nd > 0) load_stage(ya, ha, 0);
      for (int q0 = 0; q0 < nd; q0 += 2 * UH) {
        if (q0 + UH < nd) load_stage(yb, hb, q0 + UH);
        sum_stage(ya, ha, q0);
        if (q0 + 2 * UH < nd) load_stage(ya, ha, q0 + 2 * UH);
        if (q0 + UH < nd) sum_stage(yb, hb, q0 + UH);
      }
    }
    for (int q0 = 0; q0 < (kStaged ? 0 : nd); q0 += U) {
      C y[U][NR], h[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t e = entd[(int64_t)min(q0 + u, nd - 1) * a.dimdw];
        const uint32_t so = (e & kColMask) * (uint32_t)dimup * (uint32_t)ESZ;
        h[u] = g_amp_dw[e >> kColBits];
#pragma unroll
        for (int j = 0; j < NR; ++j) {
          if constexpr (ESZ == 16)
            y[u][j] = __builtin_bit_cast(
                C, __builtin_amdgcn_raw_buffer_load_b128(rv, row[j] * ESZ, so, 0));
          else
            y[u][j] = __builtin_bit_cast(
                C, __builtin_amdgcn_raw_buffer_load_b64(rv, row[j] * ESZ, so, 0));
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (q0 + u < nd) {
#pragma unroll
          for (int j = 0; j < NR; ++j) cmac<R, REAL_AMP>(ar[j], ai[j], h[u], y[u][j]);
        }
    }
    // diagonal, then the column and the amplitudes into LDS (the row
    // records and the column elements die here: fewer VGPRs in the gathers)
    const double edw = a.skip_diag ? 0.0 : a.e_dw[idw];
    const uint32_t sdw = a.skip_diag ? 0u : a.map_dw[idw];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      nu[j] = in[j] ? m[j].cnt : 0;
      const double ed = a.skip_diag ? 0.0 : m[j].e + edw + cross_energy(a.cross, m[j].s, sdw);
      ar[j] = fma((R)ed, x[j].x, ar[j]);
      ai[j] = fma((R)ed, x[j].y, ai[j]);
      if constexpr (SGN) {
        if (in[j]) {
          s_col[row[j]] = C{amp.x * x[j].x, amp.x * x[j].y};
          s_coln[row[j]] = C{-amp.x * x[j].x, -amp.x * x[j].y};
        }
      } else {
        if (in[j]) s_col[row[j]] = x[j];
      }
    }
    if constexpr (SGN) {
      if (threadIdx.x == 0) s_amp_up[0] = C{(R)0, (R)0};  // padding hops
    } else {
      if ((int)threadIdx.x < a.namp_up) s_amp_up[threadIdx.x] = amp;
      for (int t = (int)threadIdx.x + BS; t < a.namp_up; t += BS) s_amp_up[t] = g_amp_up[t];
      if (threadIdx.x == 0) s_amp_up[a.namp_up] = C{(R)0, (R)0};  // padding hops
    }
    __syncthreads();
    // up-spin hops: LDS gathers, entries through a ring of PF chunks
    int nc_w = 0;
#pragma unroll
    for (int j = 0; j < NR; ++j) nc_w = max(nc_w, (nu[j] + 3) >> 2);
    for (int off = 32; off > 0; off >>= 1) nc_w = max(nc_w, __shfl_xor(nc_w, off, 64));
    nc_w = __builtin_amdgcn_readfirstlane(nc_w);
    const int last = a.up_nc - 1;
    // signed-column entries: four 16-bit element indices per 8 bytes
    using EntT = std::conditional_t<SGN, uint2, uint4>;
    const EntT *ent4 = reinterpret_cast<const EntT *>(a.ent4_up);
    if (nc_w > 0) {
      EntT e4[PF][NR];
#pragma unroll
      for (int p = 0; p < PF; ++p)
#pragma unroll
        for (int j = 0; j < NR; ++j) e4[p][j] = ent4[(int64_t)min(p, last) * dimup + row[j]];
      for (int c0 = 0; c0 < nc_w; c0 += PF) {
#pragma unroll
        for (int p = 0; p < PF; ++p) {
          if (c0 + p < nc_w) {
#pragma unroll
            for (int j = 0; j < NR; ++j) {
              if constexpr (SGN) {
                const uint32_t ee[4] = {e4[p][j].x & 0xffffu, e4[p][j].x >> 16,
                                        e4[p][j].y & 0xffffu, e4[p][j].y >> 16};
                C y[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) y[t] = lds_load<R>(ee[t] * (uint32_t)ESZ);
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                  ar[j] += y[t].x;
                  ai[j] += y[t].y;
                }
              } else {
                const uint32_t ee[4] = {e4[p][j].x, e4[p][j].y, e4[p][j].z, e4[p][j].w};
                C y[4], h[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                  y[t] = lds_load<R>(ee[t] & kLdsAddrMask);
                  if constexpr (REAL_AMP) {
                    h[t].x = lds_load_re<R>(ee[t] >> kLdsAddrBits);
                    h[t].y = (R)0;
                  } else {
                    h[t] = lds_load<R>(ee[t] >> kLdsAddrBits);
                  }
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) cmac<R, REAL_AMP>(ar[j], ai[j], h[t], y[t]);
              }
            }
          }
          const int cn = min(c0 + p + PF, last);
#pragma unroll
          for (int j = 0; j < NR; ++j) e4[p][j] = ent4[(int64_t)cn * dimup + row[j]];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      if (in[j]) {
        C out;
        out.x = ar[j] * scale;
        out.y = ai[j] * scale;
        if (a.nt_store)
          __builtin_nontemporal_store(__builtin_bit_cast(typename Vec2<R>::T, out),
                                      reinterpret_cast<typename Vec2<R>::T *>(hv) + row[j]);
        else
          hv[row[j]] = out;
        if (lz) {
          const C xj = SGN ? vcol[row[j]] : s_col[row[j]];
          dot += (double)(xj.x * scale) * (double)out.x + (double)(xj.y * scale) * (double)out.y;
        }
      }
    }
  }
  if (lz) block_partial_lds<BS>(dot, a.partials, wp.slot, reinterpret_cast<double *>(smem));
}

// Real-vector form of k_hxv_ov (real amplitudes, vectors of dim reals) for
// short columns: each lane owns NR slots of P adjacent rows (slot s holds
// rows s*P .. s*P+P-1; P = 2 needs an even DimUp), DimUp <= BS * NR * P.
// The down-spin hops are buffer loads of whole neighbour columns (one P-row
// vector load per slot and hop) with the hop list read by uniform loads,
// issued before the LDS barrier; the column and the amplitudes go to LDS
// meanwhile, and the up-spin gathers use the one-plane hop table
// (DevSpin::ent4p).  a.vstride counts reals between chains.
template <typename R, int BS, int NR, int P>
__global__ __launch_bounds__(BS) void k_hxv_ov_real(const HxvArgs a) {
  constexpr int ESZ = sizeof(R);
  constexpr int NRW = NR * P;  // rows per lane
  constexpr int U = NRW <= 2 ? 8 : 4;
  constexpr int PF = 2;
  using V2 = typename Vec2<R>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  const BufLds lds = buf_lds(a.namp_up, a.dimup, 0, ESZ);
  R *s_amp_up = reinterpret_cast<R *>(smem);
  R *s_col = reinterpret_cast<R *>(smem + lds.col);

  const int dimup = a.dimup;
  const WgPos wp = wg_pos(a);
  if (!wp.valid) return;
  const int idw = wp.idw;
  const int64_t chain_off = (int64_t)wp.chain * a.vstride;
  LanczosDev *lz = a.lz ? a.lz + wp.chain : nullptr;
  bool done = false;
  R scale = (R)1;
  const R *vin = static_cast<const R *>(a.v0) + chain_off;
  if (lz) {
    const int k = lz->k;
    done = lz->done != 0;
    if (k & 1) vin = static_cast<const R *>(a.v1) + chain_off;
    if (!done) scale = (R)(1.0 / lz->nrm[k]);
  }
  R *hv = static_cast<R *>(a.hv) + chain_off + (int64_t)idw * dimup;
  const R *vcol = vin + (int64_t)idw * dimup;
  double dot = 0.0;

  if (!done) {
    const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<R *>(vin), (short)0, (int)(uint32_t)(a.vstride * ESZ), 0x00020000);
    const cplx<R> *g_amp_up = static_cast<const cplx<R> *>(a.amp_up);
    const cplx<R> *g_amp_dw = static_cast<const cplx<R> *>(a.amp_dw);
    const int nslot = dimup / P;
    int slot[NR], row[NRW], nu[NRW];
    bool in[NR];
    R x[NRW];
    RowMeta m[NRW];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const int sl = (int)threadIdx.x + j * BS;
      in[j] = sl < nslot;
      slot[j] = min(sl, nslot - 1);
      if constexpr (P == 2) {
        const V2 xv = reinterpret_cast<const V2 *>(vcol)[slot[j]];
        x[2 * j] = xv.x;
        x[2 * j + 1] = xv.y;
      } else {
        x[j] = vcol[slot[j]];
      }
#pragma unroll
      for (int p = 0; p < P; ++p) {
        row[j * P + p] = slot[j] * P + p;
        m[j * P + p] = a.meta_up[slot[j] * P + p];
      }
    }
    R amp = (R)0;
    if ((int)threadIdx.x < a.namp_up) amp = g_amp_up[threadIdx.x].x;
    R ar[NRW];
#pragma unroll
    for (int j = 0; j < NRW; ++j) ar[j] = (R)0;
    // down-spin hops: whole neighbour columns, offsets and amplitudes from
    // uniform loads of the column's hop list
    const int nd = a.skip_dw ? 0 : a.cnt_dw[idw];
    const uint32_t *entd = a.ent_dw + idw;
    for (int q0 = 0; q0 < nd; q0 += U) {
      R y[U][NRW], h[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t e = entd[(int64_t)min(q0 + u, nd - 1) * a.dimdw];
        const uint32_t so = (e & kColMask) * (uint32_t)dimup * (uint32_t)ESZ;
        h[u] = g_amp_dw[e >> kColBits].x;
#pragma unroll
        for (int j = 0; j < NR; ++j) {
          const int vo = slot[j] * P * ESZ;
          if constexpr (P == 2 && ESZ == 8) {
            const V2 yv = __builtin_bit_cast(V2, __builtin_amdgcn_raw_buffer_load_b128(rv, vo, so, 0));
            y[u][2 * j] = yv.x;
            y[u][2 * j + 1] = yv.y;
          } else if constexpr (P == 2) {
            const V2 yv = __builtin_bit_cast(V2, __builtin_amdgcn_raw_buffer_load_b64(rv, vo, so, 0));
            y[u][2 * j] = yv.x;
            y[u][2 * j + 1] = yv.y;
          } else if constexpr (ESZ == 8) {
            y[u][j] = __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(rv, vo, so, 0));
          } else {
            y[u][j] = __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(rv, vo, so, 0));
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (q0 + u < nd) {
#pragma unroll
          for (int j = 0; j < NRW; ++j) ar[j] = fma(h[u], y[u][j], ar[j]);
        }
    }
    // diagonal, then the column and the amplitudes into LDS
    const double edw = a.skip_diag ? 0.0 : a.e_dw[idw];
    const uint32_t sdw = a.skip_diag ? 0u : a.map_dw[idw];
#pragma unroll
    for (int j = 0; j < NRW; ++j) {
      nu[j] = in[j / P] ? m[j].cnt : 0;
      const double ed = a.skip_diag ? 0.0 : m[j].e + edw + cross_energy(a.cross, m[j].s, sdw);
      ar[j] = fma((R)ed, x[j], ar[j]);
      if (in[j / P]) s_col[row[j]] = x[j];
    }
    if ((int)threadIdx.x < a.namp_up) s_amp_up[threadIdx.x] = amp;
    for (int t = (int)threadIdx.x + BS; t < a.namp_up; t += BS) s_amp_up[t] = g_amp_up[t].x;
    if (threadIdx.x == 0) s_amp_up[a.namp_up] = (R)0;  // padding hops
    __syncthreads();
    // up-spin hops: LDS gathers, entries through a ring of PF chunks
    int nc_w = 0;
#pragma unroll
    for (int j = 0; j < NRW; ++j) nc_w = max(nc_w, (nu[j] + 3) >> 2);
    for (int off = 32; off > 0; off >>= 1) nc_w = max(nc_w, __shfl_xor(nc_w, off, 64));
    nc_w = __builtin_amdgcn_readfirstlane(nc_w);
    const int last = a.up_nc - 1;
    const uint4 *ent4 = reinterpret_cast<const uint4 *>(a.ent4_up);
    if (nc_w > 0) {
      uint4 e4[PF][NRW];
#pragma unroll
      for (int p = 0; p < PF; ++p)
#pragma unroll
        for (int j = 0; j < NRW; ++j) e4[p][j] = ent4[(int64_t)min(p, last) * dimup + row[j]];
      for (int c0 = 0; c0 < nc_w; c0 += PF) {
#pragma unroll
        for (int p = 0; p < PF; ++p) {
          if (c0 + p < nc_w) {
#pragma unroll
            for (int j = 0; j < NRW; ++j) {
              const uint32_t ee[4] = {e4[p][j].x, e4[p][j].y, e4[p][j].z, e4[p][j].w};
              R y[4], h[4];
#pragma unroll
              for (int t = 0; t < 4; ++t) {
                y[t] = lds_load_re<R>(ee[t] & kLdsAddrMask);
                h[t] = lds_load_re<R>(ee[t] >> kLdsAddrBits);
              }
#pragma unroll
              for (int t = 0; t < 4; ++t) ar[j] = fma(h[t], y[t], ar[j]);
            }
          }
          const int cn = min(c0 + p + PF, last);
#pragma unroll
          for (int j = 0; j < NRW; ++j) e4[p][j] = ent4[(int64_t)cn * dimup + row[j]];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      if (in[j]) {
        if constexpr (P == 2) {
          V2 out;
          out.x = ar[2 * j] * scale;
          out.y = ar[2 * j + 1] * scale;
          if (a.nt_store) __builtin_nontemporal_store(out, reinterpret_cast<V2 *>(hv) + slot[j]);
          else reinterpret_cast<V2 *>(hv)[slot[j]] = out;
          if (lz) {
            const V2 xv = *reinterpret_cast<const V2 *>(s_col + 2 * slot[j]);
            dot += (double)(xv.x * scale) * (double)out.x + (double)(xv.y * scale) * (double)out.y;
          }
        } else {
          const R out = ar[j] * scale;
          if (a.nt_store) __builtin_nontemporal_store(out, hv + slot[j]);
          else hv[slot[j]] = out;
          if (lz) dot += (double)(s_col[slot[j]] * scale) * (double)out;
        }
      }
    }
  }
  if (lz) block_partial_lds<BS>(dot, a.partials, wp.slot, reinterpret_cast<double *>(smem));
}

// Real-vector form of k_hxv_buf (real amplitudes, vectors of dim reals): the
// same schedule on elements of sizeof(R) -- the column staged in LDS, the
// up-spin hops gathered there through the one-plane hop table
// (DevSpin::ent4p), the down-spin hops as buffer loads of whole neighbour
// columns with the column offset in an SGPR.  Half the bytes and half the
// LDS traffic of the complex form; the excitation chains of real
// Hamiltonians run on it.  a.vstride counts reals between chains.
template <typename R, int BS, int KR>
__global__ __launch_bounds__(BS) void k_hxv_buf_real(const HxvArgs a) {
  constexpr int ESZ = sizeof(R);
  constexpr int U = KR >= 4 ? 4 : 8;
  constexpr int UH = U / 2;
  constexpr int PF = (BS == 1024 && KR == 1) ? 4 : 2;
  extern __shared__ __align__(16) unsigned char smem[];
  const int maxnd = a.dw_maxnnz;
  const BufLds lds = buf_lds(a.namp_up, a.dimup, maxnd, ESZ);
  R *s_amp_up = reinterpret_cast<R *>(smem);
  R *s_col = reinterpret_cast<R *>(smem + lds.col);
  R *s_dw_amp = reinterpret_cast<R *>(smem + lds.dw_amp);
  uint32_t *s_dw_soff = reinterpret_cast<uint32_t *>(smem + lds.dw_soff);

  const int dimup = a.dimup;
  const WgPos wp = wg_pos(a);
  if (!wp.valid) return;
  const int idw = wp.idw;
  const int64_t chain_off = (int64_t)wp.chain * a.vstride;
  LanczosDev *lz = a.lz ? a.lz + wp.chain : nullptr;
  bool done = false;
  R scale = (R)1;
  const R *vin = static_cast<const R *>(a.v0) + chain_off;
  if (lz) {
    const int k = lz->k;
    done = lz->done != 0;
    if (k & 1) vin = static_cast<const R *>(a.v1) + chain_off;
    if (!done) scale = (R)(1.0 / lz->nrm[k]);
  }
  R *hv = static_cast<R *>(a.hv) + chain_off + (int64_t)idw * dimup;
  const R *vcol = vin + (int64_t)idw * dimup;
  double dot = 0.0;

  if (!done) {
    const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<R *>(vin), (short)0, (int)(uint32_t)(a.vstride * ESZ), 0x00020000);
    const cplx<R> *g_amp_up = static_cast<const cplx<R> *>(a.amp_up);
    const cplx<R> *g_amp_dw = static_cast<const cplx<R> *>(a.amp_dw);
    const int nd = a.skip_dw ? 0 : a.cnt_dw[idw];
    for (int t = threadIdx.x; t < a.namp_up; t += BS) s_amp_up[t] = g_amp_up[t].x;
    if (threadIdx.x == 0) s_amp_up[a.namp_up] = (R)0;  // padding hops
    for (int q = threadIdx.x; q < nd; q += BS) {
      const uint32_t e = a.ent_dw[(int64_t)q * a.dimdw + idw];
      s_dw_soff[q] = (e & kColMask) * (uint32_t)dimup * (uint32_t)ESZ;
      s_dw_amp[q] = g_amp_dw[e >> kColBits].x;
    }
    stage_copy<BS, R>(dimup, [&](int f) { return vcol[f]; },
                      [&](int f, const R &x) { s_col[f] = x; });
    __syncthreads();
    const double edw = a.skip_diag ? 0.0 : a.e_dw[idw];
    const uint32_t sdw = a.skip_diag ? 0u : a.map_dw[idw];
    for (int base0 = 0; base0 < dimup; base0 += BS * KR) {
      if (base0 + (int)(threadIdx.x & ~63u) >= dimup) break;
      const int base = base0 + (int)threadIdx.x;
      int row[KR], nu[KR];
      R ar[KR], x[KR];
#pragma unroll
      for (int j = 0; j < KR; ++j) {
        row[j] = min(base + j * BS, dimup - 1);
        const RowMeta m = a.meta_up[row[j]];
        nu[j] = (base + j * BS < dimup) ? m.cnt : 0;
        x[j] = s_col[row[j]];
        const double ed = a.skip_diag ? 0.0 : m.e + edw + cross_energy(a.cross, m.s, sdw);
        ar[j] = (R)ed * x[j];
      }
      // up-spin hops: LDS gathers, entries through a ring of PF chunks
      int nc_w = 0;
#pragma unroll
      for (int j = 0; j < KR; ++j) nc_w = max(nc_w, (nu[j] + 3) >> 2);
      for (int off = 32; off > 0; off >>= 1) nc_w = max(nc_w, __shfl_xor(nc_w, off, 64));
      nc_w = __builtin_amdgcn_readfirstlane(nc_w);
      const int last = a.up_nc - 1;
      const uint4 *ent4 = reinterpret_cast<const uint4 *>(a.ent4_up);
      if (nc_w > 0) {
        uint4 e4[PF][KR];
#pragma unroll
        for (int p = 0; p < PF; ++p)
#pragma unroll
          for (int j = 0; j < KR; ++j) e4[p][j] = ent4[(int64_t)min(p, last) * dimup + row[j]];
        for (int c0 = 0; c0 < nc_w; c0 += PF) {
#pragma unroll
          for (int p = 0; p < PF; ++p) {
            if (c0 + p < nc_w) {
#pragma unroll
              for (int j = 0; j < KR; ++j) {
                const uint32_t ee[4] = {e4[p][j].x, e4[p][j].y, e4[p][j].z, e4[p][j].w};
                R y[4], h[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                  y[t] = lds_load_re<R>(ee[t] & kLdsAddrMask);
                  h[t] = lds_load_re<R>(ee[t] >> kLdsAddrBits);
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) ar[j] = fma(h[t], y[t], ar[j]);
              }
            }
            const int cn = min(c0 + p + PF, last);
#pragma unroll
            for (int j = 0; j < KR; ++j) e4[p][j] = ent4[(int64_t)cn * dimup + row[j]];
          }
        }
      }
      // down-spin hops: two stages of U/2 neighbour columns, one stage's
      // buffer loads issued before the other is summed
      R ya[UH][KR], yb[UH][KR];
      auto load_stage = [&](R(&y)[UH][KR], int q0) {
#pragma unroll
        for (int u = 0; u < UH; ++u) {
          const uint32_t so = __builtin_amdgcn_readfirstlane(s_dw_soff[min(q0 + u, nd - 1)]);
#pragma unroll
          for (int j = 0; j < KR; ++j) {
            if constexpr (ESZ == 8)
              y[u][j] = __builtin_bit_cast(
                  R, __builtin_amdgcn_raw_buffer_load_b64(rv, row[j] * ESZ, so, 0));
            else
              y[u][j] = __builtin_bit_cast(
                  R, __builtin_amdgcn_raw_buffer_load_b32(rv, row[j] * ESZ, so, 0));
          }
        }
      };
      auto sum_stage = [&](const R(&y)[UH][KR], int q0) {
#pragma unroll
        for (int u = 0; u < UH; ++u)
          if (q0 + u < nd) {
            const R h = s_dw_amp[q0 + u];
#pragma unroll
            for (int j = 0; j < KR; ++j) ar[j] = fma(h, y[u][j], ar[j]);
          }
      };
      if (nd > 0) load_stage(ya, 0);
      for (int q0 = 0; q0 < nd; q0 += 2 * UH) {
        if (q0 + UH < nd) load_stage(yb, q0 + UH);
        sum_stage(ya, q0);
        if (q0 + 2 * UH < nd) load_stage(ya, q0 + 2 * UH);
        if (q0 + UH < nd) sum_stage(yb, q0 + UH);
      }
#pragma unroll
      for (int j = 0; j < KR; ++j) {
        if (base + j * BS < dimup) {
          const R out = ar[j] * scale;
          hv[row[j]] = out;
          if (lz) dot += (double)(x[j] * scale) * (double)out;
        }
      }
    }
  }
  if (lz) block_partial_lds<BS>(dot, a.partials, wp.slot, reinterpret_cast<double *>(smem));
}

// k_hxv_buf_real with two adjacent rows per lane (rows 2t and 2t + 1 of each
// block of 2*BS rows): each down-spin hop is one 2-element load per lane
// (16 B FP64, 8 B FP32) instead of two, and the results are stored in
// pairs.  Needs an even DimUp (every column then starts on a pair
// boundary).
template <typename R, int BS>
__global__ __launch_bounds__(BS) void k_hxv_buf_real2(const HxvArgs a) {
  constexpr int ESZ = sizeof(R);
  constexpr int U = 4;
  constexpr int PF = 2;
  using R2 = typename Vec2<R>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  const int maxnd = a.dw_maxnnz;
  const BufLds lds = buf_lds(a.namp_up, a.dimup, maxnd, ESZ);
  R *s_amp_up = reinterpret_cast<R *>(smem);
  R *s_col = reinterpret_cast<R *>(smem + lds.col);
  R *s_dw_amp = reinterpret_cast<R *>(smem + lds.dw_amp);
  uint32_t *s_dw_soff = reinterpret_cast<uint32_t *>(smem + lds.dw_soff);

  const int dimup = a.dimup;
  const WgPos wp = wg_pos(a);
  if (!wp.valid) return;
  const int idw = wp.idw;
  const int64_t chain_off = (int64_t)wp.chain * a.vstride;
  LanczosDev *lz = a.lz ? a.lz + wp.chain : nullptr;
  bool done = false;
  R scale = (R)1;
  const R *vin = static_cast<const R *>(a.v0) + chain_off;
  if (lz) {
    const int k = lz->k;
    done = lz->done != 0;
    if (k & 1) vin = static_cast<const R *>(a.v1) + chain_off;
    if (!done) scale = (R)(1.0 / lz->nrm[k]);
  }
  R *hv = static_cast<R *>(a.hv) + chain_off + (int64_t)idw * dimup;
  const R *vcol = vin + (int64_t)idw * dimup;
  double dot = 0.0;

  if (!done) {
    const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<R *>(vin), (short)0, (int)(uint32_t)(a.vstride * ESZ), 0x00020000);
    const cplx<R> *g_amp_up = static_cast<const cplx<R> *>(a.amp_up);
    const cplx<R> *g_amp_dw = static_cast<const cplx<R> *>(a.amp_dw);
    const int nd = a.skip_dw ? 0 : a.cnt_dw[idw];
    for (int t = threadIdx.x; t < a.namp_up; t += BS) s_amp_up[t] = g_amp_up[t].x;
    if (threadIdx.x == 0) s_amp_up[a.namp_up] = (R)0;  // padding hops
    for (int q = threadIdx.x; q < nd; q += BS) {
      const uint32_t e = a.ent_dw[(int64_t)q * a.dimdw + idw];
      s_dw_soff[q] = (e & kColMask) * (uint32_t)dimup * (uint32_t)ESZ;
      s_dw_amp[q] = g_amp_dw[e >> kColBits].x;
    }
    stage_copy<BS, R2>(dimup / 2, [&](int f) { return reinterpret_cast<const R2 *>(vcol)[f]; },
                       [&](int f, const R2 &x) { reinterpret_cast<R2 *>(s_col)[f] = x; });
    __syncthreads();
    const double edw = a.skip_diag ? 0.0 : a.e_dw[idw];
    const uint32_t sdw = a.skip_diag ? 0u : a.map_dw[idw];
    for (int base0 = 0; base0 < dimup; base0 += 2 * BS) {
      if (base0 + 2 * (int)(threadIdx.x & ~63u) >= dimup) break;
      const int pr = base0 / 2 + (int)threadIdx.x;  // pair index
      const bool in = 2 * pr < dimup;
      const int r0 = min(2 * pr, dimup - 2);
      int row[2], nu[2];
      R ar[2], x[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        row[j] = r0 + j;
        const RowMeta m = a.meta_up[row[j]];
        nu[j] = in ? m.cnt : 0;
        x[j] = s_col[row[j]];
        const double ed = a.skip_diag ? 0.0 : m.e + edw + cross_energy(a.cross, m.s, sdw);
        ar[j] = (R)ed * x[j];
      }
      int nc_w = max((nu[0] + 3) >> 2, (nu[1] + 3) >> 2);
      for (int off = 32; off > 0; off >>= 1) nc_w = max(nc_w, __shfl_xor(nc_w, off, 64));
      nc_w = __builtin_amdgcn_readfirstlane(nc_w);
      const int last = a.up_nc - 1;
      const uint4 *ent4 = reinterpret_cast<const uint4 *>(a.ent4_up);
      if (nc_w > 0) {
        uint4 e4[PF][2];
#pragma unroll
        for (int p = 0; p < PF; ++p)
#pragma unroll
          for (int j = 0; j < 2; ++j) e4[p][j] = ent4[(int64_t)min(p, last) * dimup + row[j]];
        for (int c0 = 0; c0 < nc_w; c0 += PF) {
#pragma unroll
          for (int p = 0; p < PF; ++p) {
            if (c0 + p < nc_w) {
#pragma unroll
              for (int j = 0; j < 2; ++j) {
                const uint32_t ee[4] = {e4[p][j].x, e4[p][j].y, e4[p][j].z, e4[p][j].w};
                R y[4], h[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                  y[t] = lds_load_re<R>(ee[t] & kLdsAddrMask);
                  h[t] = lds_load_re<R>(ee[t] >> kLdsAddrBits);
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) ar[j] = fma(h[t], y[t], ar[j]);
              }
            }
            const int cn = min(c0 + p + PF, last);
#pragma unroll
            for (int j = 0; j < 2; ++j) e4[p][j] = ent4[(int64_t)cn * dimup + row[j]];
          }
        }
      }
      // down-spin hops: one 2-element load per hop for the lane's row pair
      for (int q0 = 0; q0 < nd; q0 += U) {
        R2 y[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint32_t so = __builtin_amdgcn_readfirstlane(s_dw_soff[min(q0 + u, nd - 1)]);
          if constexpr (ESZ == 8)
            y[u] = __builtin_bit_cast(
                R2, __builtin_amdgcn_raw_buffer_load_b128(rv, r0 * ESZ, so, 0));
          else
            y[u] = __builtin_bit_cast(
                R2, __builtin_amdgcn_raw_buffer_load_b64(rv, r0 * ESZ, so, 0));
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (q0 + u < nd) {
            const R h = s_dw_amp[q0 + u];
            ar[0] = fma(h, y[u].x, ar[0]);
            ar[1] = fma(h, y[u].y, ar[1]);
          }
      }
      if (in) {
        R2 out;
        out.x = ar[0] * scale;
        out.y = ar[1] * scale;
        reinterpret_cast<R2 *>(hv)[r0 / 2] = out;
        if (lz)
          dot += (double)(x[0] * scale) * (double)out.x + (double)(x[1] * scale) * (double)out.y;
      }
    }
  }
  if (lz) block_partial_lds<BS>(dot, a.partials, wp.slot, reinterpret_cast<double *>(smem));
}

// The real-vector gather pass with two columns per workgroup (columns 2p
// and 2p + 1 of the pair p = workgroup), for real vectors whose two columns
// fit the LDS together (FP32: 2 x 51.5 KB on the 12,870-row bath sector):
// every hop entry and amplitude read from the table serves both columns,
// halving the entry stream from L2.  The two columns are staged interleaved
// (element f of both at o_col + 2 f RSZ), so one 2 RSZ-byte LDS read per hop
// fetches both: the LDS serves 32 lanes per cycle whatever the width up to
// 8 bytes, so a 4-byte read per column would take twice the cycles.  The
// hop table is DevSpin::ent4q, the one-plane table's hops addressed for the
// pair (a mask and a shift per hop).  LDS: the amplitudes and a zero at 0,
// then the column pair.  One partial of <v|Hv> per pair.
//
// CPLX: the same pass on a complex vector with real amplitudes, one column
// per workgroup: the complex column is already the interleaved pair (real
// part at o_col + 2 f RSZ, imaginary part after it), so it is staged as it
// is and each hop's one 2 RSZ-byte read feeds both parts, with the sums of
// the whole-column complex pass (k_hxv_buf, real amplitudes) in the same
// order.  One partial of Re <v|Hv> per column.
template <typename R, int BS, bool CPLX>
__global__ __launch_bounds__(BS) void k_hxv_plane2(const HxvArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int PF = BS == 1024 ? 4 : 2;
  constexpr int RSZ = sizeof(R);
  const int dimup = a.dimup;
  const uint32_t o_col = (uint32_t)buf_lds(a.namp_up, 0, 0, RSZ).col;
  R *s_amp = reinterpret_cast<R *>(smem);
  cplx<R> *s_pair = reinterpret_cast<cplx<R> *>(smem + o_col);
  // CPLX: complex elements of 2 reals; the column's "pair" is its two parts
  constexpr int NR = CPLX ? 2 : 1;
  const int npair = CPLX ? a.dimdw : (a.dimdw + 1) / 2;
  const int pair = xcd_column(blockIdx.x, npair);
  const int idw0 = CPLX ? pair : 2 * pair;
  const bool two = !CPLX && idw0 + 1 < a.dimdw;
  const int idw1 = two ? idw0 + 1 : idw0;
  const int64_t chain_off = (int64_t)blockIdx.y * a.vstride * NR;
  const int64_t slot = a.pbase + (int64_t)blockIdx.y * a.pstride + blockIdx.x;
  LanczosDev *lz = a.lz ? a.lz + blockIdx.y : nullptr;
  bool done = false;
  R scale = (R)1;
  const R *vin = static_cast<const R *>(a.v0) + chain_off;
  if (lz) {
    const int k = lz->k;
    done = lz->done != 0;
    if (k & 1) vin = static_cast<const R *>(a.v1) + chain_off;
    if (!done) scale = (R)(1.0 / lz->nrm[k]);
  }
  const R *vcol0 = vin + (int64_t)idw0 * dimup * NR;
  const R *vcol1 = vin + (int64_t)idw1 * dimup * NR;
  R *hcol0 = static_cast<R *>(a.hv) + chain_off + (int64_t)idw0 * dimup * NR;
  R *hcol1 = static_cast<R *>(a.hv) + chain_off + (int64_t)idw1 * dimup * NR;
  double dot = 0.0;

  if (!done) {
    const cplx<R> *g_amp = static_cast<const cplx<R> *>(a.amp_up);
    for (int t = threadIdx.x; t < a.namp_up; t += BS) s_amp[t] = g_amp[t].x;
    if (threadIdx.x == 0) s_amp[a.namp_up] = (R)0;  // padding hops
    if constexpr (CPLX) {
      const cplx<R> *vc = reinterpret_cast<const cplx<R> *>(vcol0);
      stage_copy<BS, cplx<R>>(dimup, [&](int f) { return vc[f]; },
                              [&](int f, const cplx<R> &y) { s_pair[f] = y; });
    } else {
      stage_copy<BS, cplx<R>>(
          dimup,
          [&](int f) {
            cplx<R> y;
            y.x = vcol0[f];
            y.y = vcol1[f];
            return y;
          },
          [&](int f, const cplx<R> &y) { s_pair[f] = y; });
    }
    __syncthreads();
    const double edw0 = a.skip_diag ? 0.0 : a.e_dw[idw0];
    const double edw1 = a.skip_diag ? 0.0 : a.e_dw[idw1];
    const uint32_t sdw0 = a.skip_diag ? 0u : a.map_dw[idw0];
    const uint32_t sdw1 = a.skip_diag ? 0u : a.map_dw[idw1];
    const uint4 *ent4 = reinterpret_cast<const uint4 *>(a.ent4_up);
    const int last = a.up_nc - 1;
    for (int base0 = 0; base0 < dimup; base0 += BS) {
      if (base0 + (int)(threadIdx.x & ~63u) >= dimup) break;
      const int base = base0 + (int)threadIdx.x;
      const int row = min(base, dimup - 1);
      const RowMeta m = a.meta_up[row];
      const int nu = base < dimup ? m.cnt : 0;
      const cplx<R> xp = s_pair[row];
      const R x0 = xp.x, x1 = xp.y;
      R acc0 = (R)0, acc1 = (R)0;
      if (!a.skip_diag) {
        const R e0 = (R)(m.e + edw0 + cross_energy(a.cross, m.s, sdw0));
        acc0 = e0 * x0;
        acc1 = (CPLX ? e0 : (R)(m.e + edw1 + cross_energy(a.cross, m.s, sdw1))) * x1;
      }
      int nc_w = (nu + 3) >> 2;
      for (int off = 32; off > 0; off >>= 1) nc_w = max(nc_w, __shfl_xor(nc_w, off, 64));
      nc_w = __builtin_amdgcn_readfirstlane(nc_w);
      if (nc_w > 0) {
        uint4 e4[PF];
#pragma unroll
        for (int p = 0; p < PF; ++p) e4[p] = ent4[(int64_t)min(p, last) * dimup + row];
        for (int c0 = 0; c0 < nc_w; c0 += PF) {
#pragma unroll
          for (int p = 0; p < PF; ++p) {
            if (c0 + p < nc_w) {
              const uint32_t ee[4] = {e4[p].x, e4[p].y, e4[p].z, e4[p].w};
              cplx<R> y[4];
              R h[4];
#pragma unroll
              for (int t = 0; t < 4; ++t) {
                y[t] = lds_load<R>(ee[t] & kLdsAddrMask);
                h[t] = lds_load_re<R>(ee[t] >> kLdsAddrBits);
              }
#pragma unroll
              for (int t = 0; t < 4; ++t) {
                acc0 = fma(h[t], y[t].x, acc0);
                acc1 = fma(h[t], y[t].y, acc1);
              }
            }
            e4[p] = ent4[(int64_t)min(c0 + p + PF, last) * dimup + row];
          }
        }
      }
      if (base < dimup) {
        const R out0 = acc0 * scale, out1 = acc1 * scale;
        if constexpr (CPLX) {
          cplx<R> o;
          o.x = out0;
          o.y = out1;
          reinterpret_cast<cplx<R> *>(hcol0)[row] = o;
          if (lz)
            dot += (double)(x0 * scale) * (double)out0 + (double)(x1 * scale) * (double)out1;
        } else {
          hcol0[row] = out0;
          if (lz) dot += (double)(x0 * scale) * (double)out0;
          if (two) {
            hcol1[row] = out1;
            if (lz) dot += (double)(x1 * scale) * (double)out1;
          }
        }
      }
    }
  }
  if (lz) block_partial_lds<BS>(dot, a.partials, slot, reinterpret_cast<double *>(smem));
}

// Transposed form of H x v, for sectors far larger than the on-chip caches.
// The down-spin term sum_q H_dw(idw, jdw_q) v(iup, jdw_q) reads, for every
// column, its ~nnz neighbour columns: on a 1.66e8-state sector (54 hops per
// row) that is ~54 reads of the whole vector from HBM per product.  In the
// layout vt(idw, iup) = v(iup, idw) (idw fastest) the same term gathers
// inside one row of vt, like the up-spin term inside a column of v, so
//   1. k_hxv_buf (skip_dw):        Hv = diag v + H_up v
//   2. k_transpose:                 vt = v^T
//   3. k_hxv_buf (skip_diag, skip_dw) on vt with the down-spin tables,
//      each workgroup staging its row of vt in LDS and writing the result
//      over it:                     vt = (H_dw v)^T
//   4. k_transpose<ADD>:            Hv += vt^T
// moves 9 vector lengths of HBM traffic per product (read and write in each
// of the four) and keeps every gather in LDS.  The two gather passes write
// their partials of <v|Hv> into one row per chain, summed by k_lz_alpha.
//
// Element (i, j) of the input sits at in[j * ni + i] (i fastest), of the
// output at out[i * nj + j] (j fastest); a 64 x 64 tile goes through LDS so
// that both the loads and the stores are 64-lane runs of consecutive
// elements.  Lanczos mode (lz non-null): chain blockIdx.z reads buffer
// in1 on odd steps, and a finished chain is skipped.
struct TrArgs {
  const void *in0, *in1;
  void *out;
  int ni, nj;
  int64_t vstride;
  const LanczosDev *lz;
};

constexpr int kTile = 64;

__device__ inline void acc_elem(double &a, double b) { a += b; }
__device__ inline void acc_elem(float &a, float b) { a += b; }
__device__ inline void acc_elem(double2 &a, const double2 &b) {
  a.x += b.x;
  a.y += b.y;
}
__device__ inline void acc_elem(float2 &a, const float2 &b) {
  a.x += b.x;
  a.y += b.y;
}

// E: complex element (cplx<R>) or, in real-vector mode, a real.
// Tiles are 64 x 64 for every element size (complex128: 66.5 KB of LDS,
// two workgroups per CU), taken in super-tile order: the workgroups of one
// launch round cover SW x SW neighbouring tiles (SW * 64 * sizeof(E) = 8 or
// 4 KB runs of each row read or written), walking down a strip of SW tile
// columns before the next strip, so the streams of a round stay within few
// HBM pages.  On the 12,870 x 12,870 transposes (rows of 64 tiles in
// launch order before): complex128 1.43 -> 1.14 ms, the add form 2.11 ->
// 1.67 ms, FP64 reals 0.68 -> 0.57 ms.
template <typename E>
constexpr int tr_sw() {
  return sizeof(E) > 8 ? 8 : 16;
}

// Tile (i, j) of a ni_t x nj_t tile grid at launch position lin: strips of
// sw tile columns (the last one narrower), each walked as sw x sw super-tiles
// (the last one shorter), tile rows fastest inside a super-tile.  Every
// position maps to one tile and every tile to one position.
__device__ inline int2 tile_at(int lin, int ni_t, int nj_t, int sw) {
  const int sj = lin / (sw * ni_t);
  const int w = min(sw, nj_t - sj * sw);
  const int r = lin - sj * sw * ni_t;
  const int si = r / (sw * w);
  const int h = min(sw, ni_t - si * sw);
  const int q = r - si * sw * w;
  return make_int2(si * sw + q % h, sj * sw + q / h);
}

template <typename E, bool ADD>
__global__ __launch_bounds__(kBlock) void k_transpose(const TrArgs a) {
  constexpr int T = kTile;
  constexpr int kRows = kBlock / T;  // rows of the tile per pass of the workgroup
  constexpr int NP = T / kRows;      // passes
  __shared__ E s[T][T + 1];
  const int64_t off = (int64_t)blockIdx.z * a.vstride;
  const E *in = static_cast<const E *>(a.in0) + off;
  if (a.lz) {
    const LanczosDev *lz = a.lz + blockIdx.z;
    if (lz->done) return;
    if (lz->k & 1) in = static_cast<const E *>(a.in1) + off;
  }
  E *out = static_cast<E *>(a.out) + off;
  const int lane = threadIdx.x % T, w = threadIdx.x / T;
  const int2 tl = tile_at(blockIdx.x, (a.ni + T - 1) / T, (a.nj + T - 1) / T, tr_sw<E>());
  const int i0 = tl.x * T, j0 = tl.y * T;
  // every load of the lane is issued before the first LDS store (indices
  // clamped into the matrix), so a tile costs one HBM round trip rather
  // than NP; the same on the way out.  The LDS stores are unconditional:
  // a clamped element lands in a tile slot whose output is not stored, and
  // without a branch between them the loads stay in flight together.
  const int i = i0 + lane, ic = min(i, a.ni - 1);
  E x[NP];
#pragma unroll
  for (int t = 0; t < NP; ++t) {
    const int jc = min(j0 + w + kRows * t, a.nj - 1);
    x[t] = in[(int64_t)jc * a.ni + ic];
  }
#pragma unroll
  for (int t = 0; t < NP; ++t) s[w + kRows * t][lane] = x[t];
  __syncthreads();
  const int j = j0 + lane, jc = min(j, a.nj - 1);
#pragma unroll
  for (int t = 0; t < NP; ++t) x[t] = s[lane][w + kRows * t];
  if (ADD) {
    E o[NP];
#pragma unroll
    for (int t = 0; t < NP; ++t) {
      const int irc = min(i0 + w + kRows * t, a.ni - 1);
      o[t] = out[(int64_t)irc * a.nj + jc];
    }
#pragma unroll
    for (int t = 0; t < NP; ++t) acc_elem(x[t], o[t]);
  }
#pragma unroll
  for (int t = 0; t < NP; ++t) {
    const int ir = i0 + w + kRows * t;
    if (ir < a.ni && j < a.nj) out[(int64_t)ir * a.nj + j] = x[t];
  }
}

// Lanczos step of the transposed form: the transpose-add of the down-spin
// result fused with the vector update, so Hv never goes through HBM.  Per
// element n = idw*DimUp + iup of chain blockIdx.z:
//   h = Hv_up[n] + Wt[iup*DimDw + idw]   (Wt: the down-spin pass, in v^T)
//   u_{k+1}[n] = h - (alpha_k/nrm_k) u_k[n] - (beta_k/nrm_{k-1}) u_{k-1}[n]
//   x[n] += (y_k/nrm_k) u_k[n]            (eigenvector pass, x non-null)
// with the arithmetic of k_lanczos_update (the same sums in the same order,
// so the result equals the transpose-add followed by that update bit for
// bit); a tile of Wt is turned through LDS as in k_transpose.  E: a real
// (real-vector mode, the pad real of an odd dim untouched: it stays 0) or
// a complex element.  Partials of |u_{k+1}|^2 per tile.
struct TUpdArgs {
  void *u0, *u1;
  const void *hv, *wt;
  void *x;
  int dimup, dimdw;
  int64_t vstride;  // elements between chains
  LanczosDev *lz;
  double *partials;
};

template <typename R>
__device__ inline R tupd(R h, R u, R p, R ca, R cb, bool first) {
  return first ? h - ca * u : h - ca * u - cb * p;
}
template <typename R, typename C>
__device__ inline C tupd(const C &h, const C &u, const C &p, R ca, R cb, bool first) {
  C w;
  w.x = tupd(h.x, u.x, p.x, ca, cb, first);
  w.y = tupd(h.y, u.y, p.y, ca, cb, first);
  return w;
}
__device__ inline double tnorm2(float w) { return (double)w * (double)w; }
__device__ inline double tnorm2(double w) { return w * w; }
template <typename C>
__device__ inline double tnorm2(const C &w) {
  return (double)w.x * (double)w.x + (double)w.y * (double)w.y;
}
template <typename R>
__device__ inline void taxpy(R &x, R c, R u) {
  x += c * u;
}
template <typename R, typename C>
__device__ inline void taxpy(C &x, R c, const C &u) {
  x.x += c * u.x;
  x.y += c * u.y;
}

template <typename R, typename E>
__global__ __launch_bounds__(kBlock) void k_tadd_update(const TUpdArgs a) {
  constexpr int T = kTile;
  constexpr int kRows = kBlock / T;  // rows of the tile per pass of the workgroup
  constexpr int NP = T / kRows;      // passes
  // elements per batch of loads in the update (registers: 4 or 5 per element)
  constexpr int CH = sizeof(E) > 8 ? NP / 4 : sizeof(E) == 8 ? NP / 2 : NP;
  __shared__ E s[T][T + 1];
  const int64_t off = (int64_t)blockIdx.z * a.vstride;
  LanczosDev *lz = a.lz + blockIdx.z;
  const int k = lz->k;
  const bool done = lz->done != 0;
  double ss = 0.0;
  if (!done) {
    const E *wt = static_cast<const E *>(a.wt) + off;
    const int ni = a.dimdw, nj = a.dimup;  // element (i = idw, j = iup)
    const int lane = threadIdx.x % T, w = threadIdx.x / T;
    const int2 tl = tile_at(blockIdx.x, (ni + T - 1) / T, (nj + T - 1) / T, tr_sw<E>());
    const int i0 = tl.x * T, j0 = tl.y * T;
    // loads batched as in k_transpose (clamped indices)
    {
      const int i = i0 + lane, ic = min(i, ni - 1);
      E y[NP];
#pragma unroll
      for (int t = 0; t < NP; ++t) y[t] = wt[(int64_t)min(j0 + w + kRows * t, nj - 1) * ni + ic];
#pragma unroll
      for (int t = 0; t < NP; ++t) s[w + kRows * t][lane] = y[t];
    }
    __syncthreads();
    const E *hv = static_cast<const E *>(a.hv) + off;
    const E *uk = static_cast<const E *>((k & 1) ? a.u1 : a.u0) + off;
    E *up = static_cast<E *>((k & 1) ? a.u0 : a.u1) + off;
    E *x = static_cast<E *>(a.x);
    const R ca = (R)(lz->alpha[k] / lz->nrm[k]);
    const R cb = k > 0 ? (R)(lz->beta[k] / lz->nrm[k - 1]) : (R)0;
    const R cy = x ? (R)(lz->y[k] / lz->nrm[k]) : (R)0;
    const int j = j0 + lane, jc = min(j, nj - 1);
#pragma unroll
    for (int t0 = 0; t0 < NP; t0 += CH) {
      E h[CH], hw[CH], u[CH], p[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int ii = w + kRows * (t0 + c);
        const int64_t n = (int64_t)min(i0 + ii, ni - 1) * nj + jc;
        h[c] = s[lane][ii];
        hw[c] = hv[n];
        u[c] = uk[n];
        if (k > 0) p[c] = up[n];
      }
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int ii = w + kRows * (t0 + c);
        const int ir = i0 + ii;
        if (ir < ni && j < nj) {
          const int64_t n = (int64_t)ir * nj + j;
          acc_elem(h[c], hw[c]);
          const E wv = tupd<R>(h[c], u[c], k > 0 ? p[c] : u[c], ca, cb, k == 0);
          up[n] = wv;
          ss += tnorm2(wv);
          if (x) taxpy<R>(x[n], cy, u[c]);
        }
      }
    }
  }
  const double b = block_sum(ss);
  if (threadIdx.x == 0) a.partials[(int64_t)blockIdx.z * gridDim.x + blockIdx.x] = b;
}

// ------------------------------------------------------ Lanczos update --
struct UpdArgs {
  void *u0, *u1;  // u_k in buffer k&1, u_{k-1} (and u_{k+1}) in the other
  const void *hv;
  void *x;        // eigenvector accumulator or null
  int64_t n;
  LanczosDev *lz;  // one per chain (blockIdx.y)
  double *partials;
  int64_t vstride;
};

// NT: nontemporal loads and stores (batches far larger than the 256 MB MALL,
// whose vectors are not reread from it).
template <bool NT, typename V>
__device__ inline V ld_stream(const V *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename V>
__device__ inline void st_stream(V *p, V v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <typename R, bool NT = false>
__global__ __launch_bounds__(kBlock) void k_lanczos_update(const UpdArgs a) {
  using C = typename Vec2<R>::T;
  const int64_t off = (int64_t)blockIdx.y * a.vstride;
  LanczosDev *lz = a.lz + blockIdx.y;
  const int k = lz->k;
  const bool done = lz->done != 0;
  double ss = 0.0;
  if (!done) {
    const C *uk = static_cast<const C *>((k & 1) ? a.u1 : a.u0) + off;
    C *up = static_cast<C *>((k & 1) ? a.u0 : a.u1) + off;
    const C *hv = static_cast<const C *>(a.hv) + off;
    const R ca = (R)(lz->alpha[k] / lz->nrm[k]);
    const R cb = k > 0 ? (R)(lz->beta[k] / lz->nrm[k - 1]) : (R)0;
    C *x = static_cast<C *>(a.x);
    const R cy = x ? (R)(lz->y[k] / lz->nrm[k]) : (R)0;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += stride) {
      const C u = ld_stream<NT>(uk + i), h = ld_stream<NT>(hv + i);
      C w;
      if (k > 0) {
        const C p = ld_stream<NT>(up + i);
        w.x = h.x - ca * u.x - cb * p.x;
        w.y = h.y - ca * u.y - cb * p.y;
      } else {
        w.x = h.x - ca * u.x;
        w.y = h.y - ca * u.y;
      }
      st_stream<NT>(up + i, w);
      ss += (double)w.x * (double)w.x + (double)w.y * (double)w.y;
      if (x) {
        C xv = x[i];
        xv.x += cy * u.x;
        xv.y += cy * u.y;
        x[i] = xv;
      }
    }
  }
  block_partial(ss, a.partials + (int64_t)blockIdx.y * gridDim.x);
}

// Per-chain finishers of the two-stage sums (one workgroup per chain).
// alpha[k] = <v_k|H v_k> from the H x v partials.
__global__ __launch_bounds__(kBlock) void k_lz_alpha(LanczosDev *lz, const double *partials,
                                                     int nblk) {
  LanczosDev *c = lz + blockIdx.x;
  const double s = sum_partials(partials + (int64_t)blockIdx.x * nblk, nblk);
  if (threadIdx.x == 0) c->alpha[c->k] = c->done ? 0.0 : s;
}

// beta_{k+1} = |u_{k+1}| from the update partials, then k -> k+1; an exact
// breakdown (or NaN) ends the chain.
__global__ __launch_bounds__(kBlock) void k_lz_beta(LanczosDev *lz, const double *partials,
                                                    int nblk) {
  LanczosDev *c = lz + blockIdx.x;
  const double s = sum_partials(partials + (int64_t)blockIdx.x * nblk, nblk);
  if (threadIdx.x == 0) {
    const int k = c->k;
    const double b = sqrt(s);
    if (c->done || !(b > 1e-300)) {
      c->beta[k + 1] = 0.0;
      c->nrm[k + 1] = 1.0;
      c->done = 1;
    } else {
      c->beta[k + 1] = b;
      c->nrm[k + 1] = b;
    }
    c->k = k + 1;
  }
}

// Start of a chain: nrm[0] = |u_0|, beta[0] = 0, k = 0, not done.
__global__ __launch_bounds__(kBlock) void k_lz_start(LanczosDev *lz, const double *partials,
                                                     int nblk) {
  LanczosDev *c = lz + blockIdx.x;
  const double s = sum_partials(partials + (int64_t)blockIdx.x * nblk, nblk);
  if (threadIdx.x == 0) {
    c->nrm[0] = sqrt(s);
    c->beta[0] = 0.0;
    c->k = 0;
    c->done = 0;
  }
}

// y[0] = |x| (the eigenvector norm, y is free after the second pass).
__global__ __launch_bounds__(kBlock) void k_lz_xnorm(LanczosDev *lz, const double *partials,
                                                     int nblk) {
  const double s = sum_partials(partials, nblk);
  if (threadIdx.x == 0) lz->y[0] = sqrt(s);
}

// ------------------------------------------------------ vector helpers --
// splitmix64 hash of (seed, i) -> [-1, 1), the start vector of the GS
// Lanczos (same numbers as the oracle's start_vector before normalisation).
__device__ inline double hash_unit(uint64_t seed, uint64_t i) {
  const uint64_t g = 0x9E3779B97F4A7C15ull;
  uint64_t z = seed * g + (i + 1ull) * g;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_start(void *u, int64_t n, uint64_t seed) {
  using C = cplx<R>;
  C *p = static_cast<C *>(u);
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    C c;
    c.x = (R)hash_unit(seed, (uint64_t)i);
    c.y = (R)0;
    p[i] = c;
  }
}

// nrm[0] = |u|, k = 0, done = 0 (resets a Lanczos run on buffer 0).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_norm0(const void *u, int64_t n, LanczosDev *lz,
                                                  double *partials) {
  using C = cplx<R>;
  const C *p = static_cast<const C *>(u);
  double ss = 0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const C c = p[i];
    ss += (double)c.x * c.x + (double)c.y * c.y;
  }
  (void)lz;
  block_partial(ss, partials);
}

// x /= |x|
template <typename R>
__global__ __launch_bounds__(kBlock) void k_normalize(void *u, int64_t n, LanczosDev *lz,
                                                      double *partials) {
  using C = cplx<R>;
  C *p = static_cast<C *>(u);
  double ss = 0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const C c = p[i];
    ss += (double)c.x * c.x + (double)c.y * c.y;
  }
  // k_lz_xnorm puts the norm in y[0]; k_scale_by then divides
  (void)lz;
  block_partial(ss, partials);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_scale_by(void *u, int64_t n, const double *den) {
  using C = cplx<R>;
  C *p = static_cast<C *>(u);
  const R s = (R)(1.0 / *den);
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    C c = p[i];
    c.x *= s;
    c.y *= s;
    p[i] = c;
  }
}

template <typename RS, typename RD>
__global__ __launch_bounds__(kBlock) void k_convert(const void *src, void *dst, int64_t n) {
  const cplx<RS> *s = static_cast<const cplx<RS> *>(src);
  cplx<RD> *d = static_cast<cplx<RD> *>(dst);
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    cplx<RD> c;
    c.x = (RD)s[i].x;
    c.y = (RD)s[i].y;
    d[i] = c;
  }
}

// n reals -> n complex numbers with zero imaginary parts (src, dst distinct).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_expand_real(const R *src, cplx<R> *dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    cplx<R> c;
    c.x = src[i];
    c.y = (R)0;
    dst[i] = c;
  }
}

// --------------------------------------------------- excitation vectors --
// Gather form of O|psi> (VCA_GF_NORMAL.f90:184-198 and 624-660): for every
// target state j of sector J and every operator of the channel (all on the
// same spin), the source string is the target string with the operator
// undone; the sign is the parity of occupied orbitals below the position
// in the source string (cdg/c of VCA_SETUP.f90:724-750).  Each target has
// at most one source per operator, so no atomics are needed.
struct ExcArgs {
  void *out;
  const void *psi;
  int dimup_j, dimdw_j, dimup_i, dimdw_i;
  const uint32_t *map_j;  // the J-sector map of the acted-on spin
  const uint32_t *binom;
  int spin;
  const ExcOp *ops;  // one per chain (blockIdx.y)
  LanczosDev *lz;    // one per chain
  double *partials;
  int64_t vstride;
};

// RO (real-vector chains): psi holds real amplitudes (dim_i reals), every
// coefficient is real, and out holds dim_j reals per chain with the pad real
// of an odd dim_j set to 0.
template <typename R, bool RO>
__global__ __launch_bounds__(kBlock) void k_excite(const ExcArgs a) {
  using C = cplx<R>;
  const C *psi = static_cast<const C *>(a.psi);
  const R *psi_r = static_cast<const R *>(a.psi);
  C *out = static_cast<C *>(a.out) + (int64_t)blockIdx.y * a.vstride;
  R *out_r = static_cast<R *>(a.out) + (int64_t)blockIdx.y * a.vstride;
  const ExcOp op = a.ops[blockIdx.y];
  LanczosDev *lz = a.lz + blockIdx.y;
  const int64_t n = (int64_t)a.dimup_j * a.dimdw_j;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  double ss = 0;
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += stride) {
    const int jup = (int)(j % a.dimup_j), jdw = (int)(j / a.dimup_j);
    const uint32_t t = a.map_j[a.spin == 1 ? jup : jdw];
    double vr = 0, vi = 0;
    for (int o = 0; o < op.nops; ++o) {
      const int pos = op.pos[o];
      const uint32_t bit = 1u << pos;
      uint32_t s;
      if (op.dagger[o]) {
        if (!(t & bit)) continue;
        s = t & ~bit;
      } else {
        if (t & bit) continue;
        s = t | bit;
      }
      const double sg = parity_below(s, pos) ? -1.0 : 1.0;
      const uint32_t r = rank_string(s, a.binom);
      const int64_t i = a.spin == 1 ? (int64_t)r + (int64_t)jdw * a.dimup_i
                                    : (int64_t)jup + (int64_t)r * a.dimup_i;
      if (RO) {
        vr += sg * op.cre[o] * (double)psi_r[i];
      } else {
        const C p = psi[i];
        // coef * sg * psi
        vr += sg * (op.cre[o] * (double)p.x - op.cim[o] * (double)p.y);
        vi += sg * (op.cre[o] * (double)p.y + op.cim[o] * (double)p.x);
      }
    }
    if (RO) {
      const R c = (R)vr;
      out_r[j] = c;
      ss += (double)c * c;
    } else {
      C c;
      c.x = (R)vr;
      c.y = (R)vi;
      out[j] = c;
      ss += (double)c.x * c.x + (double)c.y * c.y;
    }
  }
  if (RO && (n & 1) && blockIdx.x == 0 && threadIdx.x == 0) out_r[n] = (R)0;
  (void)lz;
  block_partial(ss, a.partials + (int64_t)blockIdx.y * gridDim.x);
}

// Phase of a complex vector psi = e^{i phi} r (r real): sum_i psi_i^2 =
// e^{2 i phi} |r|^2.  Two-stage: per-block partials of (Re, Im) of sum psi^2.
__global__ __launch_bounds__(kBlock) void k_psi_sq(const double2 *psi, int64_t n,
                                                   double *partials) {
  double sr = 0, si = 0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double2 p = psi[i];
    sr += p.x * p.x - p.y * p.y;
    si += 2.0 * p.x * p.y;
  }
  const double br = block_sum(sr);
  const double bi = block_sum(si);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = br;
    partials[2 * blockIdx.x + 1] = bi;
  }
}

// out_i = Re(psi_i e^{-i phi}) (cp = cos phi, sp = sin phi), in RD; per-block
// partials of |Re|^2 and |Im|^2 of psi e^{-i phi}.
template <typename RD>
__global__ __launch_bounds__(kBlock) void k_psi_real(const double2 *psi, int64_t n, double cp,
                                                     double sp, RD *out, double *partials) {
  double sre = 0, sim = 0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double2 p = psi[i];
    const double re = p.x * cp + p.y * sp, im = p.y * cp - p.x * sp;
    out[i] = (RD)re;
    sre += re * re;
    sim += im * im;
  }
  const double br = block_sum(sre);
  const double bi = block_sum(sim);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = br;
    partials[2 * blockIdx.x + 1] = bi;
  }
}

// ================================================== host launch wrappers
static int grid_for(int64_t n) {
  int64_t g = (n + kBlock * 4 - 1) / (kBlock * 4);
  if (g < 1) g = 1;
  if (g > 32768) g = 32768;
  return (int)g;
}

// Hop order of a one-plane table (DevSpin::ent4p), chosen per half
// wavefront (32 consecutive rows, one per lane) so that the LDS gathers of
// one slot spread over the banks.  The LDS serves 32 lanes per cycle: an
// 8-byte read from 64 banks of 4 bytes (bank pair = address / 8 mod 32), a
// 4-byte read from 32 (bank = address / 4 mod 32); lanes reading one
// address share it (broadcast), distinct addresses in one bank are served
// one after another.  The reads of a slot: the element, 8 bytes
// (FP64 one-plane, or the FP32 column pair of k_hxv_plane2 at
// 2 ya - col), and the amplitude, rsz bytes.  The order of a row's hops is
// free (they are summed), so slot by slot each lane (the one with most hops
// left first) takes the hop whose element and amplitude addresses add the
// fewest new addresses to already used banks.  Rows keep their hops; only
// the slots move: perm[slot] is the slot of the original table the entry
// came from (a padding slot where the row has no hop left), so that the
// complex table (DevSpin::ent4, the same hops at complex element addresses,
// whose 8-byte FP32 reads fall in the same bank pairs up to a shift) can
// take the same order.  VCA_BANK_ORDER=0 keeps the tables' order.
static void order_plane_hops(std::vector<uint32_t> &ent4, const std::vector<int32_t> &cnt,
                             int dim, int nc, int rsz, uint32_t pad, uint32_t col,
                             std::vector<int32_t> &perm) {
  constexpr int dom = 32;  // lanes served in one LDS cycle
  // bank groups (32 each): the element read is 8 bytes, the amplitude read
  // rsz bytes
  auto group_e = [&](uint32_t ya) {
    const uint32_t addr = rsz == 8 ? ya : 2 * ya - col;
    return (int)((addr / 8) % 32);
  };
  auto group_a = [&](uint32_t aa) { return (int)((aa / rsz) % 32); };
  std::vector<std::vector<std::pair<uint32_t, int>>> rem(dom);
  std::vector<int> load_e(32), load_a(32), order(dom);
  perm.resize(ent4.size());
  for (size_t i = 0; i < perm.size(); ++i) perm[i] = (int32_t)(i / 4 / dim * 4 + i % 4);
  std::vector<int64_t> stamp_e((size_t)(kLdsAddrMask + 1) / rsz + 1, -1);
  std::vector<int64_t> stamp_a((size_t)kLdsAmpLimit / rsz + 1, -1);
  int64_t gen = 0;
  auto slot_of = [&](int q, int row) { return ((int64_t)(q / 4) * dim + row) * 4 + q % 4; };
  for (int i0 = 0; i0 < dim; i0 += 64) {
    const int i1 = std::min(dim, i0 + 64);
    int mx = 0;
    for (int i = i0; i < i1; ++i) mx = std::max(mx, (int)cnt[i]);
    const int nslot = std::min(4 * ((mx + 3) / 4), 4 * nc);
    for (int d0 = i0; d0 < i1; d0 += dom) {
      const int nl = std::min(i1, d0 + dom) - d0;
      for (int l = 0; l < nl; ++l) {
        rem[l].clear();
        for (int q = 0; q < cnt[d0 + l]; ++q) rem[l].push_back({ent4[slot_of(q, d0 + l)], q});
      }
      for (int q = 0; q < nslot; ++q) {
        ++gen;
        std::fill(load_e.begin(), load_e.end(), 0);
        std::fill(load_a.begin(), load_a.end(), 0);
        for (int l = 0; l < nl; ++l) order[l] = l;
        std::stable_sort(order.begin(), order.begin() + nl,
                         [&](int x, int y) { return rem[x].size() > rem[y].size(); });
        for (int k = 0; k < nl; ++k) {
          const int l = order[k];
          uint32_t w = pad;
          int from = cnt[d0 + l];  // a padding slot of the row
          if (!rem[l].empty()) {
            size_t best = 0;
            int bc = std::numeric_limits<int>::max();
            for (size_t c = 0; c < rem[l].size() && bc > 0; ++c) {
              const uint32_t e = rem[l][c].first;
              const uint32_t ya = e & kLdsAddrMask, aa = e >> kLdsAddrBits;
              const int cost = (stamp_e[ya / rsz] == gen ? 0 : load_e[group_e(ya)]) +
                               (stamp_a[aa / rsz] == gen ? 0 : load_a[group_a(aa)]);
              if (cost < bc) {
                bc = cost;
                best = c;
              }
            }
            w = rem[l][best].first;
            from = rem[l][best].second;
            rem[l][best] = rem[l].back();
            rem[l].pop_back();
          }
          const uint32_t ya = w & kLdsAddrMask, aa = w >> kLdsAddrBits;
          if (stamp_e[ya / rsz] != gen) {
            stamp_e[ya / rsz] = gen;
            ++load_e[group_e(ya)];
          }
          if (stamp_a[aa / rsz] != gen) {
            stamp_a[aa / rsz] = gen;
            ++load_a[group_a(aa)];
          }
          ent4[slot_of(q, d0 + l)] = w;
          perm[slot_of(q, d0 + l)] = from;
        }
      }
    }
  }
}

void build_spin_tables(Sector &sec, DevSpin &sp, const SpinModel &m, int ns, int npart,
                       const uint32_t *d_binom, double e_add) {
  Binom bn;
  sp.npart = npart;
  sp.dim = (int)bn(ns, npart);
  sp.nterms = (int)m.terms.size();
  sp.real_amp = m.real_amp;
  const int bound = npart * (ns - npart);
  sp.maxnnz = std::max(1, std::min(sp.nterms, bound));
  const hipStream_t st = sec.stream;
  std::vector<int2> ab(std::max(1, sp.nterms));
  std::vector<double2> amp64(2 * std::max(1, sp.nterms));
  std::vector<float2> amp32(2 * std::max(1, sp.nterms));
  for (int t = 0; t < sp.nterms; ++t) {
    ab[t] = make_int2(m.terms[t].a, m.terms[t].b);
    amp64[2 * t] = make_double2(m.terms[t].re, m.terms[t].im);
    amp64[2 * t + 1] = make_double2(-m.terms[t].re, -m.terms[t].im);
    amp32[2 * t] = make_float2((float)m.terms[t].re, (float)m.terms[t].im);
    amp32[2 * t + 1] = make_float2(-(float)m.terms[t].re, -(float)m.terms[t].im);
  }
  std::vector<double> pair((size_t)kMaxNs * kMaxNs);
  for (int a = 0; a < kMaxNs; ++a)
    for (int b = 0; b < kMaxNs; ++b) pair[(size_t)a * kMaxNs + b] = m.pair[a][b];
  int2 *d_ab;
  double *d_eps, *d_pair;
  VCA_HIP(hipMalloc(&sp.map, sizeof(uint32_t) * sp.dim));
  VCA_HIP(hipMalloc(&sp.energy, sizeof(double) * sp.dim));
  VCA_HIP(hipMalloc(&sp.cnt, sizeof(int32_t) * sp.dim));
  VCA_HIP(hipMalloc(&sp.ent, sizeof(uint32_t) * (size_t)sp.dim * sp.maxnnz));
  // slots past a row's hop count hold 0 (column 0, amplitude index 0)
  VCA_HIP(hipMemsetAsync(sp.ent, 0, sizeof(uint32_t) * (size_t)sp.dim * sp.maxnnz, st));
  VCA_HIP(hipMalloc(&sp.amp64, sizeof(double2) * amp64.size()));
  VCA_HIP(hipMalloc(&sp.amp32, sizeof(float2) * amp32.size()));
  VCA_HIP(hipMalloc(&d_ab, sizeof(int2) * ab.size()));
  VCA_HIP(hipMalloc(&d_eps, sizeof(double) * kMaxNs));
  VCA_HIP(hipMalloc(&d_pair, sizeof(double) * pair.size()));
  VCA_HIP(hipMemcpyAsync(sp.amp64, amp64.data(), sizeof(double2) * amp64.size(),
                         hipMemcpyHostToDevice, st));
  VCA_HIP(hipMemcpyAsync(sp.amp32, amp32.data(), sizeof(float2) * amp32.size(),
                         hipMemcpyHostToDevice, st));
  VCA_HIP(hipMemcpyAsync(d_ab, ab.data(), sizeof(int2) * ab.size(), hipMemcpyHostToDevice, st));
  VCA_HIP(hipMemcpyAsync(d_eps, m.eps, sizeof(double) * kMaxNs, hipMemcpyHostToDevice, st));
  VCA_HIP(hipMemcpyAsync(d_pair, pair.data(), sizeof(double) * pair.size(),
                         hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_build_spin, dim3((sp.dim + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                     ns, npart, sp.dim, sp.nterms, d_ab, d_binom, d_eps, d_pair, e_add,
                     sp.map, sp.energy, sp.cnt, sp.ent);
  VCA_HIP(hipGetLastError());
  const int nc = (sp.maxnnz + 3) / 4;
  VCA_HIP(hipMalloc(&sp.meta, sizeof(RowMeta) * sp.dim));
  VCA_HIP(hipMalloc(&sp.ent4, sizeof(uint32_t) * 4 * (size_t)nc * sp.dim));
  // LDS-address entries for the sector's element size: the column must end
  // below 2^18 bytes and the amplitudes (with the zero) within 2^14 bytes
  const int esz = sec.precision == VCA_FP64 ? (int)sizeof(double2) : (int)sizeof(float2);
  const int namp = 2 * sp.nterms;
  sp.lds_form = (size_t)(namp + 1) * esz <= kLdsAmpLimit &&
                (size_t)buf_lds(namp, 0, 0, esz).col + (size_t)sp.dim * esz <= kLdsAddrMask + 1;
  hipLaunchKernelGGL(k_pack_rows, dim3((sp.dim + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                     sp.dim, sp.maxnnz, sp.map, sp.energy, sp.cnt, sp.ent, sp.meta, sp.ent4, esz,
                     namp, sp.lds_form ? 1 : 0);
  VCA_HIP(hipGetLastError());
  // signed-column addresses (every hopping term the same real amplitude t)
  bool one_amp = sp.real_amp && sp.nterms >= 1;
  for (int t = 0; t < sp.nterms && one_amp; ++t)
    one_amp = amp64[2 * t].x == amp64[0].x && amp64[2 * t].y == 0.0;
  sp.sgn_form = one_amp && (size_t)sgn_lds(sp.dim, esz).total / esz <= 0xffff;
  if (sp.sgn_form) {
    VCA_HIP(hipMalloc(&sp.ent4s, sizeof(uint16_t) * 4 * (size_t)nc * sp.dim));
    hipLaunchKernelGGL(k_pack_rows, dim3((sp.dim + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                       sp.dim, sp.maxnnz, sp.map, sp.energy, sp.cnt, sp.ent, sp.meta, sp.ent4s,
                       esz, namp, 2);
  }
  // one-plane addresses (real amplitudes only)
  const int rsz = esz / 2;
  sp.plane_form = sp.real_amp && (size_t)(namp + 1) * rsz <= kLdsAmpLimit &&
                  (size_t)buf_lds(namp, 0, 0, rsz).col + (size_t)sp.dim * rsz <= kLdsAddrMask + 1;
  if (sp.plane_form) {
    VCA_HIP(hipMalloc(&sp.ent4p, sizeof(uint32_t) * 4 * (size_t)nc * sp.dim));
    hipLaunchKernelGGL(k_pack_rows, dim3((sp.dim + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                       sp.dim, sp.maxnnz, sp.map, sp.energy, sp.cnt, sp.ent, sp.meta, sp.ent4p,
                       rsz, namp, 1);
    VCA_HIP(hipGetLastError());
    // bank-aware hop order for the columns long enough for the plane passes
    const char *bo = std::getenv("VCA_BANK_ORDER");
    if (sp.dim >= 128 && !(bo && bo[0] == '0')) {
      std::vector<uint32_t> h4((size_t)4 * nc * sp.dim);
      std::vector<int32_t> hc(sp.dim);
      VCA_HIP(hipMemcpyAsync(h4.data(), sp.ent4p, sizeof(uint32_t) * h4.size(),
                             hipMemcpyDeviceToHost, st));
      VCA_HIP(hipMemcpyAsync(hc.data(), sp.cnt, sizeof(int32_t) * hc.size(),
                             hipMemcpyDeviceToHost, st));
      VCA_HIP(hipStreamSynchronize(st));
      const uint32_t pad = (uint32_t)buf_lds(namp, 0, 0, rsz).col |
                           ((uint32_t)(namp * rsz) << kLdsAddrBits);
      std::vector<int32_t> perm;
      order_plane_hops(h4, hc, sp.dim, nc, rsz, pad, (uint32_t)buf_lds(namp, 0, 0, rsz).col,
                       perm);
      VCA_HIP(hipMemcpyAsync(sp.ent4p, h4.data(), sizeof(uint32_t) * h4.size(),
                             hipMemcpyHostToDevice, st));
      if (sp.lds_form && esz == 8) {  // the complex64 table in the same order
        std::vector<uint32_t> c4(h4.size()), o4(h4.size());
        VCA_HIP(hipMemcpyAsync(c4.data(), sp.ent4, sizeof(uint32_t) * c4.size(),
                               hipMemcpyDeviceToHost, st));
        VCA_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < o4.size(); ++i) {
          const int row = (int)(i / 4 % (size_t)sp.dim), q = perm[i];
          o4[i] = c4[((size_t)(q / 4) * sp.dim + row) * 4 + q % 4];
        }
        VCA_HIP(hipMemcpyAsync(sp.ent4, o4.data(), sizeof(uint32_t) * o4.size(),
                               hipMemcpyHostToDevice, st));
      }
      VCA_HIP(hipStreamSynchronize(st));
    }
    // the column-pair addresses of k_hxv_plane2, where a pair fits the LDS
    const uint32_t col = (uint32_t)buf_lds(namp, 0, 0, rsz).col;
    if ((size_t)col + 2 * (size_t)align16(sp.dim * rsz) <= kLdsBig) {
      const int64_t n = (int64_t)4 * nc * sp.dim;
      VCA_HIP(hipMalloc(&sp.ent4q, sizeof(uint32_t) * n));
      hipLaunchKernelGGL(k_pair_addr, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock),
                         0, st, sp.ent4p, sp.ent4q, n, col);
      VCA_HIP(hipGetLastError());
    }
  }
  VCA_HIP(hipGetLastError());
  VCA_HIP(hipStreamSynchronize(st));
  VCA_HIP(hipFree(d_ab));
  VCA_HIP(hipFree(d_eps));
  VCA_HIP(hipFree(d_pair));
}

// Per-block partial sums of the grid reductions.  Sized once for the
// largest grid any kernel of the sector uses, so it is never reallocated
// while launches are queued; a larger request first drains the stream.
static void ensure_partials(Sector &sec, int nblocks) {
  if (sec.partials_cap >= nblocks) return;
  if (sec.partials) {
    VCA_HIP(hipStreamSynchronize(sec.stream));
    clear_graphs(sec);  // recorded launches hold the old pointer
    VCA_HIP(hipFree(sec.partials));
  }
  const int cap = std::max(nblocks, std::max(4096, sec.dw.dim) * std::max(1, sec.nb_cap));
  VCA_HIP(hipMalloc(&sec.partials, sizeof(double) * cap));
  sec.partials_cap = cap;
}

// Launch shape of H x v on a sector: one workgroup per column (k_hxv_col,
// with BS lanes and KR rows per lane) when columns are long, several
// columns per workgroup (k_hxv) otherwise.  The shape of a sector comes from
// the autotuner (sec.hxv_variant) or, before tuning, from the default below.
struct HxvShape {
  bool onecol_kernel;
  bool buf;    // buffer-load form (k_hxv_buf) when the column can be staged
  bool tform;  // transposed form (gather passes on v and v^T, k_transpose)
  bool plane;  // its gather passes one real plane at a time (k_hxv_plane)
  bool cpair;  // its gather passes whole complex columns through the pair tables
  bool ov;     // overlapped short-column form (k_hxv_ov)
  bool sgn;    // ... with the signed-column up-spin table
  int bs, kr, u, cpb, grid;
};

// Variants of the one-column kernel the autotuner times: (BS, KR, U, form),
// form 0 = k_hxv_col, 1 = k_hxv_buf (U fixed by KR there), 2 = transposed
// form with k_hxv_buf passes, 3 = transposed form with k_hxv_plane passes,
// 4 = k_hxv_ov (KR = rows per lane, NR; columns of at most NR * BS rows),
// 5 = its signed-column variant, 6 = transposed form with complex-pair
// passes (k_hxv_plane2<CPLX>: whole complex columns staged, real amplitudes,
// the column-pair hop tables of both spins).
// Forms 2 to 6 are timed only on sectors they fit (hxv_variant_ok).
static const int kVariants[][4] = {
    {256, 4, 4, 0}, {256, 2, 8, 0},  {1024, 1, 8, 0}, {1024, 2, 8, 0}, {1024, 4, 4, 0},
    {512, 2, 8, 0}, {256, 1, 16, 0}, {512, 1, 16, 0}, {1024, 1, 16, 0}, {256, 4, 4, 1},
    {256, 2, 8, 1}, {512, 2, 8, 1},  {512, 1, 8, 1},  {1024, 1, 8, 1}, {1024, 2, 8, 1},
    {512, 1, 8, 2}, {1024, 1, 8, 2}, {512, 2, 8, 2},  {1024, 2, 8, 2}, {512, 1, 8, 3},
    {1024, 1, 8, 3}, {512, 2, 4, 4}, {1024, 1, 8, 4}, {256, 2, 4, 4}, {512, 2, 4, 5},
    {1024, 1, 8, 5}, {256, 2, 4, 5}, {1024, 1, 8, 6}, {512, 1, 8, 6}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

// The transposed form pays 9 vector lengths of HBM traffic per product
// against ~2 + nnz_dw re-reads of neighbour columns by the direct form, which
// only hit on-chip while the vector fits the 256 MB MALL: offered from 2^22
// states, when both spins' columns (form 2; its buffer loads address one
// vector with 32-bit offsets) or one real plane of them (form 3, real
// amplitudes) can be staged in LDS.
constexpr int64_t kTformMinDim = 1 << 22;

static size_t hxv_esz(const Sector &sec) {
  return sec.precision == VCA_FP64 ? sizeof(double2) : sizeof(float2);
}

// VCA_TFORM_MIN_DIM lowers the size threshold (tests run the form on
// sectors small enough for the oracle)
int64_t tform_min_dim() {
  const char *e = std::getenv("VCA_TFORM_MIN_DIM");
  return e && *e ? std::atoll(e) : kTformMinDim;
}

static bool tform_fits(const Sector &sec, bool plane) {
  const size_t esz = hxv_esz(sec), rsz = esz / 2;
  auto staged = [&](const DevSpin &sp) {
    if (sp.dim < 128) return false;
    if (plane)
      return sp.plane_form && (size_t)buf_lds(2 * sp.nterms, sp.dim, 0, (int)rsz).total <= kLdsBig;
    return sp.lds_form && (size_t)buf_lds(2 * sp.nterms, sp.dim, 0, (int)esz).total <= kLdsBig;
  };
  return sec.dim >= tform_min_dim() && staged(sec.up) && staged(sec.dw) &&
         (plane || (uint64_t)sec.dim * esz < (1ull << 31));
}

bool hxv_variant_ok(const Sector &sec, int v) {
  if (v < 0 || v >= kNumVariants) return false;
  if (sec.up.dim < 128) return v == 0;  // one shape (k_hxv)
  const int form = kVariants[v][3];
  if (form == 4) {
    const size_t esz = hxv_esz(sec);
    return sec.up.lds_form && sec.up.dim <= kVariants[v][0] * kVariants[v][1] &&
           (size_t)buf_lds(2 * sec.up.nterms, sec.up.dim, 0, (int)esz).total <= kLdsBig &&
           (uint64_t)sec.dim * esz < (1ull << 31);
  }
  if (form == 5) {  // signed-column overlapped form
    const size_t esz = hxv_esz(sec);
    return sec.up.sgn_form && sec.up.real_amp && sec.dw.real_amp &&
           sec.up.dim <= kVariants[v][0] * kVariants[v][1] &&
           (size_t)sgn_lds(sec.up.dim, (int)esz).total <= kLdsBig &&
           (uint64_t)sec.dim * esz < (1ull << 31);
  }
  if (form == 6)
    return tform_fits(sec, true) && sec.up.ent4q != nullptr && sec.dw.ent4q != nullptr;
  return form < 2 || tform_fits(sec, form == 3);
}

static HxvShape hxv_shape(const Sector &sec) {
  HxvShape h{};
  if (sec.up.dim >= 128) {
    h.onecol_kernel = true;
    int v = sec.hxv_variant;
    if (!hxv_variant_ok(sec, v)) v = sec.up.dim >= 4096 ? 4 : 2;  // default
    h.bs = kVariants[v][0];
    h.kr = kVariants[v][1];
    h.u = kVariants[v][2];
    h.buf = kVariants[v][3] != 0;
    h.tform = kVariants[v][3] == 2 || kVariants[v][3] == 3 || kVariants[v][3] == 6;
    h.plane = kVariants[v][3] == 3;
    h.cpair = kVariants[v][3] == 6;
    h.ov = kVariants[v][3] == 4 || kVariants[v][3] == 5;
    h.sgn = kVariants[v][3] == 5;
    h.cpb = 1;
    h.grid = sec.dw.dim;
  } else {
    h.onecol_kernel = false;
    h.bs = kBlock;
    h.kr = 1;
    h.cpb = std::max(1, kBlock / sec.up.dim);
    h.grid = (sec.dw.dim + h.cpb - 1) / h.cpb;
  }
  return h;
}

int hxv_num_variants(const Sector &sec) { return sec.up.dim < 128 ? 1 : kNumVariants; }

int hxv_scratch_vectors(const Sector &sec) { return hxv_shape(sec).tform ? 1 : 0; }

// Partial sums per chain of one H x v: one per column, plus one per row of
// v^T in the transposed form.
// Real-vector gather pass over the columns of spin sp's table (column length
// sp.dim): two columns per workgroup (k_hxv_plane2) when both fit the LDS
// (VCA_PLANE2=0 keeps one).
static bool real_pass_two(const Sector &sec, const DevSpin &sp) {
  const char *e = std::getenv("VCA_PLANE2");
  if (e && e[0] == '0') return false;
  const int rsz = sec.precision == VCA_FP64 ? (int)sizeof(double) : (int)sizeof(float);
  return (size_t)buf_lds(2 * sp.nterms, 0, 0, rsz).col + 2 * (size_t)align16(sp.dim * rsz) <=
         kLdsBig;
}

// Workgroups of the real-vector passes: over the DimDw columns of v (up-spin
// tables), then over the DimUp rows of v^T (down-spin tables).
static int real_grid_up(const Sector &sec) {
  return real_pass_two(sec, sec.up) ? (sec.dw.dim + 1) / 2 : sec.dw.dim;
}
static int real_grid_dw(const Sector &sec) {
  return real_pass_two(sec, sec.dw) ? (sec.up.dim + 1) / 2 : sec.up.dim;
}

// Real-vector mode runs the transposed form (one real plane) where the
// sector's tuned shape is transposed, else the direct form k_hxv_buf_real.
static bool real_uses_tform(const Sector &sec) {
  return hxv_shape(sec).tform && tform_fits(sec, true);
}

static bool direct_real_fits(const Sector &sec) {
  const int rsz = sec.precision == VCA_FP64 ? (int)sizeof(double) : (int)sizeof(float);
  return sec.up.plane_form &&
         (size_t)buf_lds(2 * sec.up.nterms, sec.up.dim, sec.dw.maxnnz, rsz).total <= kLdsBig &&
         2 * (uint64_t)sec.dim * rsz < (1ull << 31);
}

static int hxv_partials(const Sector &sec, const HxvShape &h) {
  if (sec.real_vec)
    return real_uses_tform(sec) ? real_grid_up(sec) + real_grid_dw(sec) : sec.dw.dim;
  return h.tform ? sec.dw.dim + sec.up.dim : h.grid;
}

bool real_mode_ok(const Sector &sec) {
  return sec.up.dim >= 128 && sec.up.real_amp && sec.dw.real_amp &&
         (real_uses_tform(sec) || direct_real_fits(sec));
}

int64_t vec_elems(const Sector &sec) { return sec.real_vec ? (sec.dim + 1) / 2 : sec.dim; }

// Scratch vectors of the transposed form for nb chains; like the partials,
// a larger request first drains the stream and drops the recorded graphs.
static void ensure_tbuf(Sector &sec, int nb) {
  if (!hxv_shape(sec).tform || sec.tbuf_cap >= nb) return;
  if (sec.tbuf) {
    VCA_HIP(hipStreamSynchronize(sec.stream));
    clear_graphs(sec);
    VCA_HIP(hipFree(sec.tbuf));
    sec.tbuf = nullptr;
    sec.tbuf_cap = 0;
  }
  VCA_HIP(hipMalloc(&sec.tbuf, sec.vec_bytes() * (size_t)nb));
  sec.tbuf_cap = nb;
}

// Tiles of the fused transpose-add + update (k_tadd_update) per chain.
// The Lanczos update fused with the transpose-add (k_tadd_update): in
// real-vector mode on the transposed form, and for complex vectors wherever
// the H x v is the transposed form.
static bool fused_tadd(const Sector &sec) {
  return sec.real_vec ? real_uses_tform(sec) : hxv_shape(sec).tform;
}

static int tadd_tiles(const Sector &sec) {
  const int t = kTile;
  return ((sec.dw.dim + t - 1) / t) * ((sec.up.dim + t - 1) / t);
}

// Allocates the partial-sum buffer for the largest grid of a Lanczos step
// on nb chains (and the transposed form's scratch), so that launches
// recorded into a graph never allocate.
void ensure_step_partials(Sector &sec, int nb) {
  tune_real_direct(sec);  // before any launch is recorded
  const HxvShape h = hxv_shape(sec);
  int n = std::max(hxv_partials(sec, h), grid_for(sec.dim));
  if (fused_tadd(sec)) n = std::max(n, tadd_tiles(sec));
  ensure_partials(sec, n * nb);
  ensure_tbuf(sec, nb);
}

constexpr int kXcdChainsMin = 64;

template <typename R, bool REAL_AMP, bool STAGE, int BS, int KR, int U>
static void launch_col(Sector &sec, const HxvArgs &a, dim3 grid, size_t smem) {
  static std::once_flag attr_once;
  auto fn = k_hxv_col<R, REAL_AMP, STAGE, BS, KR, U>;
  std::call_once(attr_once, [&]() {
    VCA_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBig));
  });
  hipLaunchKernelGGL(fn, grid, dim3(BS), smem, sec.stream, a);
}

template <typename R, bool REAL_AMP, bool STAGE>
static void launch_col_s(Sector &sec, const HxvArgs &a, const HxvShape &h, dim3 grid,
                         size_t smem) {
  const int bs = h.bs, kr = h.kr, u = h.u;
  if (bs == 256 && kr == 4) launch_col<R, REAL_AMP, STAGE, 256, 4, 4>(sec, a, grid, smem);
  else if (bs == 256 && kr == 2) launch_col<R, REAL_AMP, STAGE, 256, 2, 8>(sec, a, grid, smem);
  else if (bs == 1024 && kr == 1 && u == 8)
    launch_col<R, REAL_AMP, STAGE, 1024, 1, 8>(sec, a, grid, smem);
  else if (bs == 1024 && kr == 2) launch_col<R, REAL_AMP, STAGE, 1024, 2, 8>(sec, a, grid, smem);
  else if (bs == 1024 && kr == 4) launch_col<R, REAL_AMP, STAGE, 1024, 4, 4>(sec, a, grid, smem);
  else if (bs == 512 && kr == 2) launch_col<R, REAL_AMP, STAGE, 512, 2, 8>(sec, a, grid, smem);
  else if (bs == 256 && kr == 1) launch_col<R, REAL_AMP, STAGE, 256, 1, 16>(sec, a, grid, smem);
  else if (bs == 512 && kr == 1) launch_col<R, REAL_AMP, STAGE, 512, 1, 16>(sec, a, grid, smem);
  else launch_col<R, REAL_AMP, STAGE, 1024, 1, 16>(sec, a, grid, smem);
}

template <typename R, bool REAL_AMP, int BS, int KR>
static void launch_buf(Sector &sec, const HxvArgs &a, dim3 grid, size_t smem) {
  static std::once_flag attr_once;
  auto fn = k_hxv_buf<R, REAL_AMP, BS, KR>;
  std::call_once(attr_once, [&]() {
    VCA_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBig));
    // the hop tables hold LDS addresses relative to 0: the dynamic LDS
    // must start there, i.e. the kernel has no static LDS
    hipFuncAttributes fa{};
    VCA_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(fn)));
    if (fa.sharedSizeBytes != 0)
      fail(VCA_ERR_HIP, "k_hxv_buf has %zu bytes of static LDS", (size_t)fa.sharedSizeBytes);
  });
  hipLaunchKernelGGL(fn, grid, dim3(BS), smem, sec.stream, a);
}

template <typename R, bool REAL_AMP>
static void launch_buf_s(Sector &sec, const HxvArgs &a, const HxvShape &h, dim3 grid,
                         size_t smem) {
  const int bs = h.bs, kr = h.kr;
  if (bs == 256 && kr == 4) launch_buf<R, REAL_AMP, 256, 4>(sec, a, grid, smem);
  else if (bs == 256 && kr == 2) launch_buf<R, REAL_AMP, 256, 2>(sec, a, grid, smem);
  else if (bs == 512 && kr == 2) launch_buf<R, REAL_AMP, 512, 2>(sec, a, grid, smem);
  else if (bs == 512 && kr == 1) launch_buf<R, REAL_AMP, 512, 1>(sec, a, grid, smem);
  else if (bs == 1024 && kr == 1) launch_buf<R, REAL_AMP, 1024, 1>(sec, a, grid, smem);
  else launch_buf<R, REAL_AMP, 1024, 2>(sec, a, grid, smem);
}

template <typename R, bool REAL_AMP, int BS, int NR, bool SGN = false>
static void launch_ov(Sector &sec, const HxvArgs &a, dim3 grid, size_t smem) {
  static std::once_flag attr_once;
  auto fn = k_hxv_ov<R, REAL_AMP, BS, NR, SGN>;
  std::call_once(attr_once, [&]() {
    VCA_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBig));
    hipFuncAttributes fa{};  // LDS addresses from 0: no static LDS
    VCA_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(fn)));
    if (fa.sharedSizeBytes != 0)
      fail(VCA_ERR_HIP, "k_hxv_ov has %zu bytes of static LDS", (size_t)fa.sharedSizeBytes);
  });
  hipLaunchKernelGGL(fn, grid, dim3(BS), smem, sec.stream, a);
}

template <typename R, bool REAL_AMP>
static void launch_ov_s(Sector &sec, const HxvArgs &a, const HxvShape &h, dim3 grid) {
  if constexpr (REAL_AMP) {
    if (h.sgn) {
      HxvArgs b = a;
      b.ent4_up = sec.up.ent4s;
      const size_t ss = (size_t)sgn_lds(a.dimup, (int)sizeof(cplx<R>)).total;
      if (h.bs == 1024) launch_ov<R, true, 1024, 1, true>(sec, b, grid, ss);
      else if (h.bs == 256) launch_ov<R, true, 256, 2, true>(sec, b, grid, ss);
      else launch_ov<R, true, 512, 2, true>(sec, b, grid, ss);
      return;
    }
  }
  const size_t smem =
      (size_t)buf_lds(a.namp_up, a.dimup, 0, (int)sizeof(cplx<R>)).total;
  if (h.bs == 1024) launch_ov<R, REAL_AMP, 1024, 1>(sec, a, grid, smem);
  else if (h.bs == 256) launch_ov<R, REAL_AMP, 256, 2>(sec, a, grid, smem);
  else launch_ov<R, REAL_AMP, 512, 2>(sec, a, grid, smem);
}

template <typename R, bool REAL_AMP>
static void launch_col_t(Sector &sec, const HxvArgs &a, bool stage, bool buf,
                         const HxvShape &h, dim3 grid, size_t smem) {
  if (buf) launch_buf_s<R, REAL_AMP>(sec, a, h, grid, smem);
  else if (stage) launch_col_s<R, REAL_AMP, true>(sec, a, h, grid, smem);
  else launch_col_s<R, REAL_AMP, false>(sec, a, h, grid, smem);
}

template <typename R, bool REAL_AMP>
static void launch_hxv_t(Sector &sec, HxvArgs &a, bool stage, bool onecol, dim3 grid,
                         size_t smem) {
  const hipStream_t st = sec.stream;
  if (stage) {
    if (onecol)
      hipLaunchKernelGGL((k_hxv<R, REAL_AMP, true, true>), grid, dim3(kBlock), smem, st, a);
    else
      hipLaunchKernelGGL((k_hxv<R, REAL_AMP, true, false>), grid, dim3(kBlock), smem, st, a);
  } else {
    if (onecol)
      hipLaunchKernelGGL((k_hxv<R, REAL_AMP, false, true>), grid, dim3(kBlock), smem, st, a);
    else
      hipLaunchKernelGGL((k_hxv<R, REAL_AMP, false, false>), grid, dim3(kBlock), smem, st, a);
  }
}

// Nontemporal loads and stores for the batched update and the overlapped
// H x v's output from 256 MB of vectors per launch (the MALL's size) up:
// nothing is found there again.  (The transposed form's transposes and
// fused update ran slower with them: bath sector real H x v 8.21 -> 8.35 ms.)
// VCA_NT_STREAMS=0/1 forces the choice.
static bool nt_streams(const Sector &sec, int nb) {
  const char *e = std::getenv("VCA_NT_STREAMS");
  if (e && *e) return e[0] == '1';
  return (size_t)nb * sec.vec_bytes() >= ((size_t)256 << 20);
}

template <typename E>
static void launch_transpose(Sector &sec, const TrArgs &t, bool add, int nb) {
  constexpr int T = kTile;
  const dim3 g((unsigned)(((t.ni + T - 1) / T) * ((t.nj + T - 1) / T)), 1, (unsigned)nb);
  if (add) hipLaunchKernelGGL((k_transpose<E, true>), g, dim3(kBlock), 0, sec.stream, t);
  else hipLaunchKernelGGL((k_transpose<E, false>), g, dim3(kBlock), 0, sec.stream, t);
  VCA_HIP(hipGetLastError());
}

template <typename R, int NPL, int BS>
static void launch_plane(Sector &sec, const HxvArgs &a, dim3 grid, size_t smem) {
  static std::once_flag attr_once;
  auto fn = k_hxv_plane<R, NPL, BS>;
  std::call_once(attr_once, [&]() {
    VCA_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBig));
    hipFuncAttributes fa{};  // LDS addresses from 0: no static LDS
    VCA_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(fn)));
    if (fa.sharedSizeBytes != 0)
      fail(VCA_ERR_HIP, "k_hxv_plane has %zu bytes of static LDS", (size_t)fa.sharedSizeBytes);
  });
  hipLaunchKernelGGL(fn, grid, dim3(BS), smem, sec.stream, a);
}

// One gather pass of the transposed form: k_hxv_buf on whole complex
// columns, (h.plane) k_hxv_plane on the two real planes of the columns, or
// (h.cpair) k_hxv_plane2<CPLX> on whole complex columns through the
// column-pair table.
template <typename R>
static void launch_buf_pass(Sector &sec, const HxvArgs &a, const HxvShape &h, bool real_amp,
                            dim3 grid) {
  if (h.cpair) {
    constexpr int rsz = (int)sizeof(R);
    const size_t smem = (size_t)buf_lds(a.namp_up, 0, 0, rsz).col +
                        2 * (size_t)align16(a.dimup * rsz);
    if (h.bs == 1024) launch_plane2<R, 1024, true>(sec, a, grid, smem);
    else launch_plane2<R, 512, true>(sec, a, grid, smem);
  } else if (h.plane) {
    const size_t smem = (size_t)buf_lds(a.namp_up, a.dimup, 0, (int)sizeof(R)).total;
    if (h.bs == 1024) launch_plane<R, 2, 1024>(sec, a, grid, smem);
    else launch_plane<R, 2, 512>(sec, a, grid, smem);
  } else {
    const size_t smem = (size_t)buf_lds(a.namp_up, a.dimup, 0, (int)sizeof(cplx<R>)).total;
    if (real_amp) launch_buf_s<R, true>(sec, a, h, grid, smem);
    else launch_buf_s<R, false>(sec, a, h, grid, smem);
  }
  VCA_HIP(hipGetLastError());
}

// The four launches of the transposed form (see k_transpose) on nb chains
// stacked [nb][dim]; the scratch vectors [nb][dim] are at tb (tb, or
// the caller's scratch).  The one-column passes run on the 2-D grid
// (columns, chains).
template <typename R>
static void launch_tform_t(Sector &sec, HxvArgs a, const HxvShape &h, bool real_amp, int nb,
                           void *tb) {
  const bool fp64 = sec.precision == VCA_FP64;
  a.xcd_chains = 0;
  a.nchains = nb;
  a.skip_dw = 1;
  a.dw_maxnnz = 0;
  // 1. Hv = diag v + H_up v (columns of v staged in LDS)
  a.skip_diag = 0;
  a.ncb = sec.dw.dim;
  if (h.plane) a.ent4_up = sec.up.ent4p;
  if (h.cpair) a.ent4_up = sec.up.ent4q;
  launch_buf_pass<R>(sec, a, h, real_amp, dim3(sec.dw.dim, nb));
  // 2. vt = v^T
  TrArgs t{};
  t.in0 = a.v0;
  t.in1 = a.v1;
  t.out = tb;
  t.ni = sec.up.dim;
  t.nj = sec.dw.dim;
  t.vstride = sec.dim;
  t.lz = a.lz;
  launch_transpose<cplx<R>>(sec, t, false, nb);
  // 3. vt = (H_dw v)^T: the down-spin tables as the gather tables, rows of
  //    vt staged in LDS and overwritten by their result
  HxvArgs b = a;
  b.v0 = b.v1 = tb;
  b.hv = tb;
  b.dimup = sec.dw.dim;
  b.dimdw = sec.up.dim;
  b.map_up = sec.dw.map;
  b.e_up = sec.dw.energy;
  b.cnt_up = sec.dw.cnt;
  b.ent_up = sec.dw.ent;
  b.meta_up = sec.dw.meta;
  b.ent4_up = h.cpair ? sec.dw.ent4q : h.plane ? sec.dw.ent4p : sec.dw.ent4;
  b.up_nc = (sec.dw.maxnnz + 3) / 4;
  b.amp_up = fp64 ? (const void *)sec.dw.amp64 : (const void *)sec.dw.amp32;
  b.namp_up = 2 * sec.dw.nterms;
  b.skip_diag = 1;
  b.ncb = sec.up.dim;
  b.pbase = sec.dw.dim;
  launch_buf_pass<R>(sec, b, h, real_amp, dim3(sec.up.dim, nb));
  // 4. Hv += vt^T; in Lanczos mode fused into the vector update
  //    (k_tadd_update, launch_lanczos_update), so Hv holds diag v + H_up v
  if (!a.lz) {
    TrArgs t2{};
    t2.in0 = t2.in1 = tb;
    t2.out = a.hv;
    t2.ni = sec.dw.dim;
    t2.nj = sec.up.dim;
    t2.vstride = sec.dim;
    t2.lz = a.lz;
    launch_transpose<cplx<R>>(sec, t2, true, nb);
  }
  if (a.lz) {
    hipLaunchKernelGGL(k_lz_alpha, dim3(nb), dim3(kBlock), 0, sec.stream, sec.lz, sec.partials,
                       (int)a.pstride);
    VCA_HIP(hipGetLastError());
  }
}

template <typename R, int BS = 1024, bool CPLX = false>
static void launch_plane2(Sector &sec, const HxvArgs &a, dim3 grid, size_t smem) {
  static std::once_flag attr_once;
  auto fn = k_hxv_plane2<R, BS, CPLX>;
  std::call_once(attr_once, [&]() {
    VCA_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBig));
    hipFuncAttributes fa{};  // LDS addresses from 0: no static LDS
    VCA_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(fn)));
    if (fa.sharedSizeBytes != 0)
      fail(VCA_ERR_HIP, "k_hxv_plane2 has %zu bytes of static LDS", (size_t)fa.sharedSizeBytes);
  });
  hipLaunchKernelGGL(fn, grid, dim3(BS), smem, sec.stream, a);
}

// One real-vector gather pass on grid workgroups per chain: one column per
// workgroup (k_hxv_plane, NPL = 1) or two (k_hxv_plane2).
template <typename R>
static void launch_real_pass(Sector &sec, HxvArgs a, const DevSpin &sp, bool two, int grid,
                             int nb) {
  constexpr int rsz = (int)sizeof(R);
  a.ent4_up = two ? sp.ent4q : sp.ent4p;
  if (two) {
    if (!sp.ent4q) fail(VCA_ERR_STATE, "no column-pair hop table");
    const size_t smem = (size_t)buf_lds(a.namp_up, 0, 0, rsz).col +
                        2 * (size_t)align16(a.dimup * rsz);
    launch_plane2<R>(sec, a, dim3(grid, nb), smem);
  } else {
    launch_plane<R, 1, 1024>(sec, a, dim3(grid, nb),
                             (size_t)buf_lds(a.namp_up, a.dimup, 0, rsz).total);
  }
  VCA_HIP(hipGetLastError());
}

template <typename R, int BS, int KR>
static void launch_buf_real_k(Sector &sec, const HxvArgs &a, dim3 grid, size_t smem) {
  static std::once_flag attr_once;
  auto fn = k_hxv_buf_real<R, BS, KR>;
  std::call_once(attr_once, [&]() {
    VCA_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBig));
    hipFuncAttributes fa{};  // LDS addresses from 0: no static LDS
    VCA_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(fn)));
    if (fa.sharedSizeBytes != 0)
      fail(VCA_ERR_HIP, "k_hxv_buf_real has %zu bytes of static LDS", (size_t)fa.sharedSizeBytes);
  });
  hipLaunchKernelGGL(fn, grid, dim3(BS), smem, sec.stream, a);
}

template <typename R, int BS, int NR, int P>
static void launch_ov_real_k(Sector &sec, const HxvArgs &a, dim3 grid, size_t smem) {
  static std::once_flag attr_once;
  auto fn = k_hxv_ov_real<R, BS, NR, P>;
  std::call_once(attr_once, [&]() {
    VCA_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBig));
    hipFuncAttributes fa{};  // LDS addresses from 0: no static LDS
    VCA_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(fn)));
    if (fa.sharedSizeBytes != 0)
      fail(VCA_ERR_HIP, "k_hxv_ov_real has %zu bytes of static LDS", (size_t)fa.sharedSizeBytes);
  });
  hipLaunchKernelGGL(fn, grid, dim3(BS), smem, sec.stream, a);
}

// Launch shapes of the real-vector direct form: (BS, rows or slots per
// lane, kind): kind 0 k_hxv_buf_real (KR rows), 1 k_hxv_buf_real2 (adjacent
// row pairs), 2 and 3 k_hxv_ov_real with single rows and with adjacent row
// pairs (DimUp <= BS x slots x rows per slot).  Pairs need an even DimUp.
static const int kRealVariants[][3] = {
    {256, 4, 0}, {256, 2, 0}, {512, 2, 0}, {512, 1, 0}, {1024, 1, 0}, {1024, 2, 0},
    {256, 1, 1}, {512, 1, 1}, {1024, 1, 1}, {512, 2, 2}, {1024, 1, 2}, {256, 4, 2},
    {512, 1, 3}, {256, 2, 3}, {1024, 1, 3}, {128, 8, 2}, {128, 4, 3}};
constexpr int kNumRealVariants = sizeof(kRealVariants) / sizeof(kRealVariants[0]);

template <typename R, int BS>
static void launch_buf_real_k2(Sector &sec, const HxvArgs &a, dim3 grid, size_t smem) {
  static std::once_flag attr_once;
  auto fn = k_hxv_buf_real2<R, BS>;
  std::call_once(attr_once, [&]() {
    VCA_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBig));
    hipFuncAttributes fa{};  // LDS addresses from 0: no static LDS
    VCA_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(fn)));
    if (fa.sharedSizeBytes != 0)
      fail(VCA_ERR_HIP, "k_hxv_buf_real2 has %zu bytes of static LDS",
           (size_t)fa.sharedSizeBytes);
  });
  hipLaunchKernelGGL(fn, grid, dim3(BS), smem, sec.stream, a);
}

// Real-vector mode, direct form: k_hxv_buf_real in the launch shape of the
// sector's tuned buffer-load variant (512 lanes x 1 row otherwise), chains
// spread or chain-affine as in the complex form.
template <typename R>
static void launch_buf_real(Sector &sec, HxvArgs a, const HxvShape &h, int nb) {
  a.ent4_up = sec.up.ent4p;
  a.ncb = sec.dw.dim;
  a.nchains = nb;
  const char *xc = std::getenv("VCA_XCD_CHAINS");
  const bool xcd = xc && *xc ? (xc[0] == '1' && nb > 1) : nb >= kXcdChainsMin;
  a.xcd_chains = xcd ? 1 : 0;
  const dim3 g3 =
      xcd ? dim3((unsigned)(sec.dw.dim * ((nb + 7) / 8 * 8)), 1) : dim3(sec.dw.dim, nb);
  const size_t smem =
      (size_t)buf_lds(a.namp_up, sec.up.dim, a.dw_maxnnz, (int)sizeof(R)).total;
  // launch shape: the tuned one (tune_real_direct) or, untuned, the complex
  // form's BS x KR, with adjacent row pairs (k_hxv_buf_real2: one 16-byte
  // load per down-spin hop for both rows) in FP64 when DimUp is even
  int bs = h.buf ? h.bs : 512, kr = h.buf ? h.kr : 1;
  bool pairs = sizeof(R) == 8 && (sec.up.dim & 1) == 0;
  int kind = pairs ? 1 : 0;
  if (sec.real_variant >= 0) {
    bs = kRealVariants[sec.real_variant][0];
    kr = kRealVariants[sec.real_variant][1];
    kind = kRealVariants[sec.real_variant][2];
    pairs = kind == 1;
  }
  if (kind >= 2) {
    const int id = bs * 100 + kr * 10 + (kind - 2);
    switch (id) {
      case 51220: launch_ov_real_k<R, 512, 2, 1>(sec, a, g3, smem); break;
      case 102410: launch_ov_real_k<R, 1024, 1, 1>(sec, a, g3, smem); break;
      case 25640: launch_ov_real_k<R, 256, 4, 1>(sec, a, g3, smem); break;
      case 51211: launch_ov_real_k<R, 512, 1, 2>(sec, a, g3, smem); break;
      case 25621: launch_ov_real_k<R, 256, 2, 2>(sec, a, g3, smem); break;
      case 102411: launch_ov_real_k<R, 1024, 1, 2>(sec, a, g3, smem); break;
      case 12880: launch_ov_real_k<R, 128, 8, 1>(sec, a, g3, smem); break;
      case 12841: launch_ov_real_k<R, 128, 4, 2>(sec, a, g3, smem); break;
      default: fail(VCA_ERR_ARG, "no real overlapped shape %d x %d", bs, kr);
    }
  } else if (pairs) {
    if (bs >= 1024) launch_buf_real_k2<R, 1024>(sec, a, g3, smem);
    else if (bs == 256) launch_buf_real_k2<R, 256>(sec, a, g3, smem);
    else launch_buf_real_k2<R, 512>(sec, a, g3, smem);
  } else if (bs == 256 && kr == 4) launch_buf_real_k<R, 256, 4>(sec, a, g3, smem);
  else if (bs == 256 && kr == 2) launch_buf_real_k<R, 256, 2>(sec, a, g3, smem);
  else if (bs == 512 && kr == 2) launch_buf_real_k<R, 512, 2>(sec, a, g3, smem);
  else if (bs == 1024 && kr == 1) launch_buf_real_k<R, 1024, 1>(sec, a, g3, smem);
  else if (bs == 1024 && kr == 2) launch_buf_real_k<R, 1024, 2>(sec, a, g3, smem);
  else launch_buf_real_k<R, 512, 1>(sec, a, g3, smem);
  VCA_HIP(hipGetLastError());
  if (a.lz) {
    hipLaunchKernelGGL(k_lz_alpha, dim3(nb), dim3(kBlock), 0, sec.stream, sec.lz, sec.partials,
                       (int)a.pstride);
    VCA_HIP(hipGetLastError());
  }
}

// Real-vector mode: the four launches of the transposed form on one real
// plane (real transposes), 1024-lane gather passes.
template <typename R>
static void launch_real_t(Sector &sec, HxvArgs a, int nb) {
  a.xcd_chains = 0;
  a.nchains = nb;
  a.skip_dw = 1;
  a.dw_maxnnz = 0;
  a.skip_diag = 0;
  a.ncb = sec.dw.dim;
  a.ent4_up = sec.up.ent4p;
  launch_real_pass<R>(sec, a, sec.up, real_pass_two(sec, sec.up), real_grid_up(sec), nb);
  TrArgs t{};
  t.in0 = a.v0;
  t.in1 = a.v1;
  t.out = sec.tbuf;
  t.ni = sec.up.dim;
  t.nj = sec.dw.dim;
  t.vstride = a.vstride;
  t.lz = a.lz;
  launch_transpose<R>(sec, t, false, nb);
  HxvArgs b = a;
  b.v0 = b.v1 = sec.tbuf;
  b.hv = sec.tbuf;
  b.dimup = sec.dw.dim;
  b.dimdw = sec.up.dim;
  b.meta_up = sec.dw.meta;
  b.ent4_up = sec.dw.ent4p;
  b.up_nc = (sec.dw.maxnnz + 3) / 4;
  b.amp_up = sec.precision == VCA_FP64 ? (const void *)sec.dw.amp64 : (const void *)sec.dw.amp32;
  b.namp_up = 2 * sec.dw.nterms;
  b.skip_diag = 1;
  b.ncb = sec.up.dim;
  b.pbase = real_grid_up(sec);
  launch_real_pass<R>(sec, b, sec.dw, real_pass_two(sec, sec.dw), real_grid_dw(sec), nb);
  // Hv += (.)^T: in Lanczos mode fused into the vector update
  // (k_tadd_update, launch_lanczos_update), so Hv_up is all H x v writes
  if (!a.lz) {
    TrArgs t2{};
    t2.in0 = t2.in1 = sec.tbuf;
    t2.out = a.hv;
    t2.ni = sec.dw.dim;
    t2.nj = sec.up.dim;
    t2.vstride = a.vstride;
    t2.lz = a.lz;
    launch_transpose<R>(sec, t2, true, nb);
  }
  if (a.lz) {
    hipLaunchKernelGGL(k_lz_alpha, dim3(nb), dim3(kBlock), 0, sec.stream, sec.lz, sec.partials,
                       (int)a.pstride);
    VCA_HIP(hipGetLastError());
  }
}

// scratch: nb vectors of the caller's for v^T, or null for sec.tbuf
// (allocated here on first use).
static void launch_hxv_tform(Sector &sec, const HxvArgs &a, const HxvShape &h, bool real_amp,
                             int nb, void *scratch) {
  if (!scratch) {
    ensure_tbuf(sec, nb);
    scratch = sec.tbuf;
  }
  if (sec.precision == VCA_FP64) launch_tform_t<double>(sec, a, h, real_amp, nb, scratch);
  else launch_tform_t<float>(sec, a, h, real_amp, nb, scratch);
}

void launch_hxv(Sector &sec, const void *v0, const void *v1, void *hv, bool lanczos, int nb,
                void *scratch) {
  HxvArgs a{};
  a.vstride = sec.dim;
  a.v0 = v0;
  a.v1 = v1;
  a.hv = hv;
  a.nt_store = nt_streams(sec, nb) ? 1 : 0;
  a.dimup = sec.up.dim;
  a.dimdw = sec.dw.dim;
  a.map_up = sec.up.map;
  a.map_dw = sec.dw.map;
  a.e_up = sec.up.energy;
  a.e_dw = sec.dw.energy;
  a.cnt_up = sec.up.cnt;
  a.cnt_dw = sec.dw.cnt;
  a.ent_up = sec.up.ent;
  a.ent_dw = sec.dw.ent;
  const bool fp64 = sec.precision == VCA_FP64;
  a.amp_up = fp64 ? (const void *)sec.up.amp64 : (const void *)sec.up.amp32;
  a.amp_dw = fp64 ? (const void *)sec.dw.amp64 : (const void *)sec.dw.amp32;
  a.namp_up = 2 * sec.up.nterms;
  a.namp_dw = 2 * sec.dw.nterms;
  a.cross = sector_cross(sec);
  a.lz = lanczos ? sec.lz : nullptr;
  const size_t esz = fp64 ? sizeof(double2) : sizeof(float2);
  const HxvShape shp = hxv_shape(sec);
  const int cpb = shp.cpb;
  a.cpb = cpb;
  const int grid = shp.grid;
  a.dw_maxnnz = sec.dw.maxnnz;
  const bool real_amp = sec.up.real_amp && sec.dw.real_amp;
  if (sec.real_vec) {
    if (!real_mode_ok(sec))
      fail(VCA_ERR_STATE, "real-vector H x v on a sector that does not fit it");
    // chains sit 2*dim reals apart (the complex layout's stride), so the
    // element-wise kernels see each chain as ceil(dim/2) complex pairs
    a.vstride = 2 * sec.dim;
    if (lanczos) ensure_partials(sec, hxv_partials(sec, shp) * nb);
    a.partials = sec.partials;
    a.meta_up = sec.up.meta;
    a.up_nc = (sec.up.maxnnz + 3) / 4;
    a.dw_maxnnz = sec.dw.maxnnz;
    a.pbase = 0;
    a.pstride = hxv_partials(sec, shp);
    if (real_uses_tform(sec)) {
      ensure_tbuf(sec, nb);
      if (fp64) launch_real_t<double>(sec, a, nb);
      else launch_real_t<float>(sec, a, nb);
    } else {
      if (fp64) launch_buf_real<double>(sec, a, shp, nb);
      else launch_buf_real<float>(sec, a, shp, nb);
    }
    return;
  }
  if (shp.onecol_kernel) {
    if (lanczos) ensure_partials(sec, hxv_partials(sec, shp) * nb);
    a.partials = sec.partials;
    a.meta_up = sec.up.meta;
    a.ent4_up = sec.up.ent4;
    a.up_nc = (sec.up.maxnnz + 3) / 4;
    a.pbase = 0;
    a.pstride = hxv_partials(sec, shp);
    if (shp.tform) {
      launch_hxv_tform(sec, a, shp, real_amp, nb, scratch);
      return;
    }
    const size_t fixed = (size_t)a.namp_up * esz + (size_t)a.dw_maxnnz * (esz + 8) + 48;
    const size_t col_bytes = (size_t)a.dimup * esz;
    const size_t budget = kLdsBig;
    const bool stage = fixed + col_bytes <= budget;
    // buffer-load form: its own LDS layout (buf_lds) and hop entries as LDS
    // addresses; buffer loads address one chain's vector with 32-bit offsets
    const size_t buf_smem = (size_t)buf_lds(a.namp_up, a.dimup, a.dw_maxnnz, (int)esz).total;
    const bool buf = shp.buf && sec.up.lds_form && buf_smem <= budget &&
                     (uint64_t)sec.dim * esz < (1ull << 31);
    const size_t smem = buf ? buf_smem : fixed + (stage ? col_bytes : 0);
    // chain-affine batches from kXcdChainsMin chains up (fewer would leave
    // XCDs idle); VCA_XCD_CHAINS=0/1 forces the choice
    const char *xc = std::getenv("VCA_XCD_CHAINS");
    const bool xcd = xc && *xc ? (xc[0] == '1' && nb > 1) : nb >= kXcdChainsMin;
    a.xcd_chains = xcd ? 1 : 0;
    a.ncb = grid;
    a.nchains = nb;
    const dim3 g3 = xcd ? dim3((unsigned)(grid * ((nb + 7) / 8 * 8)), 1) : dim3(grid, nb);
    if (shp.ov) {
      if (fp64) {
        if (real_amp) launch_ov_s<double, true>(sec, a, shp, g3);
        else launch_ov_s<double, false>(sec, a, shp, g3);
      } else {
        if (real_amp) launch_ov_s<float, true>(sec, a, shp, g3);
        else launch_ov_s<float, false>(sec, a, shp, g3);
      }
    } else if (fp64) {
      if (real_amp) launch_col_t<double, true>(sec, a, stage, buf, shp, g3, smem);
      else launch_col_t<double, false>(sec, a, stage, buf, shp, g3, smem);
    } else {
      if (real_amp) launch_col_t<float, true>(sec, a, stage, buf, shp, g3, smem);
      else launch_col_t<float, false>(sec, a, stage, buf, shp, g3, smem);
    }
    VCA_HIP(hipGetLastError());
    if (lanczos) {
      hipLaunchKernelGGL(k_lz_alpha, dim3(nb), dim3(kBlock), 0, sec.stream, sec.lz,
                         sec.partials, grid);
      VCA_HIP(hipGetLastError());
    }
    return;
  }
  const size_t amp_bytes = (size_t)(a.namp_up + a.namp_dw) * esz;
  const size_t col_bytes = (size_t)cpb * a.dimup * esz;
  const bool stage = amp_bytes + col_bytes <= kLdsBudget;
  const size_t smem = amp_bytes + (stage ? col_bytes : 0);
  if (lanczos) ensure_partials(sec, grid * nb);
  a.partials = sec.partials;
  const bool onecol = cpb == 1;
  const dim3 g3(grid, nb);
  if (fp64) {
    if (real_amp) launch_hxv_t<double, true>(sec, a, stage, onecol, g3, smem);
    else launch_hxv_t<double, false>(sec, a, stage, onecol, g3, smem);
  } else {
    if (real_amp) launch_hxv_t<float, true>(sec, a, stage, onecol, g3, smem);
    else launch_hxv_t<float, false>(sec, a, stage, onecol, g3, smem);
  }
  VCA_HIP(hipGetLastError());
  if (lanczos) {
    hipLaunchKernelGGL(k_lz_alpha, dim3(nb), dim3(kBlock), 0, sec.stream, sec.lz, sec.partials,
                       grid);
    VCA_HIP(hipGetLastError());
  }
}

void launch_lanczos_update(Sector &sec, bool accumulate_x, int nb) {
  if (fused_tadd(sec)) {
    TUpdArgs t{};
    t.u0 = sec.buf[0];
    t.u1 = sec.buf[1];
    t.hv = sec.buf[2];
    t.wt = sec.tbuf;
    t.x = accumulate_x ? sec.xvec : nullptr;
    t.dimup = sec.up.dim;
    t.dimdw = sec.dw.dim;
    t.vstride = sec.real_vec ? 2 * sec.dim : sec.dim;  // chains 2 dim reals apart
    t.lz = sec.lz;
    const int nblk = tadd_tiles(sec);
    ensure_partials(sec, nblk * nb);
    t.partials = sec.partials;
    const dim3 g((unsigned)nblk, 1, (unsigned)nb);
    const bool fp64 = sec.precision == VCA_FP64;
    if (sec.real_vec && fp64)
      hipLaunchKernelGGL((k_tadd_update<double, double>), g, dim3(kBlock), 0, sec.stream, t);
    else if (sec.real_vec)
      hipLaunchKernelGGL((k_tadd_update<float, float>), g, dim3(kBlock), 0, sec.stream, t);
    else if (fp64)
      hipLaunchKernelGGL((k_tadd_update<double, double2>), g, dim3(kBlock), 0, sec.stream, t);
    else
      hipLaunchKernelGGL((k_tadd_update<float, float2>), g, dim3(kBlock), 0, sec.stream, t);
    VCA_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_lz_beta, dim3(nb), dim3(kBlock), 0, sec.stream, sec.lz, sec.partials,
                       nblk);
    VCA_HIP(hipGetLastError());
    return;
  }
  UpdArgs a{};
  a.vstride = sec.dim;
  a.u0 = sec.buf[0];
  a.u1 = sec.buf[1];
  a.hv = sec.buf[2];
  a.x = accumulate_x ? sec.xvec : nullptr;
  a.n = vec_elems(sec);
  a.lz = sec.lz;
  const int grid = grid_for(a.n);
  ensure_partials(sec, grid * nb);
  a.partials = sec.partials;
  if (nt_streams(sec, nb)) {
    if (sec.precision == VCA_FP64)
      hipLaunchKernelGGL((k_lanczos_update<double, true>), dim3(grid, nb), dim3(kBlock), 0,
                         sec.stream, a);
    else
      hipLaunchKernelGGL((k_lanczos_update<float, true>), dim3(grid, nb), dim3(kBlock), 0,
                         sec.stream, a);
  } else if (sec.precision == VCA_FP64)
    hipLaunchKernelGGL(k_lanczos_update<double>, dim3(grid, nb), dim3(kBlock), 0, sec.stream, a);
  else
    hipLaunchKernelGGL(k_lanczos_update<float>, dim3(grid, nb), dim3(kBlock), 0, sec.stream, a);
  VCA_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_lz_beta, dim3(nb), dim3(kBlock), 0, sec.stream, sec.lz, sec.partials,
                     grid);
  VCA_HIP(hipGetLastError());
}

void launch_start_vector(Sector &sec, uint64_t seed) {
  const int64_t n = vec_elems(sec);
  const int grid = grid_for(n);
  if (sec.precision == VCA_FP64)
    hipLaunchKernelGGL(k_start<double>, dim3(grid), dim3(kBlock), 0, sec.stream, sec.buf[0], n,
                       seed);
  else
    hipLaunchKernelGGL(k_start<float>, dim3(grid), dim3(kBlock), 0, sec.stream, sec.buf[0], n,
                       seed);
  VCA_HIP(hipGetLastError());
}

void launch_norm0(Sector &sec) {
  const int64_t n = vec_elems(sec);
  const int grid = grid_for(n);
  ensure_partials(sec, grid);
  if (sec.precision == VCA_FP64)
    hipLaunchKernelGGL(k_norm0<double>, dim3(grid), dim3(kBlock), 0, sec.stream, sec.buf[0], n,
                       sec.lz, sec.partials);
  else
    hipLaunchKernelGGL(k_norm0<float>, dim3(grid), dim3(kBlock), 0, sec.stream, sec.buf[0], n,
                       sec.lz, sec.partials);
  VCA_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_lz_start, dim3(1), dim3(kBlock), 0, sec.stream, sec.lz, sec.partials,
                     grid);
  VCA_HIP(hipGetLastError());
}

void launch_scale_x(Sector &sec) {
  const int64_t n = vec_elems(sec);
  const int grid = grid_for(n);
  ensure_partials(sec, grid);
  // the norm lands in lz->y[0] (y is no longer needed after pass 2)
  const double *den = sec.lzh.y;
  if (sec.precision == VCA_FP64) {
    hipLaunchKernelGGL(k_normalize<double>, dim3(grid), dim3(kBlock), 0, sec.stream, sec.xvec,
                       n, sec.lz, sec.partials);
    hipLaunchKernelGGL(k_lz_xnorm, dim3(1), dim3(kBlock), 0, sec.stream, sec.lz, sec.partials,
                       grid);
    hipLaunchKernelGGL(k_scale_by<double>, dim3(grid), dim3(kBlock), 0, sec.stream, sec.xvec,
                       n, den);
  } else {
    hipLaunchKernelGGL(k_normalize<float>, dim3(grid), dim3(kBlock), 0, sec.stream, sec.xvec,
                       n, sec.lz, sec.partials);
    hipLaunchKernelGGL(k_lz_xnorm, dim3(1), dim3(kBlock), 0, sec.stream, sec.lz, sec.partials,
                       grid);
    hipLaunchKernelGGL(k_scale_by<float>, dim3(grid), dim3(kBlock), 0, sec.stream, sec.xvec,
                       n, den);
  }
  VCA_HIP(hipGetLastError());
}

void launch_expand_real(Sector &sec, const void *src, void *dst) {
  const int grid = grid_for(sec.dim);
  if (sec.precision == VCA_FP64)
    hipLaunchKernelGGL(k_expand_real<double>, dim3(grid), dim3(kBlock), 0, sec.stream,
                       static_cast<const double *>(src), static_cast<double2 *>(dst), sec.dim);
  else
    hipLaunchKernelGGL(k_expand_real<float>, dim3(grid), dim3(kBlock), 0, sec.stream,
                       static_cast<const float *>(src), static_cast<float2 *>(dst), sec.dim);
  VCA_HIP(hipGetLastError());
}

void launch_convert(const void *src, int src_prec, void *dst, int dst_prec, int64_t n,
                    hipStream_t st) {
  const int grid = grid_for(n);
  if (src_prec == VCA_FP64 && dst_prec == VCA_FP64)
    VCA_HIP(hipMemcpyAsync(dst, src, sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
  else if (src_prec == VCA_FP64)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid), dim3(kBlock), 0, st, src, dst, n);
  else if (dst_prec == VCA_FP64)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid), dim3(kBlock), 0, st, src, dst, n);
  else
    VCA_HIP(hipMemcpyAsync(dst, src, sizeof(float2) * n, hipMemcpyDeviceToDevice, st));
  VCA_HIP(hipGetLastError());
}

int excite_grid(int64_t dim) { return grid_for(dim); }

void launch_excite_into(const Sector &jsec, int dimup_i, int dimdw_i, const void *psi, int spin,
                        const ExcOp *d_ops, int nb, void *out, int64_t vstride, bool real,
                        LanczosDev *lz, double *partials, hipStream_t st) {
  ExcArgs a{};
  a.vstride = vstride;
  a.out = out;
  a.psi = psi;
  a.dimup_j = jsec.up.dim;
  a.dimdw_j = jsec.dw.dim;
  a.dimup_i = dimup_i;
  a.dimdw_i = dimdw_i;
  a.spin = spin;
  a.map_j = a.spin == 1 ? jsec.up.map : jsec.dw.map;
  a.binom = solver_binom(jsec.solver);
  a.ops = d_ops;
  a.lz = lz;
  a.partials = partials;
  const int grid = grid_for(jsec.dim);
  if (jsec.precision == VCA_FP64) {
    if (real)
      hipLaunchKernelGGL((k_excite<double, true>), dim3(grid, nb), dim3(kBlock), 0, st, a);
    else
      hipLaunchKernelGGL((k_excite<double, false>), dim3(grid, nb), dim3(kBlock), 0, st, a);
  } else {
    if (real)
      hipLaunchKernelGGL((k_excite<float, true>), dim3(grid, nb), dim3(kBlock), 0, st, a);
    else
      hipLaunchKernelGGL((k_excite<float, false>), dim3(grid, nb), dim3(kBlock), 0, st, a);
  }
  VCA_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_lz_start, dim3(nb), dim3(kBlock), 0, st, lz, partials, grid);
  VCA_HIP(hipGetLastError());
}

void launch_excite(Sector &jsec, int dimup_i, int dimdw_i, const void *psi, int spin,
                   const ExcOp *d_ops, int nb) {
  const int grid = grid_for(jsec.dim);
  ensure_partials(jsec, grid * nb);
  const bool ro = jsec.real_vec;
  // reals between chains in real-vector mode: the complex stride
  launch_excite_into(jsec, dimup_i, dimdw_i, psi, spin, d_ops, nb, jsec.buf[0],
                     ro ? 2 * jsec.dim : jsec.dim, ro, jsec.lz, jsec.partials, jsec.stream);
}

int real_num_variants() { return kNumRealVariants; }

bool real_variant_ok(const Sector &sec, int var) {
  if (var < 0 || var >= kNumRealVariants || !real_mode_ok(sec) || real_uses_tform(sec))
    return false;
  const int kind = kRealVariants[var][2];
  if ((kind == 1 || kind == 3) && (sec.up.dim & 1)) return false;  // row pairs
  if (kind >= 2)  // overlapped: the column in one round of lanes
    return sec.up.dim <= kRealVariants[var][0] * kRealVariants[var][1] * (kind == 3 ? 2 : 1);
  return true;
}

double real_vector_of(const void *psi, int64_t n, void *out, int prec, hipStream_t st) {
  const int grid = grid_for(n);
  double *part = nullptr;
  VCA_HIP(hipMallocAsync((void **)&part, sizeof(double) * 2 * grid, st));
  std::vector<double> h(2 * (size_t)grid);
  auto sums = [&](double &a, double &b) {
    VCA_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * 2 * grid, hipMemcpyDeviceToHost, st));
    VCA_HIP(hipStreamSynchronize(st));
    a = b = 0;
    for (int g = 0; g < grid; ++g) {
      a += h[2 * g];
      b += h[2 * g + 1];
    }
  };
  const double2 *p = static_cast<const double2 *>(psi);
  hipLaunchKernelGGL(k_psi_sq, dim3(grid), dim3(kBlock), 0, st, p, n, part);
  VCA_HIP(hipGetLastError());
  double qr, qi;
  sums(qr, qi);
  const double phi = 0.5 * std::atan2(qi, qr);
  if (prec == VCA_FP64)
    hipLaunchKernelGGL(k_psi_real<double>, dim3(grid), dim3(kBlock), 0, st, p, n, std::cos(phi),
                       std::sin(phi), static_cast<double *>(out), part);
  else
    hipLaunchKernelGGL(k_psi_real<float>, dim3(grid), dim3(kBlock), 0, st, p, n, std::cos(phi),
                       std::sin(phi), static_cast<float *>(out), part);
  VCA_HIP(hipGetLastError());
  double sre, sim;
  sums(sre, sim);
  VCA_HIP(hipFreeAsync(part, st));
  return std::sqrt(sim / std::max(sre + sim, 1e-300));
}

}  // namespace vca
