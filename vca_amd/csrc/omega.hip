// omega.hip -- the Omega integrand on the GPU (VCA_OMEGA.f90:37-75).
//
//   f(z) = (1/Nk) sum_k log|det(1 + (Delta(z) + Hloc - H_k) G'(z))|
//
// in the reference's own form: per frequency, P(z) = (Delta + Hloc) G'(z)
// once, then per (k, z) the contraction M_k = 1 + P(z) - H_k G'(z) and an
// LU of M_k.  (The algebraically equal det(G') det(G'^-1 + A_k) avoids the
// product but cancels catastrophically at large |z|, where both logarithms
// are ~ +-n log|z| and f ~ 1/|z|^2: the QAGI map samples exactly there.)
// When Hloc, H_k and the bath are spin-diagonal the determinant factorises
// over spin blocks of size Nlat*Norb (G' is spin-diagonal in the reference,
// VCA_IO/gf_cluster.f90:23-49).
//
// Kernels:
//   k_pole_sum   G_raw(z, e) = sum_p w_p / (z - p_p) over the Lehmann poles
//                of element e (all states and channels),
//                VCA_IO/gf_cluster.f90:23-49
//   k_prep       recombination G_ij = (G^mix - (1 - i chan4)(G_ii + G_jj))/2
//                (gf_cluster.f90:50-62), Delta(z) = V (z+mu-h)^-1 V^+
//                (VCA_BATH_FUNCTIONS.f90:127-150, Gauss-Jordan in LDS), and
//                P(z) = (Delta + Hloc) G'; one workgroup per (z, spin block)
//   k_logdet     per (k, z): M_k = 1 + P - H_k G' with H_k's rows held in
//                registers across a chunk of frequencies and G'(z), P(z)
//                staged in LDS; LU with partial pivoting, one matrix per
//                16-lane group (row per lane); the pivot product is kept
//                in FP64 as mantissa and exponent, one log per matrix
//   k_reduce_z   deterministic sum of the k-block partials per z
//   k_pole_grid  G_raw on the Matsubara or real-axis grid (grid points made
//                in the kernel), VCA_GF_NORMAL.f90:1024-1031
//   k_grid_sigma per (z, spin block): recombined G, G0^-1 = (z + mu) - Hloc
//                - Delta(z), Sigma = G0^-1 - G^-1, G0, in the 7-index layout
//                of impGmats / impSmats / impG0mats (VCA_GF_NORMAL.f90:
//                1045-1111)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "omega.hpp"

namespace vca {

template <typename R> struct OC2;
template <> struct OC2<double> { using T = double2; };
template <> struct OC2<float> { using T = float2; };
template <typename R> using ocplx = typename OC2<R>::T;

constexpr int kOBlock = 256;
constexpr int kNB = 16;  // padded block size of the LU kernel

// ----------------------------------------------------------- pole sums --
__global__ __launch_bounds__(kOBlock) void k_pole_sum(int nz, const double2 *z, int nelem,
                                                      const int64_t *off, const double2 *w,
                                                      const double *pole, double2 *graw) {
  const int64_t t = (int64_t)blockIdx.x * kOBlock + threadIdx.x;
  if (t >= (int64_t)nz * nelem) return;
  const int e = (int)(t % nelem), iz = (int)(t / nelem);
  const double2 zz = z[iz];
  double gr = 0, gi = 0;
  for (int64_t p = off[e]; p < off[e + 1]; ++p) {
    // w / (z - p)
    const double dr = zz.x - pole[p], di = zz.y;
    const double den = 1.0 / (dr * dr + di * di);
    const double2 ww = w[p];
    gr += (ww.x * dr + ww.y * di) * den;
    gi += (ww.y * dr - ww.x * di) * den;
  }
  graw[t] = make_double2(gr, gi);
}

// -------------------------------------------------------- per-z prep ----
// Gauss-Jordan inverse with partial pivoting of the n x n complex matrix in
// LDS (row-major a[r*ld + c]); returns log|det| (valid in all threads).
__device__ double gj_inverse(int n, double2 *a, double2 *inv, int ld, double *s_logdet,
                             int *s_piv) {
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int r = e / n, c = e % n;
    inv[r * ld + c] = make_double2(r == c ? 1.0 : 0.0, 0.0);
  }
  if (tid == 0) *s_logdet = 0.0;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (tid == 0) {
      int p = k;
      double best = -1.0;
      for (int r = k; r < n; ++r) {
        const double2 v = a[r * ld + k];
        const double m = v.x * v.x + v.y * v.y;
        if (m > best) {
          best = m;
          p = r;
        }
      }
      *s_piv = p;
      *s_logdet += 0.5 * log(best);
    }
    __syncthreads();
    const int p = *s_piv;
    if (p != k)
      for (int c = tid; c < n; c += blockDim.x) {
        double2 t1 = a[k * ld + c];
        a[k * ld + c] = a[p * ld + c];
        a[p * ld + c] = t1;
        double2 t2 = inv[k * ld + c];
        inv[k * ld + c] = inv[p * ld + c];
        inv[p * ld + c] = t2;
      }
    __syncthreads();
    const double2 piv = a[k * ld + k];
    const double dd = 1.0 / (piv.x * piv.x + piv.y * piv.y);
    const double2 pinv = make_double2(piv.x * dd, -piv.y * dd);
    __syncthreads();
    for (int c = tid; c < n; c += blockDim.x) {
      const double2 x = a[k * ld + c], y = inv[k * ld + c];
      a[k * ld + c] = make_double2(x.x * pinv.x - x.y * pinv.y, x.x * pinv.y + x.y * pinv.x);
      inv[k * ld + c] = make_double2(y.x * pinv.x - y.y * pinv.y, y.x * pinv.y + y.y * pinv.x);
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += blockDim.x) {
      const int r = e / n, c = e % n;
      if (r == k) continue;
      const double2 f = a[r * ld + k];
      const double2 x = a[k * ld + c], y = inv[k * ld + c];
      // defer the column-k update of a until all threads read f
      a[r * ld + c] = (c == k) ? a[r * ld + c]
                               : make_double2(a[r * ld + c].x - (f.x * x.x - f.y * x.y),
                                              a[r * ld + c].y - (f.x * x.y + f.y * x.x));
      inv[r * ld + c] = make_double2(inv[r * ld + c].x - (f.x * y.x - f.y * y.y),
                                     inv[r * ld + c].y - (f.x * y.y + f.y * y.x));
    }
    __syncthreads();
    for (int r = tid; r < n; r += blockDim.x)
      if (r != k) a[r * ld + k] = make_double2(0.0, 0.0);
    __syncthreads();
  }
  return *s_logdet;
}

struct PrepArgs {
  int nz, nspin_blk, nb, nlat, norb, nspin_gf;
  double chan4, xmu;
  const double2 *z;
  const double2 *graw;  // [nz][nelem], element order (ispin, ilat, jlat, iorb, jorb)
  int nelem;
  const double2 *hloc_blk;  // [nspin_blk][nb][nb] row-major
  int nbb;                  // bath block size (0: no bath)
  const double2 *hb_blk;    // [nspin_blk][nbb][nbb]
  const double2 *vb_blk;    // [nspin_blk][nb][nbb]
  double2 *gmat;            // out G'(z)  [nz][nspin_blk][16][16], zero padded
  double2 *pmat;            // out P(z) = (Delta + Hloc) G'(z), same layout
};

__global__ __launch_bounds__(64) void k_prep(const PrepArgs a) {
  __shared__ double2 s_g[kNB * kNB], s_inv[kNB * kNB];
  __shared__ double2 s_hb[kNB * kNB], s_hbinv[kNB * kNB];
  __shared__ double s_ld;
  __shared__ int s_piv;
  const int iz = blockIdx.x, sb = blockIdx.y, tid = threadIdx.x;
  const int nb = a.nb, norb = a.norb, nlat = a.nlat;
  const double2 zz = a.z[iz];
  const double2 *g = a.graw + (int64_t)iz * a.nelem;
  // G' of the block, lso order inside the block: i = iorb + ilat*Norb
  // (+ spin offset when one block holds both spins)
  for (int e = tid; e < nb * nb; e += blockDim.x) {
    const int r = e / nb, c = e % nb;
    const int ls = nlat * norb;
    const int sr = (a.nspin_blk == 1 && a.nspin_gf == 2) ? r / ls : sb;
    const int sc = (a.nspin_blk == 1 && a.nspin_gf == 2) ? c / ls : sb;
    double2 v = make_double2(0.0, 0.0);
    if (sr == sc) {
      const int rr = r % ls, cc = c % ls;
      const int il = rr / norb, io = rr % norb, jl = cc / norb, jo = cc % norb;
      auto elem = [&](int i1, int j1, int o1, int o2) {
        return g[(((sr * nlat + i1) * nlat + j1) * norb + o1) * norb + o2];
      };
      v = elem(il, jl, io, jo);
      if (!(il == jl && io == jo)) {
        const double2 gii = elem(il, il, io, io), gjj = elem(jl, jl, jo, jo);
        // (1 - i chan4) (gii + gjj)
        const double sx = gii.x + gjj.x, sy = gii.y + gjj.y;
        const double tx = sx + a.chan4 * sy, ty = sy - a.chan4 * sx;
        v = make_double2(0.5 * (v.x - tx), 0.5 * (v.y - ty));
      }
    }
    s_g[r * kNB + c] = v;
  }
  __syncthreads();
  // s_inv := Hloc (+ Delta)
  const double2 *hl = a.hloc_blk + (int64_t)sb * nb * nb;
  for (int e = tid; e < nb * nb; e += blockDim.x) {
    const int r = e / nb, c = e % nb;
    s_inv[r * kNB + c] = hl[r * nb + c];
  }
  // Delta = V (z + mu - h)^-1 V^+
  if (a.nbb > 0) {
    const int nbb = a.nbb;
    const double2 *hb = a.hb_blk + (int64_t)sb * nbb * nbb;
    for (int e = tid; e < nbb * nbb; e += blockDim.x) {
      const int r = e / nbb, c = e % nbb;
      double2 h = hb[r * nbb + c];
      double2 v = make_double2(-h.x, -h.y);
      if (r == c) {
        v.x += zz.x + a.xmu;
        v.y += zz.y;
      }
      s_hb[r * kNB + c] = v;
    }
    __syncthreads();
    gj_inverse(nbb, s_hb, s_hbinv, kNB, &s_ld, &s_piv);
    const double2 *vb = a.vb_blk + (int64_t)sb * nb * nbb;
    for (int e = tid; e < nb * nb; e += blockDim.x) {
      const int r = e / nb, c = e % nb;
      double dr = 0, di = 0;
      for (int p = 0; p < nbb; ++p) {
        const double2 vr = vb[r * nbb + p];
        for (int q = 0; q < nbb; ++q) {
          const double2 m = s_hbinv[p * kNB + q];
          const double2 vc = vb[c * nbb + q];  // conj(V(c,q)) below
          // vr * m * conj(vc)
          const double tx = vr.x * m.x - vr.y * m.y, ty = vr.x * m.y + vr.y * m.x;
          dr += tx * vc.x + ty * vc.y;
          di += ty * vc.x - tx * vc.y;
        }
      }
      s_inv[r * kNB + c].x += dr;
      s_inv[r * kNB + c].y += di;
    }
  }
  __syncthreads();
  // P = (Hloc + Delta) G'; both outputs zero padded to 16 x 16
  double2 *gout = a.gmat + ((int64_t)iz * a.nspin_blk + sb) * kNB * kNB;
  double2 *pout = a.pmat + ((int64_t)iz * a.nspin_blk + sb) * kNB * kNB;
  for (int e = tid; e < kNB * kNB; e += blockDim.x) {
    const int r = e / kNB, c = e % kNB;
    double2 gv = make_double2(0.0, 0.0), pv = make_double2(0.0, 0.0);
    if (r < nb && c < nb) {
      gv = s_g[r * kNB + c];
      for (int m = 0; m < nb; ++m) {
        const double2 x = s_inv[r * kNB + m], y = s_g[m * kNB + c];
        pv.x += x.x * y.x - x.y * y.y;
        pv.y += x.x * y.y + x.y * y.x;
      }
    }
    gout[e] = gv;
    pout[e] = pv;
  }
}

// ------------------------------------------------------- k x z log-det --
// Pivoted LU of the NBT x NBT matrices held one row per lane (lane gl of
// each NBT-lane group): row r is X(r, :) = (ar, ai) with M(r, :) = X(r, :) +
// e_cu (the identity's unit kept exact in column cu, -1 once eliminated).
// The pivot row is the active row of largest |M(r, kk)| (ties to the lowest
// row), found with an NBT-lane butterfly and broadcast to the group; rows
// used as pivots go inactive, so no swap is needed.  Returns log|det M|;
// with want_phase also the unit phase (phr, phi) of the pivot product and
// the parity of the pivot order's inversions.
template <typename R, int NBT>
__device__ __forceinline__ double lu_logdet_rows(R (&ar)[NBT], R (&ai)[NBT], int cu, int gl,
                                                 int lane, bool want_phase, double &phr,
                                                 double &phi, int &inv_parity) {
  bool active = true;
  // prod |pivot|^2 kept as mantissa * 2^e (frexp after every factor, so
  // no overflow), one log per matrix at the end
  double pm_acc = 1.0;
  int pm_exp = 0;
#pragma unroll
  for (int kk = 0; kk < NBT; ++kk) {
    const R mkr = ar[kk] + (cu == kk ? (R)1 : (R)0);  // M(r, kk)
    float mag = active ? (float)(mkr * mkr + ai[kk] * ai[kk]) : -1.0f;
    int idx = gl;
#pragma unroll
    for (int off = NBT / 2; off > 0; off >>= 1) {
      const float m2 = __shfl_xor(mag, off, NBT);
      const int i2 = __shfl_xor(idx, off, NBT);
      if (m2 > mag || (m2 == mag && i2 < idx)) {
        mag = m2;
        idx = i2;
      }
    }
    const int src = (lane & ~(NBT - 1)) | idx;
    R pr[NBT], pim[NBT];
#pragma unroll
    for (int j = kk; j < NBT; ++j) {
      pr[j] = __shfl(ar[j], src, 64);
      pim[j] = __shfl(ai[j], src, 64);
    }
    const int pcu = __shfl(cu, src, 64);
    // pivot = X(p, kk) + [its unit is in column kk], formed in FP64
    const double pvr = (double)pr[kk] + (pcu == kk ? 1.0 : 0.0);
    const double pvi = (double)pim[kk];
    const double pm = pvr * pvr + pvi * pvi;
    {
      int ex;
      pm_acc = frexp(pm_acc * pm, &ex);
      pm_exp += ex;
    }
    if (want_phase) {
      // rows still active with a lower index than the pivot row are
      // eliminated after it: one inversion each
      const uint64_t below = __ballot(active && gl < idx);
      const int gbase = lane & ~(NBT - 1);
      const uint64_t gm = (below >> gbase) & ((NBT == 64) ? ~0ull : ((1ull << NBT) - 1ull));
      inv_parity ^= __popcll(gm) & 1;
      const double rm = 1.0 / sqrt(pm);
      const double ur = pvr * rm, ui = pvi * rm;
      const double nr = phr * ur - phi * ui;
      phi = phr * ui + phi * ur;
      phr = nr;
    }
    if (gl == idx) active = false;
    if (active) {
      const R pfr = (R)pvr, pfi = pim[kk];
      const R inv = (R)1 / (pfr * pfr + pfi * pfi);
      const R lr = (mkr * pfr + ai[kk] * pfi) * inv;  // l = M(r, kk) / pivot
      const R li = (ai[kk] * pfr - mkr * pfi) * inv;
#pragma unroll
      for (int j = kk + 1; j < NBT; ++j) {
        const R prj = pr[j] + (pcu == j ? (R)1 : (R)0);  // pivot row's unit
        ar[j] -= lr * prj - li * pim[j];
        ai[j] -= lr * pim[j] + li * prj;
      }
      if (cu == kk) cu = -1;  // this row's unit sat in the eliminated column
    }
  }
  // identical in the lanes of a group
  return 0.5 * (log(pm_acc) + (double)pm_exp * 0.69314718055994530942);
}

// Each 16-lane group owns one matrix M = 1 + P(z) - H_k G'(z), lane r
// holding row r: the lane forms its row of H_k G' from its register copy
// of H_k's row and G'(z) read from LDS (the same address for the whole
// wavefront: broadcast reads).  Partial pivoting: the pivot row is the
// active row of largest |M(r,k)| (ties to the lowest row), found with a
// 16-lane butterfly, then broadcast to the group; rows already used as
// pivots go inactive, so no swap is needed and log|det| = sum log|pivot|.
// NBT = 16 or 8: lanes (rows) per matrix; H_k stored [k][block][NBT][NBT],
// G'(z) and P(z) with row stride 16.
// With phase != null the kernel also writes the phase of det(M) of every
// (z, block, k) as a unit complex: the product of pivot/|pivot| times the
// sign of the row permutation (the pivot order; its inversions are counted
// with a ballot of the active rows below the pivot row).
template <typename R, int NBT>
__global__ __launch_bounds__(kOBlock) void k_logdet(int nk, int nz, int zchunk, int nspin_blk,
                                                    const ocplx<R> *hk, const ocplx<R> *gmat,
                                                    const ocplx<R> *pmat, double *partials,
                                                    double2 *phase) {
  using C = ocplx<R>;
  constexpr int MPB = kOBlock / NBT;  // matrices per block
  const int lane = threadIdx.x & 63;
  const int gl = lane & (NBT - 1);
  const int grp = threadIdx.x / NBT;
  const int kb = blockIdx.x;
  const int sb = blockIdx.z;
  const int k = kb * MPB + grp;
  const bool kvalid = k < nk;
  __shared__ double s_red[MPB];
  __shared__ C s_gz[NBT * NBT], s_pz[NBT * NBT];
  R hr[NBT], hi[NBT];
  {
    const C *row = hk + (((int64_t)(kvalid ? k : 0) * nspin_blk + sb) * NBT + gl) * NBT;
#pragma unroll
    for (int j = 0; j < NBT; ++j) {
      const C v = row[j];
      hr[j] = kvalid ? v.x : (R)0;
      hi[j] = kvalid ? v.y : (R)0;
    }
  }
  const int zbeg = blockIdx.y * zchunk;
  const int zend = min(zbeg + zchunk, nz);
  for (int iz = zbeg; iz < zend; ++iz) {
    const int64_t zoff = ((int64_t)iz * nspin_blk + sb) * kNB * kNB;
    __syncthreads();
    for (int e = threadIdx.x; e < NBT * NBT; e += kOBlock) {
      const int r = e / NBT, c = e % NBT;
      s_gz[e] = gmat[zoff + r * kNB + c];
      s_pz[e] = pmat[zoff + r * kNB + c];
    }
    __syncthreads();
    // The row is held as X(r, :) with M(r, :) = X(r, :) + e_cu: the unit
    // of the identity stays exact (cu = its column, -1 once eliminated), so
    // the small X of high frequencies is not rounded away against 1.
    R ar[NBT], ai[NBT];
#pragma unroll
    for (int j = 0; j < NBT; ++j) {
      const C pz = s_pz[gl * NBT + j];
      ar[j] = pz.x;
      ai[j] = pz.y;
    }
    int cu = gl;
    // - H_k G' (m is uniform across the wavefront: the row element is
    // picked from registers with an indexed move, not spilled)
#pragma unroll 1
    for (int m = 0; m < NBT; ++m) {
      const R xr = hr[m], xi = hi[m];
#pragma unroll
      for (int j = 0; j < NBT; ++j) {
        const C g = s_gz[m * NBT + j];
        ar[j] -= xr * g.x - xi * g.y;
        ai[j] -= xr * g.y + xi * g.x;
      }
    }
    double phr = 1.0, phi = 0.0;  // unit phase of det(M)
    int inv_parity = 0;
    const double ld = lu_logdet_rows<R, NBT>(ar, ai, cu, gl, lane, phase != nullptr, phr, phi,
                                              inv_parity);
    if (gl == 0) s_red[grp] = kvalid ? ld : 0.0;
    if (phase && gl == 0 && kvalid) {
      const double sg = inv_parity ? -1.0 : 1.0;
      phase[((int64_t)iz * nspin_blk + sb) * nk + k] = make_double2(sg * phr, sg * phi);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0;
      for (int g = 0; g < MPB; ++g) t += s_red[g];
      partials[((int64_t)iz * nspin_blk + sb) * gridDim.x + kb] = t;
    }
  }
}

// ----------------------------------------- k x z log-det, MFMA contraction --
// The FP32 integrand with the contraction on the matrix cores: for each
// frequency, X_k = P(z) - H_k G'(z) for the 64/NBT matrices of a wavefront
// is one real GEMM on 16-row tiles of the stacked H_k rows,
//   [X_re | X_im] = [P_re | P_im] + [-H_re | -H_im] [[G_re, G_im], [-G_im, G_re]],
// (16 x 2NBT) += (16 x 2NBT) (2NBT x 2NBT), as v_mfma_f32_16x16x4f32 steps
// of 4 inner indices: per tile 2NBT/4 steps for each of the 2NBT/16 output
// column tiles.  The -H_k operand stays in registers across the chunk of
// frequencies, G'(z) and P(z) are staged in LDS per frequency, and each
// tile goes through LDS into the one-row-per-lane layout of the pivoted LU
// (lu_logdet_rows), which runs on the vector ALUs.  Operand lanes (16x16x4
// f32): A(i, k) from lane i + 16k, B(k, j) from lane j + 16k, D(4q + r, j)
// in accumulator r of lane j + 16q.  Same arguments and partial sums as
// k_logdet<float, NBT>.
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

template <int NBT>
__global__ __launch_bounds__(kOBlock) void k_logdet_mfma(int nk, int nz, int zchunk,
                                                         int nspin_blk, const float2 *hk,
                                                         const float2 *gmat, const float2 *pmat,
                                                         double *partials, double2 *phase) {
  static_assert(NBT == 8 || NBT == 16, "8- or 16-row blocks");
  constexpr int MPB = kOBlock / NBT;  // matrices per block
  constexpr int MPW = 64 / NBT;       // matrices per wavefront
  constexpr int NT = 4;               // 16-row tiles per wavefront (64 stacked rows)
  constexpr int NCT = 2 * NBT / 16;   // 16-column tiles of [re | im]
  constexpr int KC = 2 * NBT / 4;     // 4-deep steps over the inner [re | im] index
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int gl = lane & (NBT - 1);
  const int grp = threadIdx.x / NBT;
  const int kb = blockIdx.x;
  const int sb = blockIdx.z;
  const int k = kb * MPB + grp;
  const bool kvalid = k < nk;
  const int ti = lane & 15, tq = lane >> 4;
  __shared__ double s_red[MPB];
  __shared__ float2 s_gz[NBT * NBT], s_pz[NBT * NBT];
  __shared__ float s_x[kOBlock / 64][NT * NCT][16 * 17];  // rows padded to 17
  // A operand (-H_k), constant over the frequencies: tile t row ti is row
  // (t * 16 + ti) % NBT of the wavefront's matrix (t * 16 + ti) / NBT
  float a_op[NT][KC];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int row = t * 16 + ti;
    const int kk = kb * MPB + wave * MPW + row / NBT;
    const int r = row % NBT;
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      const int kin = 4 * c + tq;  // inner index: [re 0..NBT-1 | im 0..NBT-1]
      float v = 0.0f;
      if (kk < nk) {
        const float2 h = hk[(((int64_t)kk * nspin_blk + sb) * NBT + r) * NBT + (kin % NBT)];
        v = kin < NBT ? h.x : h.y;
      }
      a_op[t][c] = -v;
    }
  }
  const int zbeg = blockIdx.y * zchunk;
  const int zend = min(zbeg + zchunk, nz);
  for (int iz = zbeg; iz < zend; ++iz) {
    const int64_t zoff = ((int64_t)iz * nspin_blk + sb) * kNB * kNB;
    __syncthreads();  // the previous frequency's LU reads of s_x are done
    for (int e = threadIdx.x; e < NBT * NBT; e += kOBlock) {
      const int r = e / NBT, c = e % NBT;
      s_gz[e] = gmat[zoff + r * kNB + c];
      s_pz[e] = pmat[zoff + r * kNB + c];
    }
    __syncthreads();
    // B operand: B(kin, col) for kin = 4c + tq, col = 16 ct + ti
    float b_op[NCT][KC];
    mfma_f4 c_in[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      const int col = 16 * ct + ti;
      const int j = col % NBT;
      const bool oim = col >= NBT;
#pragma unroll
      for (int c = 0; c < KC; ++c) {
        const int kin = 4 * c + tq;
        const float2 g = s_gz[(kin % NBT) * NBT + j];
        b_op[ct][c] = kin < NBT ? (oim ? g.y : g.x) : (oim ? g.x : -g.y);
      }
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const float2 pz = s_pz[((4 * tq + rr) % NBT) * NBT + j];
        c_in[ct][rr] = oim ? pz.y : pz.x;
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        mfma_f4 acc = c_in[ct];
#pragma unroll
        for (int c = 0; c < KC; ++c)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a_op[t][c], b_op[ct][c], acc, 0, 0, 0);
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) s_x[wave][t * NCT + ct][(4 * tq + rr) * 17 + ti] = acc[rr];
      }
    }
    __syncthreads();
    // this lane's row of its matrix: wavefront row q * NBT + gl
    const int wrow = (lane / NBT) * NBT + gl;
    const int t = wrow / 16, rt = wrow % 16;
    float ar[NBT], ai[NBT];
#pragma unroll
    for (int j = 0; j < NBT; ++j) {
      ar[j] = s_x[wave][t * NCT + j / 16][rt * 17 + (j % 16)];
      ai[j] = s_x[wave][t * NCT + (NBT + j) / 16][rt * 17 + ((NBT + j) % 16)];
    }
    double phr = 1.0, phi = 0.0;
    int inv_parity = 0;
    const double ld = lu_logdet_rows<float, NBT>(ar, ai, gl, gl, lane, phase != nullptr, phr,
                                                 phi, inv_parity);
    if (gl == 0) s_red[grp] = kvalid ? ld : 0.0;
    if (phase && gl == 0 && kvalid) {
      const double sg = inv_parity ? -1.0 : 1.0;
      phase[((int64_t)iz * nspin_blk + sb) * nk + k] = make_double2(sg * phr, sg * phi);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double tsum = 0;
      for (int g = 0; g < MPB; ++g) tsum += s_red[g];
      partials[((int64_t)iz * nspin_blk + sb) * gridDim.x + kb] = tsum;
    }
  }
}

// ------------------------ k x z log-det, MFMA contraction, a matrix per lane --
// The FP32 integrand for 8-row blocks with no cross-lane step in the LU: a
// wavefront holds 64 matrices, lane l the one of k-point (l % KPW) of its
// KPW and frequency (l / KPW) of the ZPI in flight.  Per frequency the KPW
// contractions X_k = P(z) - H_k G'(z) are KPW/2 tiles of 16 stacked rows
// (two k-points each) on the matrix cores, as in k_logdet_mfma; each tile
// is stored to the wavefront's LDS region column-major (a lane's four
// accumulators are four consecutive rows of one column: one 16-byte
// store), and once all ZPI frequencies are stored each lane reads its whole
// 8 x 8 matrix into registers.  Each lane then runs the LU with partial
// pivoting by itself: the pivot row is found by compares and brought up by
// selects, the unit of the identity kept exact per row (M(r, :) = X(r, :) +
// e_cu(r)), the pivot formed and the pivot product kept in FP64 (mantissa
// and exponent).  Partial sums per (z, block of KPB k-points); phases as in
// k_logdet.
constexpr int kLane8Kpw = 8;  // k-points per wavefront of k_logdet_lane8 (x 8 frequencies)

template <int KPW, int ZPI>
__global__ __launch_bounds__(kOBlock) void k_logdet_lane8(int nk, int nz, int zchunk,
                                                          int nspin_blk, const float2 *hk,
                                                          const float2 *gmat, const float2 *pmat,
                                                          double *partials, double2 *phase) {
  static_assert(KPW * ZPI == 64 && KPW % 2 == 0, "one matrix per lane");
  constexpr int NW = kOBlock / 64;
  constexpr int KPB = NW * KPW;    // k-points per block
  constexpr int NT = KPW / 2;      // 16-row tiles of one frequency
  constexpr int MS = 8 * 16 + 4;   // LDS floats per matrix: 16 columns (re, im) of 8 rows
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int ti = lane & 15, tq = lane >> 4;
  const int kl = lane % KPW, zl = lane / KPW;  // this lane's matrix
  const int sb = blockIdx.z;
  const int kw0 = blockIdx.x * KPB + wave * KPW;
  const int k = kw0 + kl;
  const bool kvalid = k < nk;
  __shared__ float2 s_g[ZPI][64], s_p[ZPI][64];
  __shared__ __align__(16) float s_x[NW][32 * MS];  // one round: half the matrices
  __shared__ double s_red[NW][ZPI];
  float *wx = s_x[wave];
  const int zbeg = blockIdx.y * zchunk;
  const int zend = min(zbeg + zchunk, nz);
  for (int z0 = zbeg; z0 < zend; z0 += ZPI) {
    __syncthreads();  // the previous frequencies are done with s_g, s_p, s_red
    for (int e = threadIdx.x; e < ZPI * 64; e += kOBlock) {
      const int zz = e / 64, rc = e % 64;
      const int iz = z0 + zz;
      float2 g = make_float2(0.f, 0.f), pz = g;
      if (iz < zend) {
        const int64_t zoff = ((int64_t)iz * nspin_blk + sb) * kNB * kNB + (rc / 8) * kNB + rc % 8;
        g = gmat[zoff];
        pz = pmat[zoff];
      }
      s_g[zz][rc] = g;
      s_p[zz][rc] = pz;
    }
    // A operand (-H_k, reloaded per chunk from cache so that it is not held
    // through the LU): tile t row ti is row (ti & 7) of k-point 2t + (ti >> 3)
    float a_op[NT][4];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int kk = kw0 + 2 * t + (ti >> 3);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int kin = 4 * c + tq;
        float v = 0.0f;
        if (kk < nk) {
          const float2 h = hk[(((int64_t)kk * nspin_blk + sb) * 8 + (ti & 7)) * 8 + (kin & 7)];
          v = kin < 8 ? h.x : h.y;
        }
        a_op[t][c] = -v;
      }
    }
    __syncthreads();
    const int j = ti & 7;
    const bool oim = ti >= 8;
    float xr[8][8], xi[8][8];
    // two rounds of ZPI/2 frequencies (32 matrices each): the lanes of a
    // round read their matrices before the next round's tiles overwrite
    // the region (the wavefront's own LDS: its stores and loads run in order)
#pragma unroll
    for (int half = 0; half < 2; ++half) {
#pragma unroll 2
    for (int zq = 0; zq < ZPI / 2; ++zq) {
      const int zp = half * (ZPI / 2) + zq;
      // B(kin, col) for kin = 4c + tq, col = ti; C: P(z) rows
      float b_op[4];
      mfma_f4 c_in;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int kin = 4 * c + tq;
        const float2 g = s_g[zp][(kin & 7) * 8 + j];
        b_op[c] = kin < 8 ? (oim ? g.y : g.x) : (oim ? g.x : -g.y);
      }
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const float2 pz = s_p[zp][((4 * tq + rr) & 7) * 8 + j];
        c_in[rr] = oim ? pz.y : pz.x;
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        mfma_f4 acc = c_in;
#pragma unroll
        for (int c = 0; c < 4; ++c)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a_op[t][c], b_op[c], acc, 0, 0, 0);
        // D(4 tq + rr, ti): rows 4 (tq & 1) .. +3 of k-point 2t + tq / 2,
        // column (j, re/im): matrix zq * KPW + 2t + tq / 2 of the round
        const int m = zq * KPW + 2 * t + (tq >> 1);
        *reinterpret_cast<float4 *>(&wx[m * MS + (2 * j + (oim ? 1 : 0)) * 8 + 4 * (tq & 1)]) =
            make_float4(acc[0], acc[1], acc[2], acc[3]);
      }
    }
    if ((zl >= ZPI / 2) == (half == 1)) {
      const float4 *m4 =
          reinterpret_cast<const float4 *>(&wx[(lane - half * (ZPI / 2) * KPW) * MS]);
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const float4 r0 = m4[(2 * jj) * 2], r1 = m4[(2 * jj) * 2 + 1];
        const float4 i0 = m4[(2 * jj + 1) * 2], i1 = m4[(2 * jj + 1) * 2 + 1];
        xr[0][jj] = r0.x; xr[1][jj] = r0.y; xr[2][jj] = r0.z; xr[3][jj] = r0.w;
        xr[4][jj] = r1.x; xr[5][jj] = r1.y; xr[6][jj] = r1.z; xr[7][jj] = r1.w;
        xi[0][jj] = i0.x; xi[1][jj] = i0.y; xi[2][jj] = i0.z; xi[3][jj] = i0.w;
        xi[4][jj] = i1.x; xi[5][jj] = i1.y; xi[6][jj] = i1.z; xi[7][jj] = i1.w;
      }
    }
    }  // round
    // LU with partial pivoting of M = X + (unit of row r in column cu[r])
    int cu[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) cu[r] = r;
    double pm_acc = 1.0, phr = 1.0, phi = 0.0;
    int pm_exp = 0, parity = 0;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      int p = kk;
      float best;
      {
        const float mr = xr[kk][kk] + (cu[kk] == kk ? 1.0f : 0.0f);
        best = mr * mr + xi[kk][kk] * xi[kk][kk];
      }
#pragma unroll
      for (int r = kk + 1; r < 8; ++r) {
        const float mr = xr[r][kk] + (cu[r] == kk ? 1.0f : 0.0f);
        const float m2 = mr * mr + xi[r][kk] * xi[r][kk];
        if (m2 > best) {
          best = m2;
          p = r;
        }
      }
      // rows kk and p trade places; skipped when no lane of the wavefront
      // pivots (the diagonal wins wherever M is close to the identity)
      if (__any(p != kk)) {
#pragma unroll
        for (int r = kk + 1; r < 8; ++r) {
          const bool sw = p == r;
#pragma unroll
          for (int jj = kk; jj < 8; ++jj) {
            const float ar = xr[kk][jj], br = xr[r][jj], ai = xi[kk][jj], bi = xi[r][jj];
            xr[kk][jj] = sw ? br : ar;
            xr[r][jj] = sw ? ar : br;
            xi[kk][jj] = sw ? bi : ai;
            xi[r][jj] = sw ? ai : bi;
          }
          const int ca = cu[kk], cb = cu[r];
          cu[kk] = sw ? cb : ca;
          cu[r] = sw ? ca : cb;
        }
        parity ^= p != kk ? 1 : 0;
      }
      // the pivot with its unit formed in FP64 (1 + a small X keeps X's digits)
      const double pvd = (double)xr[kk][kk] + (cu[kk] == kk ? 1.0 : 0.0);
      const float pvr = (float)pvd;
      const float pvi = xi[kk][kk];
      {
        const double pm = pvd * pvd + (double)pvi * pvi;
        int ex;
        pm_acc = frexp(pm_acc * pm, &ex);
        pm_exp += ex;
        if (phase) {
          const double rm = 1.0 / sqrt(pm);
          const double ur = pvd * rm, ui = pvi * rm;
          const double nr = phr * ur - phi * ui;
          phi = phr * ui + phi * ur;
          phr = nr;
        }
      }
      const float inv = 1.0f / (pvr * pvr + pvi * pvi);
      float pr[8], pim[8];  // the pivot row with its unit
#pragma unroll
      for (int jj = kk + 1; jj < 8; ++jj) {
        pr[jj] = xr[kk][jj] + (cu[kk] == jj ? 1.0f : 0.0f);
        pim[jj] = xi[kk][jj];
      }
#pragma unroll
      for (int r = kk + 1; r < 8; ++r) {
        const float mr = xr[r][kk] + (cu[r] == kk ? 1.0f : 0.0f);
        const float mi = xi[r][kk];
        const float lr = (mr * pvr + mi * pvi) * inv;  // l = M(r, kk) / pivot
        const float li = (mi * pvr - mr * pvi) * inv;
#pragma unroll
        for (int jj = kk + 1; jj < 8; ++jj) {
          xr[r][jj] -= lr * pr[jj] - li * pim[jj];
          xi[r][jj] -= lr * pim[jj] + li * pr[jj];
        }
        if (cu[r] == kk) cu[r] = -1;  // the unit sat in the eliminated column
      }
    }
    const int iz = z0 + zl;
    double ld = 0.5 * (log(pm_acc) + (double)pm_exp * 0.69314718055994530942);
    if (!kvalid || iz >= zend) ld = 0.0;
    if (phase && kvalid && iz < zend) {
      const double sg = parity ? -1.0 : 1.0;
      phase[((int64_t)iz * nspin_blk + sb) * nk + k] = make_double2(sg * phr, sg * phi);
    }
    // sum over the wavefront's KPW k-points of each frequency, then the block's
#pragma unroll
    for (int off = KPW / 2; off > 0; off >>= 1) ld += __shfl_xor(ld, off, KPW);
    if (kl == 0) s_red[wave][zl] = ld;
    __syncthreads();
    if (threadIdx.x < ZPI && z0 + (int)threadIdx.x < zend) {
      double tsum = 0;
      for (int w = 0; w < NW; ++w) tsum += s_red[w][threadIdx.x];
      partials[((int64_t)(z0 + threadIdx.x) * nspin_blk + sb) * gridDim.x + blockIdx.x] = tsum;
    }
  }
}

// --------------------------------- k x z log-det, FP64, a matrix per lane --
// The FP64 integrand for 8-row blocks with no cross-lane step at all: as in
// k_logdet_lane8, a wavefront holds 64 matrices (lane l: k-point l % KPW of
// the wavefront's KPW, frequency l / KPW of the ZPI in flight), but each
// lane also forms its own contraction X = P(z) - H_k G'(z) on the vector
// ALUs: H_k of the block's k-points sits in LDS for the whole chunk of
// frequencies, G'(z) and P(z) of the ZPI frequencies are staged per round,
// and every LDS read of a wavefront touches only KPW or ZPI distinct
// addresses (broadcast to the other lanes).  (On MI355X the FP64 matrix
// cores run at the FP64 vector rate, so the contraction gains nothing on
// them.)  The LU is k_logdet_lane8's in FP64: partial pivoting by compares
// and selects, rows swapped only when a lane of the wavefront pivots, the
// identity's unit kept exact per row, the pivot product as FP64 mantissa and
// exponent.  Partial sums and phases as in k_logdet_lane8.
template <int KPW, int ZPI>
__global__ __launch_bounds__(kOBlock, 1) void k_logdet_lane64(int nk, int nz, int zchunk,
                                                              int nspin_blk, const double2 *hk,
                                                              const double2 *gmat,
                                                              const double2 *pmat,
                                                              double *partials, double2 *phase) {
  static_assert(KPW * ZPI == 64, "one matrix per lane");
  constexpr int NW = kOBlock / 64;
  constexpr int KPB = NW * KPW;  // k-points per block
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kl = lane % KPW, zl = lane / KPW;  // this lane's matrix
  const int sb = blockIdx.z;
  const int kb0 = blockIdx.x * KPB;
  const int kloc = wave * KPW + kl;  // k-point within the block
  const int k = kb0 + kloc;
  const bool kvalid = k < nk;
  __shared__ double2 s_h[KPB][64];
  __shared__ double2 s_g[ZPI][64], s_p[ZPI][64];
  __shared__ double s_red[NW][ZPI];
  // H_k of the block's k-points (zero past nk), for the whole chunk
  for (int e = threadIdx.x; e < KPB * 64; e += kOBlock) {
    const int kk = e / 64, rc = e % 64;
    double2 h = make_double2(0.0, 0.0);
    if (kb0 + kk < nk) h = hk[((int64_t)(kb0 + kk) * nspin_blk + sb) * 64 + rc];
    s_h[kk][rc] = h;
  }
  const int zbeg = blockIdx.y * zchunk;
  const int zend = min(zbeg + zchunk, nz);
  for (int z0 = zbeg; z0 < zend; z0 += ZPI) {
    __syncthreads();  // H staged; the previous frequencies are done with s_g, s_p, s_red
    for (int e = threadIdx.x; e < ZPI * 64; e += kOBlock) {
      const int zz = e / 64, rc = e % 64;
      const int iz = z0 + zz;
      double2 g = make_double2(0.0, 0.0), pz = g;
      if (iz < zend) {
        const int64_t zoff = ((int64_t)iz * nspin_blk + sb) * kNB * kNB + (rc / 8) * kNB + rc % 8;
        g = gmat[zoff];
        pz = pmat[zoff];
      }
      s_g[zz][rc] = g;
      s_p[zz][rc] = pz;
    }
    __syncthreads();
    // X = P - H G, row by row
    double xr[8][8], xi[8][8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      double2 h[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) h[j] = s_h[kloc][r * 8 + j];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const double2 pz = s_p[zl][r * 8 + c];
        double ar = pz.x, ai = pz.y;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const double2 g = s_g[zl][j * 8 + c];
          ar = fma(-h[j].x, g.x, ar);
          ar = fma(h[j].y, g.y, ar);
          ai = fma(-h[j].x, g.y, ai);
          ai = fma(-h[j].y, g.x, ai);
        }
        xr[r][c] = ar;
        xi[r][c] = ai;
      }
    }
    // LU with partial pivoting of M = X + (unit of row r in column cu[r])
    int cu[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) cu[r] = r;
    double pm_acc = 1.0, phr = 1.0, phi = 0.0;
    int pm_exp = 0, parity = 0;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      int p = kk;
      double best;
      {
        const double mr = xr[kk][kk] + (cu[kk] == kk ? 1.0 : 0.0);
        best = mr * mr + xi[kk][kk] * xi[kk][kk];
      }
#pragma unroll
      for (int r = kk + 1; r < 8; ++r) {
        const double mr = xr[r][kk] + (cu[r] == kk ? 1.0 : 0.0);
        const double m2 = mr * mr + xi[r][kk] * xi[r][kk];
        if (m2 > best) {
          best = m2;
          p = r;
        }
      }
      // rows kk and p trade places; skipped when no lane of the wavefront
      // pivots
      if (__any(p != kk)) {
#pragma unroll
        for (int r = kk + 1; r < 8; ++r) {
          const bool sw = p == r;
#pragma unroll
          for (int jj = kk; jj < 8; ++jj) {
            const double ar = xr[kk][jj], br = xr[r][jj], ai = xi[kk][jj], bi = xi[r][jj];
            xr[kk][jj] = sw ? br : ar;
            xr[r][jj] = sw ? ar : br;
            xi[kk][jj] = sw ? bi : ai;
            xi[r][jj] = sw ? ai : bi;
          }
          const int ca = cu[kk], cb = cu[r];
          cu[kk] = sw ? cb : ca;
          cu[r] = sw ? ca : cb;
        }
        parity ^= p != kk ? 1 : 0;
      }
      const double pvr = xr[kk][kk] + (cu[kk] == kk ? 1.0 : 0.0);
      const double pvi = xi[kk][kk];
      const double pm = pvr * pvr + pvi * pvi;
      {
        int ex;
        pm_acc = frexp(pm_acc * pm, &ex);
        pm_exp += ex;
        if (phase) {
          const double rm = 1.0 / sqrt(pm);
          const double ur = pvr * rm, ui = pvi * rm;
          const double nr = phr * ur - phi * ui;
          phi = phr * ui + phi * ur;
          phr = nr;
        }
      }
      const double inv = 1.0 / pm;
      double pr[8], pim[8];  // the pivot row with its unit
#pragma unroll
      for (int jj = kk + 1; jj < 8; ++jj) {
        pr[jj] = xr[kk][jj] + (cu[kk] == jj ? 1.0 : 0.0);
        pim[jj] = xi[kk][jj];
      }
#pragma unroll
      for (int r = kk + 1; r < 8; ++r) {
        const double mr = xr[r][kk] + (cu[r] == kk ? 1.0 : 0.0);
        const double mi = xi[r][kk];
        const double lr = (mr * pvr + mi * pvi) * inv;  // l = M(r, kk) / pivot
        const double li = (mi * pvr - mr * pvi) * inv;
#pragma unroll
        for (int jj = kk + 1; jj < 8; ++jj) {
          xr[r][jj] -= lr * pr[jj] - li * pim[jj];
          xi[r][jj] -= lr * pim[jj] + li * pr[jj];
        }
        if (cu[r] == kk) cu[r] = -1;  // the unit sat in the eliminated column
      }
    }
    const int iz = z0 + zl;
    double ld = 0.5 * (log(pm_acc) + (double)pm_exp * 0.69314718055994530942);
    if (!kvalid || iz >= zend) ld = 0.0;
    if (phase && kvalid && iz < zend) {
      const double sg = parity ? -1.0 : 1.0;
      phase[((int64_t)iz * nspin_blk + sb) * nk + k] = make_double2(sg * phr, sg * phi);
    }
    // sum over the wavefront's KPW k-points of each frequency, then the block's
#pragma unroll
    for (int off = KPW / 2; off > 0; off >>= 1) ld += __shfl_xor(ld, off, KPW);
    if (kl == 0) s_red[wave][zl] = ld;
    __syncthreads();
    if (threadIdx.x < ZPI && z0 + (int)threadIdx.x < zend) {
      double tsum = 0;
      for (int w = 0; w < NW; ++w) tsum += s_red[w][threadIdx.x];
      partials[((int64_t)(z0 + threadIdx.x) * nspin_blk + sb) * gridDim.x + blockIdx.x] = tsum;
    }
  }
}

__global__ __launch_bounds__(64) void k_reduce_z(int nz, int nspin_blk, int nkb, int nk,
                                                 const double *partials, double *f) {
  const int iz = blockIdx.x;
  if (iz >= nz) return;
  double s = 0;
  for (int sb = 0; sb < nspin_blk; ++sb) {
    double t = 0;
    const double *p = partials + ((int64_t)iz * nspin_blk + sb) * nkb;
    for (int b = threadIdx.x; b < nkb; b += 64) t += p[b];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    s += t / nk;
  }
  if (threadIdx.x == 0) f[iz] = s;
}

// ----------------------------------------------- k x z log-det, QR form --
// Cross-lane moves inside a quad with DPP (no LDS round trip): quad_perm
// controls are p0 | p1 << 2 | p2 << 4 | p3 << 6.
template <int CTRL>
__device__ inline float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ inline double dpp_mov(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)(b & 0xffffffffull), CTRL, 0xF, 0xF,
                                          false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                          (unsigned long long)(unsigned)lo));
}

// Sum over the L lanes of a group (L = 1, 2, 4; groups are aligned inside a
// quad).  Every lane of the group gets bit-identical sums.
template <int L, typename T>
__device__ inline T group_sum(T v) {
  if constexpr (L >= 2) v += dpp_mov<0xB1>(v);  // lane ^ 1
  if constexpr (L >= 4) v += dpp_mov<0x4E>(v);  // lane ^ 2
  return v;
}

// Value of lane O of the group in every lane of the group.
template <int L, int O, typename T>
__device__ inline T group_bcast(T v) {
  if constexpr (L == 1) return v;
  else if constexpr (L == 2) return dpp_mov<(O | (O << 2) | ((2 + O) << 4) | ((2 + O) << 6))>(v);
  else return dpp_mov<O * 0x55>(v);
}

// Compile-time loop: f(integral_constant<int, I>) for I = B .. E-1.
template <int B, int E>
struct StaticFor {
  template <typename F>
  __device__ static void run(F &&f) {
    if constexpr (B < E) {
      f(std::integral_constant<int, B>{});
      StaticFor<B + 1, E>::run(f);
    }
  }
};

// Same integrand as k_logdet, by Householder QR instead of pivoted LU:
// |det M| = prod_k ||x_k|| (x_k = column k from the diagonal down, after
// the previous reflections) and the phase of det M = prod_k x_k0/|x_k0|
// (each reflector I - tau v v^H, tau = 1/(||x||(||x|| + |x_0|)), has
// determinant -1, and R_kk = -e^{i arg x_0} ||x||).  QR needs no pivot
// search and no row exchange, so every cross-lane step is a DPP sum or a
// broadcast from a fixed lane.  L lanes per matrix, RL = NB/L rows per lane
// distributed cyclically (row = lane + L*q) so that the shrinking active
// part stays spread over the lanes; H_k rows in registers across the
// frequency chunk, G'(z) and P(z) staged in LDS per frequency.
template <typename R, int NB, int L>
__global__ __launch_bounds__(kOBlock) void k_logdet_qr(int nk, int nz, int zchunk,
                                                       int nspin_blk, const ocplx<R> *hk,
                                                       const ocplx<R> *gmat,
                                                       const ocplx<R> *pmat, double *partials,
                                                       double2 *phase) {
  using C = ocplx<R>;
  constexpr int RL = NB / L;
  constexpr int MPB = kOBlock / L;  // matrices per workgroup
  const int l = threadIdx.x % L;
  const int grp = threadIdx.x / L;
  const int kb = blockIdx.x;
  const int sb = blockIdx.z;
  const int k = kb * MPB + grp;
  const bool kvalid = k < nk;
  __shared__ double s_red[MPB];
  __shared__ C s_gz[NB * NB], s_pz[NB * NB];
  R hr[RL][NB], hi[RL][NB];
  {
    const C *hb = hk + ((int64_t)(kvalid ? k : 0) * nspin_blk + sb) * NB * NB;
#pragma unroll
    for (int q = 0; q < RL; ++q)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const C v = hb[(l + L * q) * NB + j];
        hr[q][j] = kvalid ? v.x : (R)0;
        hi[q][j] = kvalid ? v.y : (R)0;
      }
  }
  const int zbeg = blockIdx.y * zchunk;
  const int zend = min(zbeg + zchunk, nz);
  for (int iz = zbeg; iz < zend; ++iz) {
    const int64_t zoff = ((int64_t)iz * nspin_blk + sb) * kNB * kNB;
    __syncthreads();
    for (int e = threadIdx.x; e < NB * NB; e += kOBlock) {
      const int r = e / NB, c = e % NB;
      s_gz[e] = gmat[zoff + r * kNB + c];
      s_pz[e] = pmat[zoff + r * kNB + c];
    }
    __syncthreads();
    // M = 1 + P - H_k G'
    R mr[RL][NB], mi[RL][NB];
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      const int r = l + L * q;
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const C pz = s_pz[r * NB + j];
        mr[q][j] = (r == j ? (R)1 : (R)0) + pz.x;
        mi[q][j] = pz.y;
      }
    }
    // one row of G' live at a time (m is uniform: H_k(:, m) is picked
    // from registers with an indexed move)
#pragma unroll 1
    for (int m = 0; m < NB; ++m) {
      R hmr[RL], hmi[RL];
#pragma unroll
      for (int q = 0; q < RL; ++q) {
        hmr[q] = hr[q][m];
        hmi[q] = hi[q][m];
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const C g = s_gz[m * NB + j];
#pragma unroll
        for (int q = 0; q < RL; ++q) {
          mr[q][j] -= hmr[q] * g.x - hmi[q] * g.y;
          mi[q][j] -= hmr[q] * g.y + hmi[q] * g.x;
        }
      }
    }
    double pm_acc = 1.0;
    int pm_exp = 0;
    double phr = 1.0, phi = 0.0;
    StaticFor<0, NB>::run([&](auto kkc) {
      constexpr int kk = decltype(kkc)::value;
      constexpr int ow = kk % L, q0 = kk / L;  // lane and slot holding row kk
      R vr[RL], vi[RL];
      double n2 = 0.0;
#pragma unroll
      for (int q = q0; q < RL; ++q) {
        R xr = mr[q][kk], xi = mi[q][kk];
        if (q == q0 && l < ow) xr = xi = (R)0;  // rows above kk are done
        vr[q] = xr;
        vi[q] = xi;
        n2 += (double)xr * xr + (double)xi * xi;
      }
      n2 = group_sum<L>(n2);
      const R x0r = group_bcast<L, ow>(mr[q0][kk]);
      const R x0i = group_bcast<L, ow>(mi[q0][kk]);
      const double ax0 = sqrt((double)x0r * x0r + (double)x0i * x0i);
      const double nx = sqrt(n2);
      {
        int ex;
        pm_acc = frexp(pm_acc * n2, &ex);
        pm_exp += ex;
      }
      const double er = ax0 > 0.0 ? (double)x0r / ax0 : 1.0;
      const double ei = ax0 > 0.0 ? (double)x0i / ax0 : 0.0;
      if (phase) {
        const double t = phr * er - phi * ei;
        phi = phr * ei + phi * er;
        phr = t;
      }
      if constexpr (kk + 1 < NB) {
        const R tau = nx > 0.0 ? (R)(1.0 / (nx * (nx + ax0))) : (R)0;
        if (l == ow) {
          vr[q0] = x0r + (R)(er * nx);
          vi[q0] = x0i + (R)(ei * nx);
        }
#pragma unroll
        for (int j = kk + 1; j < NB; ++j) {
          R wr = 0, wi = 0;  // w = v^H M(:, j)
#pragma unroll
          for (int q = q0; q < RL; ++q) {
            wr += vr[q] * mr[q][j] + vi[q] * mi[q][j];
            wi += vr[q] * mi[q][j] - vi[q] * mr[q][j];
          }
          wr = group_sum<L>(wr) * tau;
          wi = group_sum<L>(wi) * tau;
#pragma unroll
          for (int q = q0; q < RL; ++q) {
            mr[q][j] -= vr[q] * wr - vi[q] * wi;
            mi[q][j] -= vr[q] * wi + vi[q] * wr;
          }
        }
      }
    });
    const double ld = 0.5 * (log(pm_acc) + (double)pm_exp * 0.69314718055994530942);
    if (l == 0) s_red[grp] = kvalid ? ld : 0.0;
    if (phase && l == 0 && kvalid)
      phase[((int64_t)iz * nspin_blk + sb) * nk + k] = make_double2(phr, phi);
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0;
      for (int g = 0; g < MPB; ++g) t += s_red[g];
      partials[((int64_t)iz * nspin_blk + sb) * gridDim.x + kb] = t;
    }
  }
}

// fim[z] = (1/Nk) sum_k arg(det M_k(z)), the principal argument of the
// product of the block phases (the imaginary part of log(det) as the
// principal complex logarithm of the full determinant).
__global__ __launch_bounds__(64) void k_reduce_phase(int nz, int nspin_blk, int nk,
                                                     const double2 *phase, double *fim) {
  const int iz = blockIdx.x;
  if (iz >= nz) return;
  double t = 0;
  for (int k = threadIdx.x; k < nk; k += 64) {
    double pr = 1.0, pi = 0.0;
    for (int sb = 0; sb < nspin_blk; ++sb) {
      const double2 u = phase[((int64_t)iz * nspin_blk + sb) * nk + k];
      const double nr = pr * u.x - pi * u.y;
      pi = pr * u.y + pi * u.x;
      pr = nr;
    }
    t += atan2(pi, pr);
  }
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  if (threadIdx.x == 0) fim[iz] = t / nk;
}

// ------------------------------------------- G, Sigma, G0 on a grid ----
// Grid point iz of the Matsubara axis (i (2 iz + 1) pi / beta) or of the
// real axis (wini + iz (wfin - wini)/(L - 1) + i eps).
struct GridSpec {
  int axis;      // 0 Matsubara, 1 real
  double pib;    // pi / beta
  double w0, dw, eps;
  int iz0;       // first grid index of this pass
};

__device__ __forceinline__ double2 grid_z(const GridSpec &g, int iz) {
  const int n = g.iz0 + iz;
  return g.axis == 0 ? make_double2(0.0, g.pib * (double)(2 * n + 1))
                     : make_double2(g.w0 + (double)n * g.dw, g.eps);
}

// 1/x for x > 0 from the hardware reciprocal estimate and two Newton steps
// r <- r (2 - x r): each squares the relative error, which takes it from
// the estimate's 2^-23 to the rounding of the last step (about 1 ulp), in
// 5 instructions against the 10 of the IEEE division sequence.  x here is
// |z - p|^2 >= Im(z)^2, positive on the Matsubara grid and for EPS > 0 on
// the real axis (EPS = 0 on a pole gives NaN, as 0/0 in the division).
__device__ __forceinline__ double recip_pos(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// graw[e][iz] = sum_p w_p / (z_iz - pole_p) over the poles of element e.
// A workgroup covers kOBlock frequencies of one element, so every lane reads
// the same pole at the same time (uniform loads) and the stores coalesce.
__global__ __launch_bounds__(kOBlock) void k_pole_grid(int nz, GridSpec gs, const int64_t *off,
                                                       const double2 *w, const double *pole,
                                                       double2 *graw) {
  const int e = blockIdx.y;
  const int iz = blockIdx.x * kOBlock + threadIdx.x;
  const double2 zz = grid_z(gs, iz < nz ? iz : 0);
  const double di = zz.y, di2 = di * di;
  double gr = 0, gi = 0;
#pragma unroll 4
  for (int64_t p = off[e]; p < off[e + 1]; ++p) {
    const double dr = zz.x - pole[p];
    const double den = recip_pos(fma(dr, dr, di2));
    const double2 ww = w[p];
    gr = fma(fma(ww.x, dr, ww.y * di), den, gr);
    gi = fma(fma(ww.y, dr, -ww.x * di), den, gi);
  }
  if (iz < nz) graw[(int64_t)e * nz + iz] = make_double2(gr, gi);
}

struct SigmaArgs {
  int nz, nspin_blk, nb, nlat, norb, nspin, nelem;
  double chan4, xmu;
  GridSpec gs;
  const double2 *graw;      // [nelem][nz]
  const double2 *hloc_blk;  // [nspin_blk][nb][nb]
  int nbb;
  const double2 *hb_blk, *vb_blk;
  // outputs [Nlat][Nlat][Nspin][Nspin][Norb][Norb][nz], first index fastest
  double2 *g, *sigma, *g0;
};

// One workgroup per (frequency, spin block): the block of G (off-diagonal
// elements recombined from the mixed channels), G0^-1 = (z + mu) - Hloc -
// Delta(z) with Delta = V (z + mu - h)^-1 V^+, Sigma = G0^-1 - G^-1 and
// G0 = (G0^-1)^-1, all in LDS, written in the element order of the 7-index
// arrays.
__global__ __launch_bounds__(64) void k_grid_sigma(const SigmaArgs a) {
  __shared__ double2 s_g[kNB * kNB], s_a[kNB * kNB], s_t[kNB * kNB], s_inv[kNB * kNB];
  __shared__ double2 s_hb[kNB * kNB], s_hbinv[kNB * kNB];
  __shared__ double s_ld;
  __shared__ int s_piv;
  const int iz = blockIdx.x, sb = blockIdx.y, tid = threadIdx.x;
  const int nb = a.nb, no = a.norb, nl = a.nlat, nsp = a.nspin;
  const int ls = nl * no;
  const bool mixed = a.nspin_blk == 1 && nsp == 2;
  const double2 zz = grid_z(a.gs, iz);
  auto gel = [&](int sp, int i1, int j1, int o1, int o2) {
    return a.graw[(int64_t)((((sp * nl + i1) * nl + j1) * no + o1) * no + o2) * a.nz + iz];
  };
  for (int e = tid; e < nb * nb; e += blockDim.x) {
    const int r = e / nb, c = e % nb;
    const int sr = mixed ? r / ls : sb, sc = mixed ? c / ls : sb;
    double2 v = make_double2(0.0, 0.0);
    if (sr == sc) {
      const int rr = r % ls, cc = c % ls;
      const int il = rr / no, io = rr % no, jl = cc / no, jo = cc % no;
      v = gel(sr, il, jl, io, jo);
      if (!(il == jl && io == jo)) {
        const double2 gii = gel(sr, il, il, io, io), gjj = gel(sr, jl, jl, jo, jo);
        const double sx = gii.x + gjj.x, sy = gii.y + gjj.y;
        const double tx = sx + a.chan4 * sy, ty = sy - a.chan4 * sx;
        v = make_double2(0.5 * (v.x - tx), 0.5 * (v.y - ty));
      }
    }
    s_g[r * kNB + c] = v;
    s_t[r * kNB + c] = v;
    const double2 h = a.hloc_blk[((int64_t)sb * nb + r) * nb + c];
    double2 g0i = make_double2(-h.x, -h.y);
    if (r == c) {
      g0i.x += zz.x + a.xmu;
      g0i.y += zz.y;
    }
    s_a[r * kNB + c] = g0i;
  }
  if (a.nbb > 0) {
    const int nbb = a.nbb;
    const double2 *hb = a.hb_blk + (int64_t)sb * nbb * nbb;
    for (int e = tid; e < nbb * nbb; e += blockDim.x) {
      const int r = e / nbb, c = e % nbb;
      const double2 h = hb[r * nbb + c];
      double2 v = make_double2(-h.x, -h.y);
      if (r == c) {
        v.x += zz.x + a.xmu;
        v.y += zz.y;
      }
      s_hb[r * kNB + c] = v;
    }
    __syncthreads();
    gj_inverse(nbb, s_hb, s_hbinv, kNB, &s_ld, &s_piv);
    const double2 *vb = a.vb_blk + (int64_t)sb * nb * nbb;
    for (int e = tid; e < nb * nb; e += blockDim.x) {
      const int r = e / nb, c = e % nb;
      double dr = 0, di = 0;
      for (int p = 0; p < nbb; ++p) {
        const double2 vr = vb[r * nbb + p];
        for (int q = 0; q < nbb; ++q) {
          const double2 m = s_hbinv[p * kNB + q];
          const double2 vc = vb[c * nbb + q];
          const double tx = vr.x * m.x - vr.y * m.y, ty = vr.x * m.y + vr.y * m.x;
          dr += tx * vc.x + ty * vc.y;
          di += ty * vc.x - tx * vc.y;
        }
      }
      s_a[r * kNB + c].x -= dr;
      s_a[r * kNB + c].y -= di;
    }
  }
  __syncthreads();
  // position of block entry (r, c) in the 7-index output
  auto oidx = [&](int r, int c) {
    const int sr = mixed ? r / ls : sb, sc = mixed ? c / ls : sb;
    const int rr = r % ls, cc = c % ls;
    const int il = rr / no, io = rr % no, jl = cc / no, jo = cc % no;
    return (int64_t)il +
           (int64_t)nl * (jl + (int64_t)nl * (sr + (int64_t)nsp *
                                                       (sc + (int64_t)nsp *
                                                                 (io + (int64_t)no *
                                                                           (jo + (int64_t)no * iz)))));
  };
  if (a.sigma) {
    gj_inverse(nb, s_t, s_inv, kNB, &s_ld, &s_piv);  // G^-1
    for (int e = tid; e < nb * nb; e += blockDim.x) {
      const int r = e / nb, c = e % nb;
      const double2 x = s_a[r * kNB + c], y = s_inv[r * kNB + c];
      a.sigma[oidx(r, c)] = make_double2(x.x - y.x, x.y - y.y);
    }
    __syncthreads();
  }
  if (a.g)
    for (int e = tid; e < nb * nb; e += blockDim.x) {
      const int r = e / nb, c = e % nb;
      a.g[oidx(r, c)] = s_g[r * kNB + c];
    }
  if (a.g0) {
    for (int e = tid; e < nb * nb; e += blockDim.x) {
      const int r = e / nb, c = e % nb;
      s_t[r * kNB + c] = s_a[r * kNB + c];
    }
    __syncthreads();
    gj_inverse(nb, s_t, s_inv, kNB, &s_ld, &s_piv);
    for (int e = tid; e < nb * nb; e += blockDim.x) {
      const int r = e / nb, c = e % nb;
      a.g0[oidx(r, c)] = s_inv[r * kNB + c];
    }
  }
}

// ================================================================ host ===
float OmegaDev::grid_gf(int axis, int L, double beta, double wini, double wfin, double eps,
                        double2 *g, double2 *sigma, double2 *g0) {
  if (L <= 0 || !(g || sigma || g0)) return 0.0f;
  const hipStream_t st = stream;
  const int nlso = host.nlat * host.nspin * host.norb;
  const int64_t per_z = (int64_t)nlso * nlso;
  const int chunk = std::min(L, 4096);
  double2 *d_raw = nullptr, *d_out[3] = {nullptr, nullptr, nullptr};
  double2 *h_out[3] = {g, sigma, g0};
  VCA_HIP(hipMalloc(&d_raw, sizeof(double2) * (size_t)chunk * std::max(host.nelem, 1)));
  for (int i = 0; i < 3; ++i)
    if (h_out[i]) VCA_HIP(hipMalloc(&d_out[i], sizeof(double2) * (size_t)chunk * per_z));
  GridSpec gs{};
  gs.axis = axis;
  gs.pib = M_PI / beta;
  gs.w0 = wini;
  gs.dw = L > 1 ? (wfin - wini) / (double)(L - 1) : 0.0;
  gs.eps = eps;
  hipEvent_t e0, e1;
  VCA_HIP(hipEventCreate(&e0));
  VCA_HIP(hipEventCreate(&e1));
  float ms_total = 0.0f;
  for (int z0 = 0; z0 < L; z0 += chunk) {
    const int nz = std::min(chunk, L - z0);
    gs.iz0 = z0;
    VCA_HIP(hipEventRecord(e0, st));
    hipLaunchKernelGGL(k_pole_grid, dim3((nz + kOBlock - 1) / kOBlock, host.nelem),
                       dim3(kOBlock), 0, st, nz, gs, d_off, d_w, d_pole, d_raw);
    VCA_HIP(hipGetLastError());
    VCA_HIP(hipEventRecord(e1, st));
    for (int i = 0; i < 3; ++i)
      if (d_out[i])
        VCA_HIP(hipMemsetAsync(d_out[i], 0, sizeof(double2) * (size_t)nz * per_z, st));
    SigmaArgs sa{};
    sa.nz = nz;
    sa.nspin_blk = host.nspin_blk;
    sa.nb = host.nb;
    sa.nlat = host.nlat;
    sa.norb = host.norb;
    sa.nspin = host.nspin;
    sa.nelem = host.nelem;
    sa.chan4 = host.chan4;
    sa.xmu = host.xmu;
    sa.gs = gs;
    sa.graw = d_raw;
    sa.hloc_blk = d_hloc;
    sa.nbb = host.nbb;
    sa.hb_blk = d_hb;
    sa.vb_blk = d_vb;
    sa.g = d_out[0];
    sa.sigma = d_out[1];
    sa.g0 = d_out[2];
    hipLaunchKernelGGL(k_grid_sigma, dim3(nz, host.nspin_blk), dim3(64), 0, st, sa);
    VCA_HIP(hipGetLastError());
    for (int i = 0; i < 3; ++i)
      if (d_out[i])
        VCA_HIP(hipMemcpyAsync(h_out[i] + (int64_t)z0 * per_z, d_out[i],
                               sizeof(double2) * (size_t)nz * per_z, hipMemcpyDeviceToHost, st));
    VCA_HIP(hipStreamSynchronize(st));
    float ms = 0.0f;
    VCA_HIP(hipEventElapsedTime(&ms, e0, e1));
    ms_total += ms;
  }
  VCA_HIP(hipEventDestroy(e0));
  VCA_HIP(hipEventDestroy(e1));
  for (void *p : {(void *)d_raw, (void *)d_out[0], (void *)d_out[1], (void *)d_out[2]})
    if (p) VCA_HIP(hipFree(p));
  return ms_total;
}

OmegaDev::~OmegaDev() { release(); }

void OmegaDev::release() {
  for (void *p : {(void *)d_off, (void *)d_w, (void *)d_pole, (void *)d_hk64, (void *)d_hk32,
                  (void *)d_hloc, (void *)d_hb, (void *)d_vb, (void *)d_z, (void *)d_graw,
                  (void *)d_g64, (void *)d_g32, (void *)d_p64, (void *)d_p32, (void *)d_part,
                  (void *)d_f, (void *)d_phase, (void *)d_fim})
    if (p) (void)hipFree(p);
  d_phase = nullptr;
  d_fim = nullptr;
  d_off = nullptr;
  d_w = nullptr;
  d_pole = nullptr;
  d_hk64 = nullptr;
  d_hk32 = nullptr;
  d_hloc = nullptr;
  d_hb = nullptr;
  d_vb = nullptr;
  d_z = nullptr;
  d_graw = nullptr;
  d_g64 = nullptr;
  d_g32 = nullptr;
  d_p64 = nullptr;
  d_p32 = nullptr;
  d_part = nullptr;
  d_f = nullptr;
  zcap = 0;
}

void OmegaDev::upload(const OmegaHost &h, hipStream_t st) {
  release();
  host = h;
  stream = st;
  const int64_t np = (int64_t)h.w.size();
  VCA_HIP(hipMalloc(&d_off, sizeof(int64_t) * h.off.size()));
  VCA_HIP(hipMalloc(&d_w, sizeof(double2) * std::max<int64_t>(np, 1)));
  VCA_HIP(hipMalloc(&d_pole, sizeof(double) * std::max<int64_t>(np, 1)));
  VCA_HIP(hipMemcpyAsync(d_off, h.off.data(), sizeof(int64_t) * h.off.size(),
                         hipMemcpyHostToDevice, st));
  if (np) {
    VCA_HIP(hipMemcpyAsync(d_w, h.w.data(), sizeof(double2) * np, hipMemcpyHostToDevice, st));
    VCA_HIP(hipMemcpyAsync(d_pole, h.pole.data(), sizeof(double) * np, hipMemcpyHostToDevice,
                           st));
  }
  // H_k blocks repacked with the kernel's row stride (8 when blocks fit 8)
  nbt = h.nb <= 8 ? 8 : kNB;
  const size_t hk_n = (size_t)h.nk * h.nspin_blk * nbt * nbt;
  std::vector<double2> hk64(hk_n);
  std::vector<float2> hk32(hk_n);
  for (size_t kb = 0; kb < (size_t)h.nk * h.nspin_blk; ++kb)
    for (int r = 0; r < nbt; ++r)
      for (int c = 0; c < nbt; ++c) {
        const double2 v = h.hk[(kb * kNB + r) * kNB + c];
        hk64[(kb * nbt + r) * nbt + c] = v;
        hk32[(kb * nbt + r) * nbt + c] = make_float2((float)v.x, (float)v.y);
      }
  VCA_HIP(hipMalloc(&d_hk64, sizeof(double2) * hk_n));
  VCA_HIP(hipMalloc(&d_hk32, sizeof(float2) * hk_n));
  VCA_HIP(hipMemcpyAsync(d_hk64, hk64.data(), sizeof(double2) * hk_n, hipMemcpyHostToDevice,
                         st));
  VCA_HIP(hipMemcpyAsync(d_hk32, hk32.data(), sizeof(float2) * hk_n, hipMemcpyHostToDevice, st));
  VCA_HIP(hipMalloc(&d_hloc, sizeof(double2) * h.hloc.size()));
  VCA_HIP(hipMemcpyAsync(d_hloc, h.hloc.data(), sizeof(double2) * h.hloc.size(),
                         hipMemcpyHostToDevice, st));
  if (h.nbb > 0) {
    VCA_HIP(hipMalloc(&d_hb, sizeof(double2) * h.hb.size()));
    VCA_HIP(hipMalloc(&d_vb, sizeof(double2) * h.vb.size()));
    VCA_HIP(hipMemcpyAsync(d_hb, h.hb.data(), sizeof(double2) * h.hb.size(),
                           hipMemcpyHostToDevice, st));
    VCA_HIP(hipMemcpyAsync(d_vb, h.vb.data(), sizeof(double2) * h.vb.size(),
                           hipMemcpyHostToDevice, st));
  }
  VCA_HIP(hipStreamSynchronize(st));
}

void OmegaDev::ensure_z(int nz) {
  if (nz <= zcap) return;
  for (void *p : {(void *)d_z, (void *)d_graw, (void *)d_g64, (void *)d_g32, (void *)d_p64,
                  (void *)d_p32, (void *)d_part, (void *)d_f, (void *)d_phase, (void *)d_fim})
    if (p) VCA_HIP(hipFree(p));
  d_phase = nullptr;
  d_fim = nullptr;
  zcap = std::max(nz, 64);
  const int mpb = kOBlock / nbt;
  const int nkb = (host.nk + mpb - 1) / mpb;
  VCA_HIP(hipMalloc(&d_z, sizeof(double2) * zcap));
  VCA_HIP(hipMalloc(&d_graw, sizeof(double2) * (size_t)zcap * host.nelem));
  const size_t nm = (size_t)zcap * host.nspin_blk * kNB * kNB;
  VCA_HIP(hipMalloc(&d_g64, sizeof(double2) * nm));
  VCA_HIP(hipMalloc(&d_p64, sizeof(double2) * nm));
  VCA_HIP(hipMalloc(&d_g32, sizeof(float2) * nm));
  VCA_HIP(hipMalloc(&d_p32, sizeof(float2) * nm));
  VCA_HIP(hipMalloc(&d_part, sizeof(double) * (size_t)zcap * host.nspin_blk * nkb));
  VCA_HIP(hipMalloc(&d_f, sizeof(double) * zcap));
}

__global__ void k_b_to_f32(int64_t n, const double2 *s, float2 *d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = make_float2((float)s[i].x, (float)s[i].y);
}

void OmegaDev::integrand(int nz, const double2 *z, int precision, double *f, float *ms_logdet,
                         double *fim) {
  if (nz <= 0) return;
  ensure_z(nz);
  if (fim && !d_phase) {
    VCA_HIP(hipMalloc(&d_phase, sizeof(double2) * (size_t)zcap * host.nspin_blk * host.nk));
    VCA_HIP(hipMalloc(&d_fim, sizeof(double) * zcap));
  }
  double2 *phase = fim ? d_phase : nullptr;
  const hipStream_t st = stream;
  VCA_HIP(hipMemcpyAsync(d_z, z, sizeof(double2) * nz, hipMemcpyHostToDevice, st));
  {
    const int64_t nt = (int64_t)nz * host.nelem;
    hipLaunchKernelGGL(k_pole_sum, dim3((unsigned)((nt + kOBlock - 1) / kOBlock)),
                       dim3(kOBlock), 0, st, nz, d_z, host.nelem, d_off, d_w, d_pole, d_graw);
    VCA_HIP(hipGetLastError());
  }
  PrepArgs pa{};
  pa.nz = nz;
  pa.nspin_blk = host.nspin_blk;
  pa.nb = host.nb;
  pa.nlat = host.nlat;
  pa.norb = host.norb;
  pa.nspin_gf = host.nspin;
  pa.chan4 = host.chan4;
  pa.xmu = host.xmu;
  pa.z = d_z;
  pa.graw = d_graw;
  pa.nelem = host.nelem;
  pa.hloc_blk = d_hloc;
  pa.nbb = host.nbb;
  pa.hb_blk = d_hb;
  pa.vb_blk = d_vb;
  pa.gmat = d_g64;
  pa.pmat = d_p64;
  hipLaunchKernelGGL(k_prep, dim3(nz, host.nspin_blk), dim3(64), 0, st, pa);
  VCA_HIP(hipGetLastError());
  // FP64 8-row blocks: the QR kernel (4 lanes per matrix).  FP32, 16-row
  // blocks, or VCA_OMEGA_LU=1: the pivoted-LU kernel (a row per lane).  In
  // FP32 the LU is the more accurate of the two: M is close to the
  // identity, so its multipliers are small, while a Householder reflection
  // always makes O(1) updates (3x the rounding error of the Omega integral
  // in FP32 on the 2x2 BHZ grid).
  const char *lu_env = std::getenv("VCA_OMEGA_LU");
  // FP64 8-row blocks: a matrix per lane (k_logdet_lane64); VCA_OMEGA_LANE=0
  // keeps the QR kernel
  const char *ln_env = std::getenv("VCA_OMEGA_LANE");
  const bool lane64 = nbt == 8 && precision == VCA_FP64 && !(lu_env && lu_env[0] == '1') &&
                      !(ln_env && ln_env[0] == '0');
  const bool use_qr =
      nbt == 8 && precision == VCA_FP64 && !(lu_env && lu_env[0] == '1') && !lane64;
  // FP32: the contraction on the matrix cores (VCA_OMEGA_MFMA=0: on the
  // vector ALUs, k_logdet); 8-row blocks with one matrix per lane
  // (VCA_OMEGA_MFMA=row: a row per lane, k_logdet_mfma<8>)
  const char *mf_env = std::getenv("VCA_OMEGA_MFMA");
  const bool mfma = precision != VCA_FP64 && !(mf_env && mf_env[0] == '0');
  const bool lane8 = mfma && nbt == 8 && !(mf_env && mf_env[0] == 'r');
  const int lanes = use_qr ? 4 : nbt;
  const int mpb = (lane8 || lane64) ? (kOBlock / 64) * kLane8Kpw : kOBlock / lanes;
  const int nkb = (host.nk + mpb - 1) / mpb;
  const int zchunk = 16;
  const dim3 grid(nkb, (nz + zchunk - 1) / zchunk, host.nspin_blk);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms_logdet) {
    VCA_HIP(hipEventCreate(&e0));
    VCA_HIP(hipEventCreate(&e1));
  }
  if (precision == VCA_FP64) {
    if (e0) VCA_HIP(hipEventRecord(e0, st));
    if (lane64)
      hipLaunchKernelGGL((k_logdet_lane64<kLane8Kpw, 64 / kLane8Kpw>), grid, dim3(kOBlock), 0, st,
                         host.nk, nz, zchunk, host.nspin_blk, d_hk64, d_g64, d_p64, d_part, phase);
    else if (use_qr)
      hipLaunchKernelGGL((k_logdet_qr<double, 8, 4>), grid, dim3(kOBlock), 0, st, host.nk, nz,
                         zchunk, host.nspin_blk, d_hk64, d_g64, d_p64, d_part, phase);
    else if (nbt == 8)
      hipLaunchKernelGGL((k_logdet<double, 8>), grid, dim3(kOBlock), 0, st, host.nk, nz, zchunk,
                         host.nspin_blk, d_hk64, d_g64, d_p64, d_part, phase);
    else
      hipLaunchKernelGGL((k_logdet<double, 16>), grid, dim3(kOBlock), 0, st, host.nk, nz,
                         zchunk, host.nspin_blk, d_hk64, d_g64, d_p64, d_part, phase);
  } else {
    const int64_t nm = (int64_t)nz * host.nspin_blk * kNB * kNB;
    hipLaunchKernelGGL(k_b_to_f32, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, st, nm,
                       d_g64, d_g32);
    hipLaunchKernelGGL(k_b_to_f32, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, st, nm,
                       d_p64, d_p32);
    if (e0) VCA_HIP(hipEventRecord(e0, st));
    if (lane8) {
      hipLaunchKernelGGL((k_logdet_lane8<kLane8Kpw, 64 / kLane8Kpw>), grid, dim3(kOBlock), 0, st, host.nk, nz,
                         zchunk, host.nspin_blk, d_hk32, d_g32, d_p32, d_part, phase);
    } else if (nbt == 8) {
      if (mfma)
        hipLaunchKernelGGL((k_logdet_mfma<8>), grid, dim3(kOBlock), 0, st, host.nk, nz, zchunk,
                           host.nspin_blk, d_hk32, d_g32, d_p32, d_part, phase);
      else
        hipLaunchKernelGGL((k_logdet<float, 8>), grid, dim3(kOBlock), 0, st, host.nk, nz,
                           zchunk, host.nspin_blk, d_hk32, d_g32, d_p32, d_part, phase);
    } else {
      if (mfma)
        hipLaunchKernelGGL((k_logdet_mfma<16>), grid, dim3(kOBlock), 0, st, host.nk, nz, zchunk,
                           host.nspin_blk, d_hk32, d_g32, d_p32, d_part, phase);
      else
        hipLaunchKernelGGL((k_logdet<float, 16>), grid, dim3(kOBlock), 0, st, host.nk, nz,
                           zchunk, host.nspin_blk, d_hk32, d_g32, d_p32, d_part, phase);
    }
  }
  VCA_HIP(hipGetLastError());
  if (e1) VCA_HIP(hipEventRecord(e1, st));
  hipLaunchKernelGGL(k_reduce_z, dim3(nz), dim3(64), 0, st, nz, host.nspin_blk, nkb, host.nk,
                     d_part, d_f);
  VCA_HIP(hipGetLastError());
  if (fim) {
    hipLaunchKernelGGL(k_reduce_phase, dim3(nz), dim3(64), 0, st, nz, host.nspin_blk, host.nk,
                       d_phase, d_fim);
    VCA_HIP(hipGetLastError());
    VCA_HIP(hipMemcpyAsync(fim, d_fim, sizeof(double) * nz, hipMemcpyDeviceToHost, st));
  }
  VCA_HIP(hipMemcpyAsync(f, d_f, sizeof(double) * nz, hipMemcpyDeviceToHost, st));
  VCA_HIP(hipStreamSynchronize(st));
  if (ms_logdet) {
    VCA_HIP(hipEventElapsedTime(ms_logdet, e0, e1));
    VCA_HIP(hipEventDestroy(e0));
    VCA_HIP(hipEventDestroy(e1));
  }
}

}  // namespace vca
