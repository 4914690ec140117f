// Lowest eigenpairs of a sector Hamiltonian: thick-restart Lanczos with
// full reorthogonalisation (the Hermitian form of Krylov-Schur), the
// equivalent of the ARPACK call SciFortran's sp_eigh makes for the
// reference at VCA_DIAG.f90:153-170 (which='SA', ncv = Nblock, maxiter =
// Nitermax, tol = lanc_tolerance, VCA_DIAG.f90:100-102).
//
// Device side: the basis V[0..m] (complex128, [m+1][dim]) and four
// memory-bound kernels -- a multi-dot that reads every basis vector once
// per group of G vectors, a fused w -= V h with |w|^2, a basis rotation
// V Y and a scale.  The Lanczos steps of one restart cycle run without a
// host round trip: every step orthogonalises twice (classical Gram-Schmidt
// twice, so no decision on the host), keeps its coefficients and norm in
// device memory and normalises by the norm read on the device.  Host side,
// once per cycle: the m x m projected matrix T = V^H H V from those
// coefficients, the breakdown check (a step whose norm fell to rounding is
// redone from a fresh direction, and the steps after it rerun), its
// eigenpairs (herm_eig) and the restart logic.
//
// Convergence test of a Ritz pair (theta, y): |beta_m| |y_m| <= tol *
// max(eps^(2/3), |theta|), ARPACK's dsconv criterion.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "vca_internal.hpp"

namespace vca {

namespace {

constexpr int kEBS = 256;  // lanes per workgroup
constexpr int kEG = 8;     // basis vectors per pass of the rotation
constexpr int kED = 16;    // basis vectors per pass of the multi-dot
constexpr int kEMaxBlocks = 1024;

__device__ inline double e_wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over the workgroup, valid in thread 0.
__device__ inline double e_block_sum(double v, double *s_part) {
  v = e_wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) s_part[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kEBS / 64; ++w) r += s_part[w];
  return r;
}

// Per-workgroup partial dots <V_j, w> for j in [j0, j0+nj), nj <= kED:
// pr[(2 j + c) * nblk + b], c = 0 real, 1 imaginary part.  With self != 0
// also |w|^2 into column self_col.  All 2 nj + 1 sums are reduced together:
// wavefront butterflies, then one LDS exchange across the 4 wavefronts
// (one barrier for the whole set).
__global__ __launch_bounds__(kEBS) void k_mdot(const double2 *V, int64_t dim, int j0, int nj,
                                               const double2 *w, int self_col, double *pr) {
  constexpr int NV = 2 * kED + 1;
  __shared__ double s_part[kEBS / 64][NV];
  double acc[NV];
#pragma unroll
  for (int g = 0; g < NV; ++g) acc[g] = 0;
  const int64_t stride = (int64_t)gridDim.x * kEBS;
  for (int64_t x = (int64_t)blockIdx.x * kEBS + threadIdx.x; x < dim; x += stride) {
    const double2 wx = w[x];
    acc[2 * kED] += wx.x * wx.x + wx.y * wx.y;
#pragma unroll
    for (int g = 0; g < kED; ++g)
      if (g < nj) {
        const double2 v = V[(int64_t)(j0 + g) * dim + x];
        acc[2 * g] += v.x * wx.x + v.y * wx.y;  // conj(v) * w
        acc[2 * g + 1] += v.x * wx.y - v.y * wx.x;
      }
  }
#pragma unroll
  for (int g = 0; g < NV; ++g) acc[g] = e_wave_sum(acc[g]);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int g = 0; g < NV; ++g) s_part[wid][g] = acc[g];
  __syncthreads();
  const int nblk = gridDim.x;
  const int t = threadIdx.x;
  if (t < 2 * nj || (t == 2 * kED && self_col >= 0)) {
    double s = 0;
    for (int wv = 0; wv < kEBS / 64; ++wv) s += s_part[wv][t];
    const int64_t col = t == 2 * kED ? self_col : 2 * j0 + t;
    pr[col * nblk + blockIdx.x] = s;
  }
}

// out[c] = sum_b pr[c * nblk + b] in block order (deterministic), one
// workgroup per column c.
__global__ __launch_bounds__(kEBS) void k_colsum(const double *pr, int nblk, double *out) {
  __shared__ double s_part[kEBS / 64];
  double s = 0;
  for (int b = threadIdx.x; b < nblk; b += kEBS) s += pr[(int64_t)blockIdx.x * nblk + b];
  s = e_block_sum(s, s_part);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// w -= sum_{j<n} h_j V_j with h = (re, im) pairs; pr[b] = partial |w|^2.
__global__ __launch_bounds__(kEBS) void k_gs_sub(const double2 *V, int64_t dim, int n,
                                                 const double *h, double2 *w, double *pr) {
  __shared__ double s_part[kEBS / 64];
  extern __shared__ double2 s_h[];
  for (int j = threadIdx.x; j < n; j += kEBS) s_h[j] = make_double2(h[2 * j], h[2 * j + 1]);
  __syncthreads();
  double ss = 0;
  const int64_t stride = (int64_t)gridDim.x * kEBS;
  for (int64_t x = (int64_t)blockIdx.x * kEBS + threadIdx.x; x < dim; x += stride) {
    double2 a = w[x];
    for (int j = 0; j < n; ++j) {
      const double2 v = V[(int64_t)j * dim + x];
      const double2 c = s_h[j];
      a.x -= c.x * v.x - c.y * v.y;
      a.y -= c.x * v.y + c.y * v.x;
    }
    w[x] = a;
    ss += a.x * a.x + a.y * a.y;
  }
  const double s = e_block_sum(ss, s_part);
  if (threadIdx.x == 0) pr[blockIdx.x] = s;
}

// out[i] = sum_{j<m} Y[j + (i0 + i) * ldy] V_j for i < ni <= kEG.
__global__ __launch_bounds__(kEBS) void k_rotate(const double2 *V, int64_t dim, int m,
                                                 const double2 *Y, int ldy, int i0, int ni,
                                                 double2 *out) {
  extern __shared__ double2 s_y[];  // [ni][m]
  for (int t = threadIdx.x; t < ni * m; t += kEBS) {
    const int i = t / m, j = t % m;
    s_y[t] = Y[j + (int64_t)(i0 + i) * ldy];
  }
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * kEBS;
  for (int64_t x = (int64_t)blockIdx.x * kEBS + threadIdx.x; x < dim; x += stride) {
    double ar[kEG], ai[kEG];
#pragma unroll
    for (int g = 0; g < kEG; ++g) ar[g] = ai[g] = 0;
    for (int j = 0; j < m; ++j) {
      const double2 v = V[(int64_t)j * dim + x];
#pragma unroll
      for (int g = 0; g < kEG; ++g)
        if (g < ni) {
          const double2 c = s_y[g * m + j];
          ar[g] += c.x * v.x - c.y * v.y;
          ai[g] += c.x * v.y + c.y * v.x;
        }
    }
#pragma unroll
    for (int g = 0; g < kEG; ++g)
      if (g < ni) out[(int64_t)g * dim + x] = make_double2(ar[g], ai[g]);
  }
}

__global__ __launch_bounds__(kEBS) void k_scale2(double2 *v, int64_t dim, double s) {
  const int64_t stride = (int64_t)gridDim.x * kEBS;
  for (int64_t x = (int64_t)blockIdx.x * kEBS + threadIdx.x; x < dim; x += stride) {
    double2 a = v[x];
    a.x *= s;
    a.y *= s;
    v[x] = a;
  }
}

// v *= 1/sqrt(*nrm2) with the squared norm read from device memory (left
// as is when it is not positive: the host redoes such a step).
__global__ __launch_bounds__(kEBS) void k_scale_nrm(double2 *v, int64_t dim, const double *nrm2) {
  const double n2 = *nrm2;
  if (!(n2 > 0.0)) return;
  const double s = 1.0 / sqrt(n2);
  const int64_t stride = (int64_t)gridDim.x * kEBS;
  for (int64_t x = (int64_t)blockIdx.x * kEBS + threadIdx.x; x < dim; x += stride) {
    double2 a = v[x];
    a.x *= s;
    a.y *= s;
    v[x] = a;
  }
}

// Deterministic pseudo-random vector: splitmix64 of (seed, i) in [-1, 1).
__global__ __launch_bounds__(kEBS) void k_hash_vec(double2 *v, int64_t dim, uint64_t seed) {
  const int64_t stride = (int64_t)gridDim.x * kEBS;
  for (int64_t x = (int64_t)blockIdx.x * kEBS + threadIdx.x; x < dim; x += stride) {
    const uint64_t g = 0x9E3779B97F4A7C15ull;
    uint64_t z = seed * g + ((uint64_t)x + 1ull) * g;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    v[x] = make_double2((double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0, 0.0);
  }
}

int eig_grid(int64_t dim) {
  int64_t g = (dim + kEBS * 4 - 1) / (kEBS * 4);
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, kEMaxBlocks));
}

// Device workspace of one eigensolve.
struct Work {
  Sector &sec;
  int64_t dim;
  int m;      // basis size (ncv)
  int nblk;
  double2 *V = nullptr;  // [m+1][dim]
  double2 *W = nullptr;  // [m][dim] rotation target
  double *pr = nullptr;  // [(2(m+1)+1)][nblk]
  double *red = nullptr; // [2(m+1)+2]
  double2 *Y = nullptr;  // [m][m]
  // per step j of a cycle: the coefficients of both Gram-Schmidt passes,
  // [j][pass][2(m+1)], and the squared norm after them, [j]
  double *coef = nullptr, *nrm2 = nullptr;
  int ldc() const { return 2 * (m + 1); }
  explicit Work(Sector &s, int m_) : sec(s), dim(s.dim), m(m_), nblk(eig_grid(s.dim)) {
    const hipStream_t st = s.stream;
    V = static_cast<double2 *>(dev_alloc_async(sizeof(double2) * (size_t)(m + 1) * dim, st));
    W = static_cast<double2 *>(dev_alloc_async(sizeof(double2) * (size_t)m * dim, st));
    pr = static_cast<double *>(
        dev_alloc_async(sizeof(double) * (size_t)(2 * (m + 1) + 1) * nblk, st));
    red = static_cast<double *>(dev_alloc_async(sizeof(double) * (2 * (m + 1) + 2), st));
    Y = static_cast<double2 *>(dev_alloc_async(sizeof(double2) * (size_t)m * m, st));
    coef = static_cast<double *>(dev_alloc_async(sizeof(double) * (size_t)m * 2 * ldc(), st));
    nrm2 = static_cast<double *>(dev_alloc_async(sizeof(double) * (size_t)m, st));
  }
  // stream-ordered frees: every use of the buffers is on sec.stream
  ~Work() {
    for (void *q : {(void *)V, (void *)W, (void *)pr, (void *)red, (void *)Y, (void *)coef,
                    (void *)nrm2})
      if (q) (void)hipFreeAsync(q, sec.stream);
  }
  double2 *vec(int j) const { return V + (size_t)j * dim; }

  // h[j] = <V_j, w> for j < n into dst[0..2n), |w|^2 into dst[2n] if self
  void mdot(int n, const double2 *w, bool self, double *dst = nullptr) {
    const hipStream_t st = sec.stream;
    if (!dst) dst = red;
    for (int j0 = 0; j0 < n; j0 += kED) {
      const int nj = std::min(kED, n - j0);
      hipLaunchKernelGGL(k_mdot, dim3(nblk), dim3(kEBS), 0, st, V, dim, j0, nj, w,
                         (self && j0 == 0) ? 2 * n : -1, pr);
    }
    if (n == 0 && self)
      hipLaunchKernelGGL(k_mdot, dim3(nblk), dim3(kEBS), 0, st, V, dim, 0, 0, w, 0, pr);
    hipLaunchKernelGGL(k_colsum, dim3(2 * n + (self ? 1 : 0)), dim3(kEBS), 0, st, pr, nblk, dst);
    VCA_HIP(hipGetLastError());
  }
  // w -= V[0..n) h with h = coefs[0..2n) (default red), |w|^2 into *nrm_dst
  // (default red[2n+1])
  void sub(int n, double2 *w, const double *coefs = nullptr, double *nrm_dst = nullptr) {
    const hipStream_t st = sec.stream;
    if (!coefs) coefs = red;
    if (!nrm_dst) nrm_dst = red + 2 * n + 1;
    hipLaunchKernelGGL(k_gs_sub, dim3(nblk), dim3(kEBS), sizeof(double2) * std::max(1, n), st,
                       V, dim, n, coefs, w, pr);
    hipLaunchKernelGGL(k_colsum, dim3(1), dim3(kEBS), 0, st, pr, nblk, nrm_dst);
    VCA_HIP(hipGetLastError());
  }
  // One Lanczos step without a host round trip: w = H V_j, two classical
  // Gram-Schmidt passes against V[0..j] (coefficients to coef[j]), |w|^2 to
  // nrm2[j], then w /= |w| on the device.
  void device_step(int j) {
    double2 *w = vec(j + 1);
    launch_hxv(sec, vec(j), vec(j), w, false);
    for (int pass = 0; pass < 2; ++pass) {
      double *c = coef + ((size_t)j * 2 + pass) * ldc();
      mdot(j + 1, w, false, c);
      sub(j + 1, w, c, nrm2 + j);
    }
    hipLaunchKernelGGL(k_scale_nrm, dim3(nblk), dim3(kEBS), 0, sec.stream, w, dim, nrm2 + j);
    VCA_HIP(hipGetLastError());
  }
  // out[i] = V[0..mm) Y[:, i0 + i] for i < ni (Y already on the device)
  void rotate(int mm, int ldy, int i0, int ni, double2 *out) {
    for (int g0 = 0; g0 < ni; g0 += kEG) {
      const int n = std::min(kEG, ni - g0);
      hipLaunchKernelGGL(k_rotate, dim3(nblk), dim3(kEBS), sizeof(double2) * n * mm,
                         sec.stream, V, dim, mm, Y, ldy, i0 + g0, n, out + (size_t)g0 * dim);
    }
    VCA_HIP(hipGetLastError());
  }
  void scale(double2 *v, double s) {
    hipLaunchKernelGGL(k_scale2, dim3(nblk), dim3(kEBS), 0, sec.stream, v, dim, s);
    VCA_HIP(hipGetLastError());
  }
  void host(double *dst, int n) {
    VCA_HIP(hipMemcpyAsync(dst, red, sizeof(double) * n, hipMemcpyDeviceToHost, sec.stream));
    VCA_HIP(hipStreamSynchronize(sec.stream));
  }

  // Orthogonalise w against V[0..n) (classical Gram-Schmidt with one DGKS
  // correction when the norm drops below 1/sqrt(2) of its value), adds the
  // coefficients to h[0..n) and returns |w|.
  double orthogonalise(int n, double2 *w, zc *h) {
    std::vector<double> r(2 * n + 2);
    mdot(n, w, true);
    sub(n, w);
    host(r.data(), 2 * n + 2);
    for (int j = 0; j < n; ++j) h[j] += zc(r[2 * j], r[2 * j + 1]);
    const double before = std::sqrt(std::max(0.0, r[2 * n]));
    double after = std::sqrt(std::max(0.0, r[2 * n + 1]));
    if (n > 0 && after < 0.7071067811865476 * before) {
      mdot(n, w, false);
      sub(n, w);
      host(r.data(), 2 * n + 2);
      for (int j = 0; j < n; ++j) h[j] += zc(r[2 * j], r[2 * j + 1]);
      after = std::sqrt(std::max(0.0, r[2 * n + 1]));
    }
    return after;
  }
};

}  // namespace

// nev lowest eigenpairs of the FP64 sector sec.  start: device complex128
// start vector or null for the hash start vector (seed 5001, the start
// vector of the ground-state Lanczos).  Eigenvectors are returned as
// separate device allocations (caller frees), normalised.
EigsOut eigs_lanczos(Sector &sec, int nev, int ncv, double tol, int maxiter, const void *start,
                     std::vector<double> &evals, std::vector<void *> &evecs) {
  if (sec.precision != VCA_FP64) fail(VCA_ERR_ARG, "eigs: the sector must be FP64");
  const int64_t dim = sec.dim;
  if (nev < 1 || nev >= dim) fail(VCA_ERR_ARG, "eigs: nev=%d out of range for dim %lld", nev,
                                 (long long)dim);
  int m = (int)std::min<int64_t>(std::max(ncv, nev + 2), dim);
  if (m <= nev) fail(VCA_ERR_ARG, "eigs: basis size %d must exceed nev=%d", m, nev);
  Work wk(sec, m);
  const hipStream_t st = sec.stream;
  const double eps23 = std::pow(2.220446049250313e-16, 2.0 / 3.0);

  // v_0
  if (start)
    VCA_HIP(hipMemcpyAsync(wk.vec(0), start, sizeof(double2) * dim, hipMemcpyDeviceToDevice, st));
  else
    hipLaunchKernelGGL(k_hash_vec, dim3(wk.nblk), dim3(kEBS), 0, st, wk.vec(0), dim,
                       (uint64_t)5001);
  {
    std::vector<zc> h0(1);
    const double n0 = wk.orthogonalise(0, wk.vec(0), h0.data());
    if (!(n0 > 0)) fail(VCA_ERR_ARG, "eigs: zero start vector");
    wk.scale(wk.vec(0), 1.0 / n0);
  }

  std::vector<zc> T((size_t)m * m, zc(0.0));  // column-major projected matrix
  auto t_at = [&](int i, int j) -> zc & { return T[(size_t)i + (size_t)j * m]; };
  int k = 0;  // vectors kept by the last restart
  double beta_m = 0;
  std::vector<double> theta(m);
  std::vector<zc> Yh((size_t)m * m);
  EigsOut out{};
  uint64_t reseed = 5002;
  const int ldc = wk.ldc();
  std::vector<double> hc((size_t)m * 2 * ldc), hn(m);
  for (int it = 0; it < std::max(1, maxiter); ++it) {
    out.restarts = it;
    int j0 = k;
    while (j0 < m) {
      // steps j0..m-1 on the device, then their coefficients in one copy
      for (int j = j0; j < m; ++j) wk.device_step(j);
      out.hxv += m - j0;
      VCA_HIP(hipMemcpyAsync(hc.data() + (size_t)j0 * 2 * ldc, wk.coef + (size_t)j0 * 2 * ldc,
                             sizeof(double) * (size_t)(m - j0) * 2 * ldc,
                             hipMemcpyDeviceToHost, st));
      VCA_HIP(hipMemcpyAsync(hn.data() + j0, wk.nrm2 + j0, sizeof(double) * (m - j0),
                             hipMemcpyDeviceToHost, st));
      VCA_HIP(hipStreamSynchronize(st));
      int next = m;
      for (int j = j0; j < m; ++j) {
        const double *c1 = hc.data() + ((size_t)j * 2) * ldc;
        const double *c2 = c1 + ldc;
        double tnorm = 0;
        for (int i = 0; i <= j; ++i) {
          const zc h(c1[2 * i] + c2[2 * i], c1[2 * i + 1] + c2[2 * i + 1]);
          t_at(i, j) = h;
          t_at(j, i) = std::conj(h);
          tnorm = std::max(tnorm, std::abs(h));
        }
        t_at(j, j) = zc(t_at(j, j).real(), 0.0);
        double b = std::sqrt(std::max(0.0, hn[j]));
        const bool breakdown = b <= 1e-14 * std::max(1.0, tnorm);
        if (breakdown) b = 0;
        if (j + 1 < m) {
          t_at(j + 1, j) = zc(b, 0.0);
          t_at(j, j + 1) = zc(b, 0.0);
        } else {
          beta_m = b;
        }
        if (breakdown) {
          // invariant subspace: V_{j+1} from a fresh direction orthogonal to
          // V, and the steps after it again
          if (j + 1 < m) {
            double2 *w = wk.vec(j + 1);
            hipLaunchKernelGGL(k_hash_vec, dim3(wk.nblk), dim3(kEBS), 0, st, w, dim, reseed++);
            std::vector<zc> hz(j + 1, zc(0.0));
            const double nz = wk.orthogonalise(j + 1, w, hz.data());
            if (!(nz > 0)) fail(VCA_ERR_NOCONV, "eigs: could not extend the basis");
            wk.scale(w, 1.0 / nz);
            next = j + 1;
          }
          break;
        }
      }
      j0 = next;
    }
    Yh = T;
    herm_eig(m, Yh.data(), theta.data());
    int nconv = 0;
    for (int i = 0; i < nev; ++i) {
      const double res = beta_m * std::abs(Yh[(size_t)(m - 1) + (size_t)i * m]);
      if (res <= tol * std::max(eps23, std::fabs(theta[i]))) ++nconv;
    }
    out.nconv = nconv;
    if (nconv == nev || it + 1 >= std::max(1, maxiter)) break;
    // thick restart: keep the k lowest Ritz vectors and the residual vector
    k = std::min(m - 2, nev + (m - nev) / 3);
    k = std::max(k, nev);
    VCA_HIP(hipMemcpyAsync(wk.Y, Yh.data(), sizeof(double2) * (size_t)m * m,
                           hipMemcpyHostToDevice, st));
    wk.rotate(m, m, 0, k, wk.W);
    VCA_HIP(hipMemcpyAsync(wk.V, wk.W, sizeof(double2) * (size_t)k * dim,
                           hipMemcpyDeviceToDevice, st));
    VCA_HIP(hipMemcpyAsync(wk.vec(k), wk.vec(m), sizeof(double2) * dim,
                           hipMemcpyDeviceToDevice, st));
    std::fill(T.begin(), T.end(), zc(0.0));
    for (int i = 0; i < k; ++i) t_at(i, i) = zc(theta[i], 0.0);
    if (beta_m == 0) {
      // exact invariant subspace found with unconverged pairs: fresh direction
      hipLaunchKernelGGL(k_hash_vec, dim3(wk.nblk), dim3(kEBS), 0, st, wk.vec(k), dim,
                         reseed++);
      std::vector<zc> hz(k, zc(0.0));
      const double nz = wk.orthogonalise(k, wk.vec(k), hz.data());
      if (!(nz > 0)) fail(VCA_ERR_NOCONV, "eigs: could not extend the basis");
      wk.scale(wk.vec(k), 1.0 / nz);
    }
  }
  // eigenvectors
  VCA_HIP(hipMemcpyAsync(wk.Y, Yh.data(), sizeof(double2) * (size_t)m * m,
                         hipMemcpyHostToDevice, st));
  wk.rotate(m, m, 0, nev, wk.W);
  evals.assign(theta.begin(), theta.begin() + nev);
  evecs.clear();
  for (int i = 0; i < nev; ++i) {
    void *d = dev_alloc_async(sizeof(double2) * dim, st);
    VCA_HIP(hipMemcpyAsync(d, wk.W + (size_t)i * dim, sizeof(double2) * dim,
                           hipMemcpyDeviceToDevice, st));
    evecs.push_back(d);
  }
  VCA_HIP(hipStreamSynchronize(st));
  return out;
}

}  // namespace vca
