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
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "rows.hpp"
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
  hipStream_t st;  // every launch and copy of the solve
  explicit Work(Sector &s, int m_, hipStream_t stream = nullptr)
      : sec(s), dim(s.dim), m(m_), nblk(eig_grid(s.dim)), st(stream ? stream : s.stream) {
    V = static_cast<double2 *>(dev_alloc_async(sizeof(double2) * (size_t)(m + 1) * dim, st));
    W = static_cast<double2 *>(dev_alloc_async(sizeof(double2) * (size_t)m * dim, st));
    pr = static_cast<double *>(
        dev_alloc_async(sizeof(double) * (size_t)(2 * (m + 1) + 1) * nblk, st));
    red = static_cast<double *>(dev_alloc_async(sizeof(double) * (2 * (m + 1) + 2), st));
    Y = static_cast<double2 *>(dev_alloc_async(sizeof(double2) * (size_t)m * m, st));
    coef = static_cast<double *>(dev_alloc_async(sizeof(double) * (size_t)m * 2 * ldc(), st));
    nrm2 = static_cast<double *>(dev_alloc_async(sizeof(double) * (size_t)m, st));
  }
  // stream-ordered frees: every use of the buffers is on st
  ~Work() {
    for (void *q : {(void *)V, (void *)W, (void *)pr, (void *)red, (void *)Y, (void *)coef,
                    (void *)nrm2})
      if (q) (void)hipFreeAsync(q, st);
  }
  double2 *vec(int j) const { return V + (size_t)j * dim; }

  // h[j] = <V_j, w> for j < n into dst[0..2n), |w|^2 into dst[2n] if self
  void mdot(int n, const double2 *w, bool self, double *dst = nullptr) {
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
    hipLaunchKernelGGL(k_scale_nrm, dim3(nblk), dim3(kEBS), 0, st, w, dim, nrm2 + j);
    VCA_HIP(hipGetLastError());
  }
  // out[i] = V[0..mm) Y[:, i0 + i] for i < ni (Y already on the device)
  void rotate(int mm, int ldy, int i0, int ni, double2 *out) {
    for (int g0 = 0; g0 < ni; g0 += kEG) {
      const int n = std::min(kEG, ni - g0);
      hipLaunchKernelGGL(k_rotate, dim3(nblk), dim3(kEBS), sizeof(double2) * n * mm, st, V,
                         dim, mm, Y, ldy, i0 + g0, n, out + (size_t)g0 * dim);
    }
    VCA_HIP(hipGetLastError());
  }
  void scale(double2 *v, double s) {
    hipLaunchKernelGGL(k_scale2, dim3(nblk), dim3(kEBS), 0, st, v, dim, s);
    VCA_HIP(hipGetLastError());
  }
  void host(double *dst, int n) {
    VCA_HIP(hipMemcpyAsync(dst, red, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    VCA_HIP(hipStreamSynchronize(st));
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

// One thick-restart eigensolve as a state machine, so that a single solve
// (eigs_lanczos) and a batch of solves whose Lanczos steps share launches
// (eigs_lanczos_batch) run the same host logic: the steps j0..m-1 of a
// restart cycle run on the device (either way), their coefficients come to
// the host, absorb() builds the projected matrix and handles breakdowns,
// and end_cycle() takes the Ritz pairs and restarts or finishes.
struct Trl {
  Sector &sec;
  int nev, m;
  double tol;
  int maxiter;
  Work wk;
  double eps23;
  std::vector<zc> T;  // column-major projected matrix
  int k = 0;          // vectors kept by the last restart
  int j0 = 0;         // first step of the cycle still to run
  int it = 0;         // restart cycle
  double beta_m = 0;
  std::vector<double> theta;
  std::vector<zc> Yh;
  EigsOut out{};
  uint64_t reseed = 5002;
  std::vector<double> hc, hn;
  bool done = false;

  Trl(Sector &s, int nev_, int ncv, double tol_, int maxiter_, hipStream_t st)
      : sec(s), nev(nev_), m(0), tol(tol_), maxiter(maxiter_),
        wk(s, basis_size(s, nev_, ncv), st), eps23(std::pow(2.220446049250313e-16, 2.0 / 3.0)) {
    m = wk.m;
    T.assign((size_t)m * m, zc(0.0));
    theta.resize(m);
    Yh.resize((size_t)m * m);
    hc.resize((size_t)m * 2 * wk.ldc());
    hn.resize(m);
  }
  static int basis_size(const Sector &s, int nev, int ncv) {
    if (s.precision != VCA_FP64) fail(VCA_ERR_ARG, "eigs: the sector must be FP64");
    if (nev < 1 || nev >= s.dim)
      fail(VCA_ERR_ARG, "eigs: nev=%d out of range for dim %lld", nev, (long long)s.dim);
    const int m = (int)std::min<int64_t>(std::max(ncv, nev + 2), s.dim);
    if (m <= nev) fail(VCA_ERR_ARG, "eigs: basis size %d must exceed nev=%d", m, nev);
    return m;
  }
  zc &t_at(int i, int j) { return T[(size_t)i + (size_t)j * m]; }

  // v_0: the start vector (device complex128) or the hash vector of seed 5001
  void start(const void *v0) {
    const hipStream_t st = wk.st;
    if (v0)
      VCA_HIP(hipMemcpyAsync(wk.vec(0), v0, sizeof(double2) * wk.dim, hipMemcpyDeviceToDevice,
                             st));
    else
      hipLaunchKernelGGL(k_hash_vec, dim3(wk.nblk), dim3(kEBS), 0, st, wk.vec(0), wk.dim,
                         (uint64_t)5001);
    std::vector<zc> h0(1);
    const double n0 = wk.orthogonalise(0, wk.vec(0), h0.data());
    if (!(n0 > 0)) fail(VCA_ERR_ARG, "eigs: zero start vector");
    wk.scale(wk.vec(0), 1.0 / n0);
  }
  // queue the copies of steps j0..m-1's coefficients to the host
  void fetch() {
    const int ldc = wk.ldc();
    VCA_HIP(hipMemcpyAsync(hc.data() + (size_t)j0 * 2 * ldc, wk.coef + (size_t)j0 * 2 * ldc,
                           sizeof(double) * (size_t)(m - j0) * 2 * ldc, hipMemcpyDeviceToHost,
                           wk.st));
    VCA_HIP(hipMemcpyAsync(hn.data() + j0, wk.nrm2 + j0, sizeof(double) * (m - j0),
                           hipMemcpyDeviceToHost, wk.st));
  }
  // after the steps and fetch() (synchronised): the projected matrix; a
  // step whose norm fell to rounding gets a fresh direction and the steps
  // after it run again.  Returns true when the cycle's basis is complete.
  bool absorb() {
    const int ldc = wk.ldc();
    out.hxv += m - j0;
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
          hipLaunchKernelGGL(k_hash_vec, dim3(wk.nblk), dim3(kEBS), 0, wk.st, w, wk.dim,
                             reseed++);
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
    return j0 >= m;
  }
  // The cycle's Ritz pairs: converged (or out of restarts) -> done, else
  // the thick restart keeps the k lowest Ritz vectors and the residual
  // vector.  end_cycle_host() takes the decision and prepares the projected
  // matrix (true: restart); restart_device() rotates the basis,
  // restart_fresh() handles an exact invariant subspace and starts the next
  // cycle.
  bool end_cycle_host() {
    out.restarts = it;
    Yh = T;
    herm_eig(m, Yh.data(), theta.data());
    int nconv = 0;
    for (int i = 0; i < nev; ++i) {
      const double res = beta_m * std::abs(Yh[(size_t)(m - 1) + (size_t)i * m]);
      if (res <= tol * std::max(eps23, std::fabs(theta[i]))) ++nconv;
    }
    out.nconv = nconv;
    if (nconv == nev || it + 1 >= std::max(1, maxiter)) {
      done = true;
      return false;
    }
    k = std::min(m - 2, nev + (m - nev) / 3);
    k = std::max(k, nev);
    std::fill(T.begin(), T.end(), zc(0.0));
    for (int i = 0; i < k; ++i) t_at(i, i) = zc(theta[i], 0.0);
    return true;
  }
  void restart_device() {
    const hipStream_t st = wk.st;
    VCA_HIP(hipMemcpyAsync(wk.Y, Yh.data(), sizeof(double2) * (size_t)m * m,
                           hipMemcpyHostToDevice, st));
    wk.rotate(m, m, 0, k, wk.W);
    VCA_HIP(hipMemcpyAsync(wk.V, wk.W, sizeof(double2) * (size_t)k * wk.dim,
                           hipMemcpyDeviceToDevice, st));
    VCA_HIP(hipMemcpyAsync(wk.vec(k), wk.vec(m), sizeof(double2) * wk.dim,
                           hipMemcpyDeviceToDevice, st));
  }
  void restart_fresh() {
    if (beta_m == 0) {
      // exact invariant subspace found with unconverged pairs: fresh direction
      hipLaunchKernelGGL(k_hash_vec, dim3(wk.nblk), dim3(kEBS), 0, wk.st, wk.vec(k), wk.dim,
                         reseed++);
      std::vector<zc> hz(k, zc(0.0));
      const double nz = wk.orthogonalise(k, wk.vec(k), hz.data());
      if (!(nz > 0)) fail(VCA_ERR_NOCONV, "eigs: could not extend the basis");
      wk.scale(wk.vec(k), 1.0 / nz);
    }
    j0 = k;
    ++it;
  }
  void end_cycle() {
    if (!end_cycle_host()) return;
    restart_device();
    restart_fresh();
  }
  // the nev Ritz vectors: rotated into W, then copied to separate device
  // allocations (caller frees)
  void rotate_eigenvectors() {
    VCA_HIP(hipMemcpyAsync(wk.Y, Yh.data(), sizeof(double2) * (size_t)m * m,
                           hipMemcpyHostToDevice, wk.st));
    wk.rotate(m, m, 0, nev, wk.W);
  }
  void copy_eigenvectors(std::vector<double> &evals, std::vector<void *> &evecs) {
    const hipStream_t st = wk.st;
    evals.assign(theta.begin(), theta.begin() + nev);
    evecs.clear();
    for (int i = 0; i < nev; ++i) {
      void *d = dev_alloc_async(sizeof(double2) * wk.dim, st);
      VCA_HIP(hipMemcpyAsync(d, wk.W + (size_t)i * wk.dim, sizeof(double2) * wk.dim,
                             hipMemcpyDeviceToDevice, st));
      evecs.push_back(d);
    }
  }
  void eigenvectors(std::vector<double> &evals, std::vector<void *> &evecs) {
    rotate_eigenvectors();
    copy_eigenvectors(evals, evecs);
    VCA_HIP(hipStreamSynchronize(wk.st));
  }
};

// ------------------------------------------------------------------------
// Batched steps: one launch per kernel of a Lanczos step covers every solve
// of the batch, each at its own step j (jtab, -1: idle this step) on its
// own sector.  The vector kernels keep each solve's own grid (nblk
// workgroups striding over its vector, as the single-solve kernels), so the
// Gram-Schmidt sums are the single solve's; H x v is the row-per-lane form
// of rows.hpp.
struct EbJob {
  RowTab tab;
  double2 *V;     // [m+1][dim]
  double2 *W;     // [m][dim] rotation target
  int64_t dim;
  int m, nblk, ldc;
  double *pr;     // [(2(m+1)+1)][nblk]
  double *coef;   // [m][2][ldc]
  double *nrm2;   // [m]
};

struct EbArgs {
  const EbJob *jobs;
  const int2 *items;  // (job, index in the job)
  const int *jtab;    // step of each job, -1: idle
  int pass;           // Gram-Schmidt pass
  CrossTerm cross;
};

// V_{j+1} = H V_j, one row per lane.
template <bool REAL_AMP>
__global__ __launch_bounds__(kEBS) void k_eb_hxv(const EbArgs a) {
  extern __shared__ __align__(16) unsigned char eb_smem[];
  const int2 it = a.items[blockIdx.x];
  const int j = a.jtab[it.x];
  if (j < 0) return;
  const EbJob &jb = a.jobs[it.x];
  const RowTab tab = jb.tab;
  double2 *s_amu = reinterpret_cast<double2 *>(eb_smem);
  double2 *s_amd = s_amu + tab.namp_up;
  rows::stage_amps<double, kEBS>(tab, s_amu, s_amd);
  __syncthreads();
  const int64_t r = (int64_t)it.y * kEBS + threadIdx.x;
  if (r >= jb.dim) return;
  const int idw = (int)(r / tab.dimup);
  const int iup = (int)(r - (int64_t)idw * tab.dimup);
  RowMeta mu;
  const double ed = rows::diag(tab, a.cross, iup, idw, mu);
  const double2 *vin = jb.V + (int64_t)j * jb.dim;
  const double2 *col = vin + (int64_t)idw * tab.dimup;
  const double2 x = col[iup];
  double ar = ed * x.x, ai = ed * x.y;
  rows::hops<double, REAL_AMP, double2>(tab, vin, col, iup, idw, mu.cnt, tab.cnt_dw[idw], s_amu,
                                        s_amd, ar, ai);
  jb.V[(int64_t)(j + 1) * jb.dim + r] = make_double2(ar, ai);
}

// Partial dots <V_i, w>, w = V_{j+1}, for the group of kED basis vectors
// i0 = kED g ...: k_mdot's sums, block b of the job's grid.
__global__ __launch_bounds__(kEBS) void k_eb_mdot(const EbArgs a) {
  constexpr int NV = 2 * kED;
  __shared__ double s_part[kEBS / 64][NV];
  const int2 it = a.items[blockIdx.x];
  const int j = a.jtab[it.x];
  if (j < 0) return;
  const EbJob &jb = a.jobs[it.x];
  const int nblk = jb.nblk;
  const int g = it.y / nblk, b = it.y - g * nblk;
  const int n = j + 1, j0 = kED * g;
  if (j0 >= n) return;
  const int nj = min(kED, n - j0);
  const int64_t dim = jb.dim;
  const double2 *V = jb.V;
  const double2 *w = V + (int64_t)(j + 1) * dim;
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0;
  const int64_t stride = (int64_t)nblk * kEBS;
  for (int64_t x = (int64_t)b * kEBS + threadIdx.x; x < dim; x += stride) {
    const double2 wx = w[x];
#pragma unroll
    for (int q = 0; q < kED; ++q)
      if (q < nj) {
        const double2 v = V[(int64_t)(j0 + q) * dim + x];
        acc[2 * q] += v.x * wx.x + v.y * wx.y;  // conj(v) * w
        acc[2 * q + 1] += v.x * wx.y - v.y * wx.x;
      }
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = e_wave_sum(acc[q]);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) s_part[wid][q] = acc[q];
  __syncthreads();
  const int t = threadIdx.x;
  if (t < 2 * nj) {
    double s = 0;
    for (int wv = 0; wv < kEBS / 64; ++wv) s += s_part[wv][t];
    jb.pr[(int64_t)(2 * j0 + t) * nblk + b] = s;
  }
}

// coef[j][pass][c] = sum_b pr[c][b] for c < 2 (j + 1), in block order.
__global__ __launch_bounds__(kEBS) void k_eb_colsum(const EbArgs a) {
  __shared__ double s_part[kEBS / 64];
  const int2 it = a.items[blockIdx.x];
  const int j = a.jtab[it.x];
  if (j < 0) return;
  const EbJob &jb = a.jobs[it.x];
  const int c = it.y;
  if (c >= 2 * (j + 1)) return;
  double s = 0;
  for (int b = threadIdx.x; b < jb.nblk; b += kEBS) s += jb.pr[(int64_t)c * jb.nblk + b];
  s = e_block_sum(s, s_part);
  if (threadIdx.x == 0) jb.coef[((int64_t)j * 2 + a.pass) * jb.ldc + c] = s;
}

// w -= sum_{i<=j} h_i V_i with the pass's coefficients; pr[b] = partial
// |w|^2 (k_gs_sub's sums).
__global__ __launch_bounds__(kEBS) void k_eb_sub(const EbArgs a) {
  __shared__ double s_part[kEBS / 64];
  extern __shared__ double2 s_h[];
  const int2 it = a.items[blockIdx.x];
  const int j = a.jtab[it.x];
  if (j < 0) return;
  const EbJob &jb = a.jobs[it.x];
  const int n = j + 1, b = it.y;
  const double *h = jb.coef + ((int64_t)j * 2 + a.pass) * jb.ldc;
  for (int q = threadIdx.x; q < n; q += kEBS) s_h[q] = make_double2(h[2 * q], h[2 * q + 1]);
  __syncthreads();
  const int64_t dim = jb.dim;
  double2 *w = jb.V + (int64_t)(j + 1) * dim;
  double ss = 0;
  const int64_t stride = (int64_t)jb.nblk * kEBS;
  for (int64_t x = (int64_t)b * kEBS + threadIdx.x; x < dim; x += stride) {
    double2 y = w[x];
    for (int q = 0; q < n; ++q) {
      const double2 v = jb.V[(int64_t)q * dim + x];
      const double2 c = s_h[q];
      y.x -= c.x * v.x - c.y * v.y;
      y.y -= c.x * v.y + c.y * v.x;
    }
    w[x] = y;
    ss += y.x * y.x + y.y * y.y;
  }
  const double s = e_block_sum(ss, s_part);
  if (threadIdx.x == 0) jb.pr[b] = s;
}

// nrm2[j] = sum_b pr[b], one workgroup per job.
__global__ __launch_bounds__(kEBS) void k_eb_nrm(const EbArgs a) {
  __shared__ double s_part[kEBS / 64];
  const int2 it = a.items[blockIdx.x];
  const int j = a.jtab[it.x];
  if (j < 0) return;
  const EbJob &jb = a.jobs[it.x];
  double s = 0;
  for (int b = threadIdx.x; b < jb.nblk; b += kEBS) s += jb.pr[b];
  s = e_block_sum(s, s_part);
  if (threadIdx.x == 0) jb.nrm2[j] = s;
}

// V_{j+1} /= sqrt(nrm2[j]) (left as is when not positive).
__global__ __launch_bounds__(kEBS) void k_eb_scale(const EbArgs a) {
  const int2 it = a.items[blockIdx.x];
  const int j = a.jtab[it.x];
  if (j < 0) return;
  const EbJob &jb = a.jobs[it.x];
  const double n2 = jb.nrm2[j];
  if (!(n2 > 0.0)) return;
  const double sc = 1.0 / sqrt(n2);
  double2 *w = jb.V + (int64_t)(j + 1) * jb.dim;
  const int64_t stride = (int64_t)jb.nblk * kEBS;
  for (int64_t x = (int64_t)it.y * kEBS + threadIdx.x; x < jb.dim; x += stride) {
    double2 y = w[x];
    y.x *= sc;
    y.y *= sc;
    w[x] = y;
  }
}

// Basis rotation of the jobs that restart or finish this round: W[i] =
// V[0..m) Y[:, i] for i < rtab[job] (k_rotate's sums), Y from the pooled
// per-job matrices (ldy = m).
__global__ __launch_bounds__(kEBS) void k_eb_rotate(const EbArgs a, const int *rtab,
                                                    const double2 *Ypool, int64_t ystride) {
  extern __shared__ double2 s_y[];  // [ni][m]
  const int2 it = a.items[blockIdx.x];
  const int nout = rtab[it.x];
  if (nout <= 0) return;
  const EbJob &jb = a.jobs[it.x];
  const int nblk = jb.nblk, m = jb.m;
  const int g = it.y / nblk, b = it.y - g * nblk;
  const int i0 = kEG * g;
  if (i0 >= nout) return;
  const int ni = min(kEG, nout - i0);
  const double2 *Y = Ypool + it.x * ystride;
  for (int t = threadIdx.x; t < ni * m; t += kEBS) {
    const int i = t / m, j = t % m;
    s_y[t] = Y[j + (int64_t)(i0 + i) * m];
  }
  __syncthreads();
  const int64_t dim = jb.dim;
  const int64_t stride = (int64_t)nblk * kEBS;
  for (int64_t x = (int64_t)b * kEBS + threadIdx.x; x < dim; x += stride) {
    double ar[kEG], ai[kEG];
#pragma unroll
    for (int q = 0; q < kEG; ++q) ar[q] = ai[q] = 0;
    for (int j = 0; j < m; ++j) {
      const double2 v = jb.V[(int64_t)j * dim + x];
#pragma unroll
      for (int q = 0; q < kEG; ++q)
        if (q < ni) {
          const double2 c = s_y[q * m + j];
          ar[q] += c.x * v.x - c.y * v.y;
          ai[q] += c.x * v.y + c.y * v.x;
        }
    }
#pragma unroll
    for (int q = 0; q < kEG; ++q)
      if (q < ni) jb.W[(int64_t)(i0 + q) * dim + x] = make_double2(ar[q], ai[q]);
  }
}

// Restart copies: V[i] = W[i] for i < k, then V[k] = V[m] (ktab[job] = k,
// -1: no restart this round).
__global__ __launch_bounds__(kEBS) void k_eb_restart(const EbArgs a, const int *ktab) {
  const int2 it = a.items[blockIdx.x];
  const int k = ktab[it.x];
  if (k < 0) return;
  const EbJob &jb = a.jobs[it.x];
  const int64_t dim = jb.dim;
  const int64_t stride = (int64_t)jb.nblk * kEBS;
  for (int64_t x = (int64_t)it.y * kEBS + threadIdx.x; x < dim; x += stride) {
    for (int i = 0; i < k; ++i) jb.V[(int64_t)i * dim + x] = jb.W[(int64_t)i * dim + x];
    jb.V[(int64_t)k * dim + x] = jb.V[(int64_t)jb.m * dim + x];
  }
}

// fn(0..n-1) on up to 16 host threads (static slices); fn writes only
// per-index state.
template <typename Fn>
void eb_parallel(int n, Fn &&fn) {
  const int nth = std::min<int>(
      n, (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency())));
  if (nth <= 1) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nth; ++t)
    th.emplace_back([&, t]() {
      for (int i = t; i < n; i += nth) fn(i);
    });
  for (auto &x : th) x.join();
}

template <typename T>
T *eb_upload(const std::vector<T> &h, hipStream_t st) {
  T *d = nullptr;
  VCA_HIP(hipMalloc(&d, sizeof(T) * std::max<size_t>(1, h.size())));
  VCA_HIP(hipMemcpyAsync(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st));
  return d;
}

}  // namespace

// nev lowest eigenpairs of the FP64 sector sec.  start: device complex128
// start vector or null for the hash start vector (seed 5001, the start
// vector of the ground-state Lanczos).  Eigenvectors are returned as
// separate device allocations (caller frees), normalised.
EigsOut eigs_lanczos(Sector &sec, int nev, int ncv, double tol, int maxiter, const void *start,
                     std::vector<double> &evals, std::vector<void *> &evecs) {
  Trl s(sec, nev, ncv, tol, maxiter, sec.stream);
  s.start(start);
  while (!s.done) {
    while (s.j0 < s.m) {
      for (int j = s.j0; j < s.m; ++j) s.wk.device_step(j);
      s.fetch();
      VCA_HIP(hipStreamSynchronize(s.wk.st));
      s.absorb();
    }
    s.end_cycle();
  }
  s.eigenvectors(evals, evecs);
  return s.out;
}

void eigs_lanczos_batch(std::vector<EigsJob> &jobs, const CrossTerm &cross, bool real_amp,
                        hipStream_t st) {
  const int nj = (int)jobs.size();
  if (nj == 0) return;
  std::vector<std::unique_ptr<Trl>> tr(nj);
  auto fail_job = [&](int q) {
    jobs[q].code = VCA_ERR_HIP;
    jobs[q].msg = "eigs batch: unknown error";
    try {
      throw;
    } catch (const Error &e) {
      jobs[q].code = e.code;
      jobs[q].msg = last_error_message();
    } catch (...) {
    }
    tr[q].reset();
  };
  for (int q = 0; q < nj; ++q) {
    try {
      tr[q] = std::make_unique<Trl>(*jobs[q].sec, jobs[q].nev, jobs[q].ncv, jobs[q].tol,
                                    jobs[q].maxiter, st);
      tr[q]->start(jobs[q].start);
    } catch (...) {
      fail_job(q);
    }
  }
  // per-job descriptors, pooled coefficient / norm / rotation arrays (one
  // copy each way per round), and the static work items of every kernel
  int maxm = 1, namp = 1;
  for (int q = 0; q < nj; ++q)
    if (tr[q]) maxm = std::max(maxm, tr[q]->m);
  const int64_t cstride = (int64_t)maxm * 4 * (maxm + 1);  // m x 2 passes x ldc
  const int64_t ystride = (int64_t)maxm * maxm;
  double *d_coef = nullptr, *d_nrm2 = nullptr;
  double2 *d_Y = nullptr;
  VCA_HIP(hipMalloc(&d_coef, sizeof(double) * cstride * nj));
  VCA_HIP(hipMalloc(&d_nrm2, sizeof(double) * maxm * nj));
  VCA_HIP(hipMalloc(&d_Y, sizeof(double2) * ystride * nj));
  std::vector<double> h_coef((size_t)cstride * nj), h_nrm2((size_t)maxm * nj);
  std::vector<double2> h_Y((size_t)ystride * nj);
  std::vector<EbJob> desc(nj);
  std::vector<int2> it_rows, it_vec, it_mdot, it_col, it_one, it_rot;
  for (int q = 0; q < nj; ++q) {
    EbJob &d = desc[q];
    d = EbJob{};
    if (!tr[q]) continue;
    const Work &w = tr[q]->wk;
    d.tab = row_tab(*jobs[q].sec);
    d.V = w.V;
    d.W = w.W;
    d.dim = w.dim;
    d.m = w.m;
    d.nblk = w.nblk;
    d.ldc = w.ldc();
    d.pr = w.pr;
    d.coef = d_coef + q * cstride;
    d.nrm2 = d_nrm2 + q * maxm;
    namp = std::max(namp, d.tab.namp_up + d.tab.namp_dw);
    for (int64_t b = 0; b < (d.dim + kEBS - 1) / kEBS; ++b) it_rows.push_back(make_int2(q, (int)b));
    for (int b = 0; b < d.nblk; ++b) it_vec.push_back(make_int2(q, b));
    for (int g = 0; g < (d.m + kED - 1) / kED; ++g)
      for (int b = 0; b < d.nblk; ++b) it_mdot.push_back(make_int2(q, g * d.nblk + b));
    for (int g = 0; g < (d.m + kEG - 1) / kEG; ++g)
      for (int b = 0; b < d.nblk; ++b) it_rot.push_back(make_int2(q, g * d.nblk + b));
    for (int c = 0; c < 2 * d.m; ++c) it_col.push_back(make_int2(q, c));
    it_one.push_back(make_int2(q, 0));
  }
  EbJob *d_desc = eb_upload(desc, st);
  int2 *d_rows = eb_upload(it_rows, st), *d_vec = eb_upload(it_vec, st);
  int2 *d_mdot = eb_upload(it_mdot, st), *d_col = eb_upload(it_col, st);
  int2 *d_one = eb_upload(it_one, st), *d_rot = eb_upload(it_rot, st);
  int *d_jtab = nullptr, *d_rtab = nullptr;
  VCA_HIP(hipMalloc(&d_rtab, sizeof(int) * 2 * nj));
  size_t jtab_cap = 0;
  std::vector<int> jtab, rtab(2 * nj);
  const size_t smem_amp = sizeof(double2) * namp, smem_h = sizeof(double2) * maxm;
  EbArgs a0{};
  a0.jobs = d_desc;
  a0.cross = cross;
  auto live = [&](int q) { return tr[q] && !tr[q]->done; };
  for (;;) {
    int tsteps = 0;
    for (int q = 0; q < nj; ++q)
      if (live(q)) tsteps = std::max(tsteps, tr[q]->m - tr[q]->j0);
    if (tsteps == 0) break;
    jtab.assign((size_t)tsteps * nj, -1);
    for (int q = 0; q < nj; ++q)
      if (live(q))
        for (int t = 0; t < tr[q]->m - tr[q]->j0; ++t) jtab[(size_t)t * nj + q] = tr[q]->j0 + t;
    if (jtab.size() > jtab_cap) {
      if (d_jtab) VCA_HIP(hipFree(d_jtab));
      jtab_cap = jtab.size();
      VCA_HIP(hipMalloc(&d_jtab, sizeof(int) * jtab_cap));
    }
    VCA_HIP(hipMemcpyAsync(d_jtab, jtab.data(), sizeof(int) * jtab.size(), hipMemcpyHostToDevice,
                           st));
    for (int t = 0; t < tsteps; ++t) {
      EbArgs a = a0;
      a.jtab = d_jtab + (size_t)t * nj;
      a.items = d_rows;
      if (real_amp)
        hipLaunchKernelGGL(k_eb_hxv<true>, dim3(it_rows.size()), dim3(kEBS), smem_amp, st, a);
      else
        hipLaunchKernelGGL(k_eb_hxv<false>, dim3(it_rows.size()), dim3(kEBS), smem_amp, st, a);
      for (int pass = 0; pass < 2; ++pass) {
        a.pass = pass;
        a.items = d_mdot;
        hipLaunchKernelGGL(k_eb_mdot, dim3(it_mdot.size()), dim3(kEBS), 0, st, a);
        a.items = d_col;
        hipLaunchKernelGGL(k_eb_colsum, dim3(it_col.size()), dim3(kEBS), 0, st, a);
        a.items = d_vec;
        hipLaunchKernelGGL(k_eb_sub, dim3(it_vec.size()), dim3(kEBS), smem_h, st, a);
        a.items = d_one;
        hipLaunchKernelGGL(k_eb_nrm, dim3(it_one.size()), dim3(kEBS), 0, st, a);
      }
      a.items = d_vec;
      hipLaunchKernelGGL(k_eb_scale, dim3(it_vec.size()), dim3(kEBS), 0, st, a);
      VCA_HIP(hipGetLastError());
    }
    VCA_HIP(hipMemcpyAsync(h_coef.data(), d_coef, sizeof(double) * h_coef.size(),
                           hipMemcpyDeviceToHost, st));
    VCA_HIP(hipMemcpyAsync(h_nrm2.data(), d_nrm2, sizeof(double) * h_nrm2.size(),
                           hipMemcpyDeviceToHost, st));
    VCA_HIP(hipStreamSynchronize(st));
    // host side of every job (on host threads: the projected eigenproblems
    // are independent); restarts and finishes share two launches
    std::fill(rtab.begin(), rtab.end(), 0);
    std::fill(rtab.begin() + nj, rtab.end(), -1);
    std::vector<int> state(nj, 0);  // 1 restart, 2 finish
    eb_parallel(nj, [&](int q) {
      if (!live(q)) return;
      Trl &x = *tr[q];
      try {
        const int ldc = x.wk.ldc();
        std::copy(h_coef.begin() + q * cstride + (size_t)x.j0 * 2 * ldc,
                  h_coef.begin() + q * cstride + (size_t)x.m * 2 * ldc,
                  x.hc.begin() + (size_t)x.j0 * 2 * ldc);
        std::copy(h_nrm2.begin() + q * maxm + x.j0, h_nrm2.begin() + q * maxm + x.m,
                  x.hn.begin() + x.j0);
        if (!x.absorb()) return;  // a breakdown: the rest of the cycle again
        const bool restart = x.end_cycle_host();
        std::copy(x.Yh.begin(), x.Yh.end(), reinterpret_cast<zc *>(h_Y.data() + q * ystride));
        rtab[q] = restart ? x.k : x.nev;
        if (restart) rtab[nj + q] = x.k;
        state[q] = restart ? 1 : 2;
      } catch (...) {
        fail_job(q);
        rtab[q] = 0;
        rtab[nj + q] = -1;
      }
    });
    std::vector<int> restarting, finishing;
    for (int q = 0; q < nj; ++q) {
      if (state[q] == 1) restarting.push_back(q);
      if (state[q] == 2) finishing.push_back(q);
    }
    if (!restarting.empty() || !finishing.empty()) {
      VCA_HIP(hipMemcpyAsync(d_Y, h_Y.data(), sizeof(double2) * h_Y.size(),
                             hipMemcpyHostToDevice, st));
      VCA_HIP(hipMemcpyAsync(d_rtab, rtab.data(), sizeof(int) * rtab.size(),
                             hipMemcpyHostToDevice, st));
      EbArgs a = a0;
      a.items = d_rot;
      hipLaunchKernelGGL(k_eb_rotate, dim3(it_rot.size()), dim3(kEBS),
                         sizeof(double2) * kEG * maxm, st, a, d_rtab, d_Y, ystride);
      a.items = d_vec;
      hipLaunchKernelGGL(k_eb_restart, dim3(it_vec.size()), dim3(kEBS), 0, st, a, d_rtab + nj);
      VCA_HIP(hipGetLastError());
    }
    for (int q : restarting) {
      if (!tr[q]) continue;
      try {
        tr[q]->restart_fresh();
      } catch (...) {
        fail_job(q);
      }
    }
    for (int q : finishing) {
      if (!tr[q]) continue;
      try {
        tr[q]->copy_eigenvectors(jobs[q].evals, jobs[q].evecs);
        jobs[q].out = tr[q]->out;
        jobs[q].code = VCA_OK;
      } catch (...) {
        fail_job(q);
      }
    }
    // the host copies of this round's tables stay alive until the stream
    // has read them
    VCA_HIP(hipStreamSynchronize(st));
  }
  for (void *p : {(void *)d_desc, (void *)d_rows, (void *)d_vec, (void *)d_mdot, (void *)d_col,
                  (void *)d_one, (void *)d_rot, (void *)d_jtab, (void *)d_rtab, (void *)d_coef,
                  (void *)d_nrm2, (void *)d_Y})
    if (p) VCA_HIP(hipFree(p));
  for (int q = 0; q < nj; ++q) tr[q].reset();
}

}  // namespace vca
