// vca_internal.hpp -- shared types of the MI355X VCA solver core.
//
// Layout of a sector (nup, ndw) in HBM:
//   state i = iup + idw * DimUp (0-based), the reference's layout
//   (VCA_SETUP.f90:505-518); a vector is complex128 (FP64) or complex64
//   (MIXED), interleaved (re, im).
//   Per spin sigma: map[dim] (ascending strings of popcount n, the order of
//   build_sector, VCA_SETUP.f90:679-684), energy[dim] (one-spin diagonal
//   energy), cnt[dim] (hops per row) and ent[maxnnz][dim] (k-major ELL of
//   the one-spin hop matrix H_sigma(row, col): col in bits 0..20, signed
//   amplitude index 2*term+neg in bits 21..31).
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vca_gpu.h"

namespace vca {

using zc = std::complex<double>;

constexpr int kMaxNs = 24;        // strings fit 32 bits; dims fit 21 bits
constexpr int kColBits = 21;
constexpr uint32_t kColMask = (1u << kColBits) - 1u;
constexpr int kMaxTerms = 1024;   // 2*terms must fit the 11 high bits
constexpr int kBinomN = kMaxNs + 1;
// Hop entries of the LDS-staged H x v (DevSpin::ent4): bits 0-17 the LDS byte
// address of the gathered element, bits 18-31 that of its signed amplitude.
constexpr int kLdsAddrBits = 18;
constexpr uint32_t kLdsAddrMask = (1u << kLdsAddrBits) - 1u;
constexpr uint32_t kLdsAmpLimit = 1u << (32 - kLdsAddrBits);  // amplitude table bytes
// dynamic LDS of the large-column kernels (of the 160 KB per CU)
constexpr size_t kLdsBig = 150 * 1024;

struct Error {
  int code;
};
[[noreturn]] void fail(int code, const char *fmt, ...);
void set_error(const char *fmt, ...);

#define VCA_HIP(call)                                                       \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      ::vca::fail(VCA_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,   \
                  hipGetErrorString(e_));                                   \
  } while (0)

// ---------------------------------------------------------------- model --
// One off-diagonal one-body term c^+_a c_b (0-based orbital positions) of
// one spin, with its amplitude.
struct Term {
  int a, b;
  double re, im;
};

struct SpinModel {
  std::vector<Term> terms;      // nonzero terms, reference insertion order
  double eps[kMaxNs];           // linear energy per orbital (incl. -xmu, hf)
  double pair[kMaxNs][kMaxNs];  // same-spin density-density, a < b
  bool real_amp;
};

// Hamiltonian of the cluster in the form the kernels consume:
//   H = sum_sigma sum_{a!=b} T_sigma(a,b) c^+_a c_b
//     + sum_sigma E_sigma(string_sigma) + cross(up, dw) + e_const
// cross(up,dw) = sum_{io,jo} wx[io][jo] popc((up>>io) & (dw>>jo) & site_mask)
struct ModelData {
  int ns = 0, nlat = 0, norb = 0, nspin = 0, nlat_bath = 0, norb_bath = 0;
  SpinModel spin[2];
  double wx[5][5] = {};
  uint32_t site_mask = 0;
  double e_const = 0;
};

// Binomial table C(n, k), n, k < kBinomN, row-major [n][k].
struct Binom {
  uint32_t c[kBinomN * kBinomN];
  Binom();
  uint32_t operator()(int n, int k) const {
    return (k < 0 || k > n || n < 0) ? 0u : c[n * kBinomN + k];
  }
};

// ------------------------------------------------- Fock-space arithmetic --
// Rank of a string among the strings of equal popcount in ascending order:
// rank(s) = sum_k C(p_k, k) over set-bit positions p_1 < p_2 < ... (k from
// 1).  Equals binary_search(map, s) - 1 (VCA_SETUP.f90:946-963) on the
// ascending map, checked by brute force in the tests.
__host__ __device__ inline uint32_t rank_string(uint32_t s, const uint32_t *binom) {
  uint32_t r = 0;
  int k = 1;
  while (s) {
    int p = __builtin_ctz(s);
    r += binom[p * kBinomN + k];
    ++k;
    s &= s - 1u;
  }
  return r;
}

__host__ __device__ inline uint32_t unrank_string(int ns, int n, uint32_t r,
                                                  const uint32_t *binom) {
  uint32_t s = 0;
  int p = ns - 1;
  for (int k = n; k >= 1; --k) {
    while (binom[p * kBinomN + k] > r) --p;
    s |= 1u << p;
    r -= binom[p * kBinomN + k];
    --p;
  }
  return s;
}

// Parity of the number of set bits of s strictly below position pos.
__host__ __device__ inline uint32_t parity_below(uint32_t s, int pos) {
  return (uint32_t)__builtin_popcount(s & ((1u << pos) - 1u)) & 1u;
}

// Hops into row string s: for every term (a <- b) with a occupied and b
// empty in s, the column string j = s - a + b and the sign of
// c^+_a c_b acting on j, computed as the reference does: c(b) on j then
// c^+(a) on the result (VCA_SETUP.f90:724-750), each sign the parity of the
// occupied orbitals below the position.  Writes packed entries to
// out[q*stride] and returns their count.
__host__ __device__ inline int row_hops(uint32_t s, int nterms, const int2 *ab,
                                        const uint32_t *binom, uint32_t *out,
                                        int64_t stride) {
  int q = 0;
  for (int t = 0; t < nterms; ++t) {
    const int a = ab[t].x, b = ab[t].y;
    if (((s >> a) & 1u) == 0u || ((s >> b) & 1u) != 0u) continue;
    const uint32_t j = (s ^ (1u << a)) | (1u << b);
    const uint32_t k1 = j & ~(1u << b);
    const uint32_t neg = parity_below(j, b) ^ parity_below(k1, a);
    const uint32_t col = rank_string(j, binom);
    out[(int64_t)q * stride] = col | ((uint32_t)(2 * t) + neg) << kColBits;
    ++q;
  }
  return q;
}

__host__ __device__ inline double spin_energy(uint32_t s, int ns, const double *eps,
                                              const double *pair /*[ns][ns]*/) {
  double e = 0;
  for (int a = 0; a < ns; ++a)
    if ((s >> a) & 1u) {
      e += eps[a];
      for (int b = a + 1; b < ns; ++b)
        if ((s >> b) & 1u) e += pair[a * kMaxNs + b];
    }
  return e;
}

// Host copy of one spin table (dense path and CPU tests).
struct HostSpinTable {
  int npart = 0, dim = 0, maxnnz = 0;
  std::vector<uint32_t> map;
  std::vector<double> energy;
  std::vector<int32_t> cnt;
  std::vector<uint32_t> ent;
};

// ------------------------------------------------------------ device ----
// One excitation chain's operator O = sum_o coef_o (c or c^+)_{pos_o}, all
// on the spin of the launch (VCA_GF_NORMAL.f90:184-198, 624-660).
struct ExcOp {
  int nops;
  int pos[2], dagger[2];
  double cre[2], cim[2];
};

struct LanczosDev {
  int32_t k;     // current step
  int32_t done;  // exact breakdown seen
  uint32_t counter[4];
  double *alpha, *beta, *nrm, *y;  // device arrays
};

// Per-row diagonal energy, string and hop count in one 16-byte record.
struct alignas(16) RowMeta {
  double e;
  uint32_t s;
  int32_t cnt;
};

struct DevSpin {
  int npart = 0, dim = 0, maxnnz = 0, nterms = 0;
  bool real_amp = true;
  uint32_t *map = nullptr;
  double *energy = nullptr;
  int32_t *cnt = nullptr;
  uint32_t *ent = nullptr;
  RowMeta *meta = nullptr;   // [dim]
  // [ceil(maxnnz/4)][dim][4]: 4 hops of a row per 16 bytes, as LDS byte
  // addresses for the staged kernel of the sector's precision (amplitude
  // table with a zero entry at LDS 0, then the column); padding entries
  // address row 0 and the zero amplitude.  lds_form: the addresses fit.
  uint32_t *ent4 = nullptr;
  bool lds_form = false;
  // The same hops addressed for one real plane staged in LDS (elements and
  // amplitudes of sizeof(real) bytes), read by the plane kernel of the
  // transposed H x v on spins with real amplitudes; plane_form: usable.
  uint32_t *ent4p = nullptr;
  // ent4p's hops addressed for two columns staged interleaved (element f of
  // both at the column offset + 2 f sizeof(real)), read by the two-column
  // plane kernel; null where a column pair does not fit the LDS.
  uint32_t *ent4q = nullptr;
  // Signed-column entries (every hopping term the same real amplitude t, so
  // a hop's amplitude is +t or -t): each hop is the LDS address of its element in a staged copy
  // of the column scaled by +t or by -t, padding hops the zero at address 0.
  // sgn_form: built (the two copies fit the address bits).
  uint32_t *ent4s = nullptr;
  bool sgn_form = false;
  bool plane_form = false;
  double2 *amp64 = nullptr;  // [2*nterms]: +amp, -amp
  float2 *amp32 = nullptr;
};

struct CrossTerm {
  double wx[5][5];
  uint32_t site_mask;
  int norb;
};

struct Solver;

struct Sector {
  Solver *solver = nullptr;
  int nup = 0, ndw = 0, precision = VCA_FP64;
  int64_t dim = 0;
  DevSpin up, dw;
  hipStream_t stream = nullptr;
  bool own_stream = true;  // false: a stream of the solver's, not destroyed here
  // Lanczos workspace (lazily allocated) for nb_cap chains of lz_cap steps
  int lz_cap = 0, nb_cap = 0;
  int hxv_variant = -1;  // launch shape of H x v picked by the autotuner
  int real_variant = -1;  // the same for the real-vector direct form
  void *buf[3] = {nullptr, nullptr, nullptr};  // u_even, u_odd, hv [nb][dim]
  void *xvec = nullptr;                        // eigenvector accumulator
  double *partials = nullptr;
  int partials_cap = 0;
  // transposed form of H x v: one scratch vector per chain [tbuf_cap][dim]
  void *tbuf = nullptr;
  int tbuf_cap = 0;
  LanczosDev *lz = nullptr;  // device structs [nb_cap]
  LanczosDev lzh{};          // host copy of chain 0's device array pointers
  std::vector<LanczosDev> lzh_all;  // host copies for every chain
  ExcOp *d_ops = nullptr;    // [nb_cap]
  double *lz_arrays = nullptr;
  // Real-vector mode of a ground-state run (real H, real start vector):
  // buf[*] and xvec hold dim reals instead of dim complex numbers, H x v is
  // the transposed form on one real plane, and the element-wise kernels
  // (update, norms, scaling, conversion) see the reals as ceil(dim/2)
  // complex pairs, the odd pad element kept 0.  One chain only.
  bool real_vec = false;
  // captured chunks of Lanczos steps, keyed by (steps, chains, accumulate,
  // real-vector mode)
  struct StepGraph {
    int nsteps, nb;
    bool accx, real;
    hipGraphExec_t exec;
  };
  std::vector<StepGraph> graphs;
  size_t vec_bytes() const {
    return (size_t)dim * (precision == VCA_FP64 ? sizeof(double2) : sizeof(float2));
  }
  ~Sector();
};

// Real-vector mode of a sector for the lifetime of the scope (Sector::
// real_vec), switched off again on every exit, errors included.
struct RealVecScope {
  Sector &sec;
  RealVecScope(Sector &s, bool on) : sec(s) { sec.real_vec = on; }
  ~RealVecScope() { sec.real_vec = false; }
  RealVecScope(const RealVecScope &) = delete;
  RealVecScope &operator=(const RealVecScope &) = delete;
};

// Kernel launch wrappers (kernels.hip)
void build_spin_tables(Sector &sec, DevSpin &sp, const SpinModel &m, int ns,
                       int npart, const uint32_t *d_binom, double e_add);
CrossTerm sector_cross(const Sector &sec);
const uint32_t *solver_binom(const Solver *s);
// scratch (complex vectors, transposed form only): nb vectors of the
// sector's element size for v^T, or null for the sector's own (sec.tbuf,
// allocated on first use)
void launch_hxv(Sector &sec, const void *v0, const void *v1, void *hv, bool lanczos,
                int nb = 1, void *scratch = nullptr);
void launch_lanczos_update(Sector &sec, bool accumulate_x, int nb = 1);
void ensure_step_partials(Sector &sec, int nb);
int hxv_num_variants(const Sector &sec);
bool hxv_variant_ok(const Sector &sec, int v);  // launch shape v can run on sec
// vectors of device scratch per chain that the sector's H x v shape needs
int hxv_scratch_vectors(const Sector &sec);
// Real-vector mode (Sector::real_vec) can run on the sector: real
// amplitudes and the one-plane transposed H x v fits.
bool real_mode_ok(const Sector &sec);
// Launch shapes of the real-vector direct form (used when the transposed
// form is not): their number, whether shape i runs this sector, and the
// autotuner that picks one per sector shape (never inside a graph capture).
int real_num_variants();
bool real_variant_ok(const Sector &sec, int var);
void tune_real_direct(Sector &sec);
// sector size from which the transposed H x v and the real-vector ground
// state are used (2^22 states; VCA_TFORM_MIN_DIM overrides)
int64_t tform_min_dim();
// complex elements the element-wise kernels cover (ceil(dim/2) in
// real-vector mode)
int64_t vec_elems(const Sector &sec);
// dst[i] = src[i] + 0i for the dim reals of src (same precision)
void launch_expand_real(Sector &sec, const void *src, void *dst);
// out (n reals in precision prec) = Re(psi e^{-i phi}) for the complex128
// device vector psi, phi = arg(sum psi_i^2) / 2 its global phase; returns
// |Im(psi e^{-i phi})| / |psi| (rounding level when psi is a real vector
// times a phase).
double real_vector_of(const void *psi, int64_t n, void *out, int prec, hipStream_t st);
void clear_graphs(Sector &sec);
void launch_start_vector(Sector &sec, uint64_t seed);
void launch_norm0(Sector &sec);  // nrm[0] = ||buf[0]||, k = 0, done = 0
void launch_scale_x(Sector &sec);  // xvec /= ||xvec||
void launch_convert(const void *src, int src_prec, void *dst, int dst_prec,
                    int64_t n, hipStream_t st);
void launch_excite(Sector &jsec, int dimup_i, int dimdw_i, const void *psi, int spin,
                   const ExcOp *d_ops, int nb);
CrossTerm make_cross(const ModelData &md);
// Partial sums per chain of the excitation's norm (grid of k_excite).
int excite_grid(int64_t dim);
// k_excite into an explicit buffer: nb chains' start vectors O|psi> at out
// (vstride elements apart; reals when real), their norms into lz[c].nrm[0];
// partials holds excite_grid(jsec.dim) * nb doubles.
void launch_excite_into(const Sector &jsec, int dimup_i, int dimdw_i, const void *psi, int spin,
                        const ExcOp *d_ops, int nb, void *out, int64_t vstride, bool real,
                        LanczosDev *lz, double *partials, hipStream_t st);

// Chains of uneven target sectors run together (chains.hip): every chain
// of the set advances one Lanczos step per two launches, captured in
// 16-step graphs.
struct ChainJob {
  Sector *js;   // target sector (its spin tables; its precision is the set's)
  bool real;    // real vectors (real Hamiltonian, real state, real operator)
  int nsteps;   // Lanczos steps of the chain
  ExcOp op;     // excitation operator on the kept state
  int spin;     // spin the operator acts on (1 up, 2 down)
};
struct ChainOut {
  std::vector<double> alpha, beta;  // nsteps each; beta[0] = 0
  double nrm0 = 0;                  // |O psi|
};
struct ChainSetStats {
  double seconds = 0;     // GPU time of the Lanczos steps
  int64_t launches = 0;   // kernel launches of those steps
  int64_t hxv_count = 0;  // chain steps
  double hxv_states = 0;  // sum over chain steps of the sector dimension
  int64_t device_bytes = 0;  // vector memory of the set
  int steps = 0;          // steps of the longest chain
  int column_items = 0;   // segments whose H x v items are LDS-staged columns
};
// Device bytes of one chain in a set (its three vectors).
int64_t chain_set_bytes(const Sector &js, bool real, int prec);
// Runs every job to its step count; psi is the kept state (complex, of the
// set's precision) on the source sector (dimup_i x dimdw_i), psi_r its real
// form (null when no job is real).
ChainSetStats run_chain_set(const std::vector<ChainJob> &jobs, int prec, int dimup_i,
                            int dimdw_i, const void *psi, const void *psi_r,
                            const CrossTerm &cross, bool real_amp, hipStream_t st,
                            std::vector<ChainOut> &out);

// Lowest eigenpairs by thick-restart Lanczos (eigs.hip).
struct EigsOut {
  int nconv;     // converged among the nev requested
  int restarts;  // restarts used
  int64_t hxv;   // H x v applications
};
EigsOut eigs_lanczos(Sector &sec, int nev, int ncv, double tol, int maxiter, const void *start,
                     std::vector<double> &evals, std::vector<void *> &evecs);
// Many eigensolves at once (the small sectors of a sweep): the same
// thick-restart solve per job, with the Lanczos steps of all jobs in shared
// launches on stream st.  A job's failure is reported in its code and msg.
struct EigsJob {
  Sector *sec = nullptr;  // FP64 sector
  int nev = 0, ncv = 0, maxiter = 0;
  double tol = 0;
  const void *start = nullptr;  // device start vector or null (hash vector)
  std::vector<double> evals;
  std::vector<void *> evecs;    // device allocations, caller frees
  EigsOut out{};
  int code = VCA_OK;
  std::string msg;
};
void eigs_lanczos_batch(std::vector<EigsJob> &jobs, const CrossTerm &cross, bool real_amp,
                        hipStream_t st);
// The calling thread's last error message (set by fail / set_error).
std::string last_error_message();

// Stream-ordered device memory (hipMallocAsync / hipFreeAsync) for the
// eigenvectors and the eigensolver workspaces: a plain hipFree waits for the
// whole device, which stalls sector solves running concurrently on other
// streams.  Free only what the host has already synchronised with.
void *dev_alloc_async(size_t bytes, hipStream_t st);
void dev_free_async(void *p, hipStream_t st);

// Host linear algebra (linalg.cpp)
// Symmetric tridiagonal eigenproblem, diag d[n], offdiag e[n] with e[i]
// coupling i-1 and i (e[0] unused).  Eigenvalues ascending in w; if z is
// not null, z is n x n column-major eigenvectors.
void tridiag_eig(int n, const double *d, const double *e, double *w, double *z);
// Eigenvalues (ascending) and the first component of each eigenvector,
// z0[j] = Z(0, j): what the continued-fraction weights need, in O(n^2).
void tridiag_eig_first(int n, const double *d, const double *e, double *w, double *z0);
double lanczos_of_sum(int n, const double *a1, const double *b1, double n1, const double *a2,
                      const double *b2, double n2, double *al, double *be);
double lanczos_of_quadratures(int n, const double *x1, const double *z1, double n1,
                              const double *x2, const double *z2, double n2, double *al,
                              double *be);
// Lowest eigenvalue of the leading m x m block.
double tridiag_lowest(int m, const double *d, const double *e);
// Complex Hermitian eigenproblem (column-major a, n x n, overwritten by the
// eigenvectors), eigenvalues ascending in w.
void herm_eig(int n, zc *a, double *w);

// Pole file of the cluster Green's function (gfprime.cpp).  One entry per
// element: [state][channel] -> poles and weights, complex as in the file.
struct GfChanIO {
  std::vector<zc> w, p;
};
using GfEntry = std::vector<std::vector<GfChanIO>>;
void gfprime_write(const std::string &path, bool formatted, const std::vector<GfEntry> &el);
std::vector<GfEntry> gfprime_read(const std::string &path, bool formatted, size_t nelem);

}  // namespace vca
