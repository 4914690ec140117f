// solver.cpp -- host orchestration and C ABI of the VCA solver core.
//
// The flow of vca_solve follows vca_solve_serial (VCA_MAIN.f90:124-215):
//   1. input ingestion: Hloc, H_k, bath (VCA_AUX_FUNX.f90:124-164,
//      VCA_BATH_SETUP.f90:112-138) -> one-body term tables per spin;
//   2. sector scan and T=0 ground-state selection (VCA_DIAG.f90:83-235):
//      dense Hermitian eigensolve for small sectors, GPU Lanczos otherwise;
//   3. excitation chains per kept state (VCA_GF_NORMAL.f90:37-104,121-958):
//      start vector built on the GPU, GPU Lanczos, host tridiagonal
//      eigensolve -> Lehmann poles and weights (:970-1022);
//   4. Omega integrand on the GPU for batches of frequencies and the
//      frequency quadrature on the host (VCA_OMEGA.f90:123-188);
//   5. Omega = E0 - integral (VCA_MAIN.f90:183-197).
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <condition_variable>
#include <exception>
#include <mutex>
#include <tuple>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "omega.hpp"
#include "vca_internal.hpp"

namespace vca {

// ------------------------------------------------------------- errors ---
static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

std::string last_error_message() { return g_err; }

void fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  throw Error{code};
}

Binom::Binom() {
  for (int n = 0; n < kBinomN; ++n)
    for (int k = 0; k < kBinomN; ++k) {
      uint32_t v;
      if (k == 0) v = 1;
      else if (n == 0) v = 0;
      else v = c[(n - 1) * kBinomN + (k - 1)] + c[(n - 1) * kBinomN + k];
      c[n * kBinomN + k] = v;
    }
}

// ------------------------------------------------------------- solver ---
struct GfChan {
  std::vector<zc> w;
  std::vector<double> p;
};
using ElemKey = std::array<int, 5>;  // ilat, jlat, ispin, iorb, jorb (1-based)

struct KeptState {
  double e;
  int nup, ndw;
  void *d_vec;  // complex128 on device
};

struct Solver {
  vca_params p{};
  std::vector<int32_t> sectors;
  int ns = 0, nlso = 0;
  std::vector<zc> hloc, bath_h, bath_v, hk;
  int64_t nk = 0;
  bool has_hloc = false, has_bath = false;
  ModelData md;
  bool md_valid = false;
  uint32_t *d_binom = nullptr;
  hipStream_t stream = nullptr;
  hipStream_t batch_stream = nullptr;  // a sweep's batched eigensolves
  hipStream_t dense_stream = nullptr;  // a sweep's dense-sector copies
  // one stream per sweep worker thread, kept for the solver's lifetime: the
  // sectors a worker solves run on it (no stream is created and destroyed
  // per sector, and no stream-ordered allocation or free is left on a
  // destroyed stream)
  std::vector<hipStream_t> worker_streams;
  vca_results res{};
  bool solved = false;
  std::map<ElemKey, std::vector<std::vector<GfChan>>> gf;
  double emin = 0, zeta = 1, max_exc = 0;
  std::unique_ptr<OmegaDev> om;
  bool om_ready = false;
  bool gf_loaded = false;  // pole store read from a file (vca_read_gfprime)
  std::mutex res_mu;       // res updates from the concurrent sector solves
  double t_start = 0;      // steady-clock seconds at the start of vca_solve
  std::vector<KeptState> states;
  ~Solver();
};

void *dev_alloc_async(size_t bytes, hipStream_t st) {
  void *p = nullptr;
  VCA_HIP(hipMallocAsync(&p, std::max<size_t>(bytes, 1), st));
  return p;
}

void dev_free_async(void *p, hipStream_t st) {
  if (p) VCA_HIP(hipFreeAsync(p, st));
}

static void free_states(Solver &s) {
  for (auto &st : s.states)
    if (st.d_vec) (void)hipFreeAsync(st.d_vec, s.stream);
  s.states.clear();
}

Solver::~Solver() {
  free_states(*this);
  om.reset();
  if (d_binom) (void)hipFree(d_binom);
  for (hipStream_t st : {batch_stream, dense_stream})
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  for (hipStream_t st : worker_streams) {
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
  if (stream) (void)hipStreamDestroy(stream);
}

const uint32_t *solver_binom(const Solver *s) { return s->d_binom; }

void clear_graphs(Sector &sec) {
  for (auto &g : sec.graphs) (void)hipGraphExecDestroy(g.exec);
  sec.graphs.clear();
}

Sector::~Sector() {
  if (stream) (void)hipStreamSynchronize(stream);
  clear_graphs(*this);
  for (DevSpin *sp : {&up, &dw})
    for (void *q : {(void *)sp->map, (void *)sp->energy, (void *)sp->cnt, (void *)sp->ent,
                    (void *)sp->amp64, (void *)sp->amp32, (void *)sp->meta, (void *)sp->ent4,
                    (void *)sp->ent4p, (void *)sp->ent4q, (void *)sp->ent4s})
      if (q) (void)hipFree(q);
  for (void *q : {buf[0], buf[1], buf[2], xvec, tbuf, (void *)partials, (void *)lz,
                  (void *)lz_arrays, (void *)d_ops})
    if (q) (void)hipFree(q);
  if (stream && own_stream) (void)hipStreamDestroy(stream);
}

CrossTerm sector_cross(const Sector &sec) {
  const ModelData &md = sec.solver->md;
  CrossTerm c{};
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) c.wx[i][j] = md.wx[i][j];
  c.site_mask = md.site_mask;
  c.norb = md.norb;
  return c;
}

// rank-6 column-major index, 0-based indices
static inline size_t idx6(int i1, int i2, int i3, int i4, int i5, int i6, int d1, int d2,
                          int d3, int d4, int d5) {
  return (size_t)i1 +
         (size_t)d1 * (i2 + (size_t)d2 * (i3 + (size_t)d3 * (i4 + (size_t)d4 * (i5 + (size_t)d5 * i6))));
}

// Derive the term tables from Hloc and the bath, in the insertion order of
// the stored path (H_cluster_up.f90:108-186, H_cluster_dw.f90:8-84) and the
// diagonal decomposition of H_cluster_local.f90:11-90.
static void derive_model(Solver &s) {
  const vca_params &p = s.p;
  ModelData &md = s.md;
  md = ModelData{};
  md.ns = s.ns;
  md.nlat = p.nlat;
  md.norb = p.norb;
  md.nspin = p.nspin;
  md.nlat_bath = p.nlat_bath;
  md.norb_bath = p.norb_bath;
  const int nl = p.nlat, no = p.norb, nsp = p.nspin, nlb = p.nlat_bath, nob = p.norb_bath;
  const bool bath = nlb > 0 && nob > 0;
  auto H = [&](int il, int jl, int si, int sj, int io, int jo) {
    return s.hloc[idx6(il, jl, si, sj, io, jo, nl, nl, nsp, nsp, no)];
  };
  auto BH = [&](int il, int jl, int si, int sj, int io, int jo) {
    return s.bath_h[idx6(il, jl, si, sj, io, jo, nlb, nlb, nsp, nsp, nob)];
  };
  auto BV = [&](int il, int jl, int si, int sj, int io, int jo) {
    return s.bath_v[idx6(il, jl, si, sj, io, jo, nl, nlb, nsp, nsp, no)];
  };
  auto imp = [&](int il, int io) { return io + il * no; };          // 0-based
  auto bpos = [&](int ib, int ob) { return nl * no + ib * nob + ob; };  // 0-based
  const double hf_orb = p.hfmode ? 1.0 : 0.0;
  for (int sp = 0; sp < 2; ++sp) {
    const int si = sp == 0 ? 0 : nsp - 1;  // block (1,1) or (Nspin,Nspin)
    SpinModel &m = md.spin[sp];
    m.terms.clear();
    m.real_amp = true;
    for (int a = 0; a < kMaxNs; ++a) {
      m.eps[a] = 0;
      for (int b = 0; b < kMaxNs; ++b) m.pair[a][b] = 0;
    }
    auto add = [&](int a, int b, zc v) {
      if (v == zc(0.0)) return;
      if (a == b) return;
      if (v.imag() != 0.0) m.real_amp = false;
      m.terms.push_back(Term{a, b, v.real(), v.imag()});
    };
    for (int il = 0; il < nl; ++il)
      for (int jl = 0; jl < nl; ++jl)
        for (int io = 0; io < no; ++io)
          for (int jo = 0; jo < no; ++jo) add(imp(il, io), imp(jl, jo), H(il, jl, si, si, io, jo));
    if (bath) {
      for (int il = 0; il < nlb; ++il)
        for (int jl = 0; jl < nlb; ++jl)
          for (int io = 0; io < nob; ++io)
            for (int jo = 0; jo < nob; ++jo)
              add(bpos(il, io), bpos(jl, jo), BH(il, jl, si, si, io, jo));
      for (int il = 0; il < nl; ++il)
        for (int io = 0; io < no; ++io)
          for (int jl = 0; jl < nlb; ++jl)
            for (int jo = 0; jo < nob; ++jo) {
              const zc v = BV(il, jl, si, si, io, jo);
              add(imp(il, io), bpos(jl, jo), v);
              add(bpos(jl, jo), imp(il, io), std::conj(v));
            }
    }
    if ((int)m.terms.size() > kMaxTerms)
      fail(VCA_ERR_ARG, "%zu one-body terms exceed the table limit %d", m.terms.size(), kMaxTerms);
    // linear energies
    for (int il = 0; il < nl; ++il)
      for (int io = 0; io < no; ++io) {
        const zc d = H(il, il, si, si, io, io);
        if (d.imag() != 0.0) fail(VCA_ERR_ARG, "Hloc has a complex diagonal element");
        m.eps[imp(il, io)] = d.real() - p.xmu -
                             hf_orb * (0.5 * p.uloc[io] + 0.5 * (2.0 * p.ust - p.jh) * (no - 1));
        for (int jo = io + 1; jo < no; ++jo) m.pair[imp(il, io)][imp(il, jo)] = p.ust - p.jh;
      }
    if (bath)
      for (int ib = 0; ib < nlb; ++ib)
        for (int ob = 0; ob < nob; ++ob) {
          const zc d = BH(ib, ib, si, si, ob, ob);
          if (d.imag() != 0.0) fail(VCA_ERR_ARG, "bath_h has a complex diagonal element");
          m.eps[bpos(ib, ob)] = d.real() - p.xmu;
        }
  }
  for (int io = 0; io < no; ++io)
    for (int jo = 0; jo < no; ++jo) md.wx[io][jo] = io == jo ? p.uloc[io] : p.ust;
  md.site_mask = 0;
  for (int il = 0; il < nl; ++il) md.site_mask |= 1u << (il * no);
  md.e_const = 0;
  if (p.hfmode) {
    for (int il = 0; il < nl; ++il) {
      for (int io = 0; io < no; ++io) md.e_const += 0.25 * p.uloc[io];
      md.e_const += 0.25 * (2.0 * p.ust - p.jh) * (no * (no - 1) / 2);
    }
  }
  s.md_valid = true;
}

// ---------------------------------------------------------- autotuner ---
// Picks the launch shape of H x v (lanes per workgroup, rows per lane) for
// a sector shape by timing every variant on the sector itself (events, a
// warm-up launch then the best of three), once per process and shape.
// VCA_HXV_VARIANT=<i> forces variant i; VCA_TUNE=0 keeps the default.
struct TuneKey {
  int dimup, dimdw, prec, real_amp, nt_up, nt_dw;
  bool operator<(const TuneKey &o) const {
    return std::tie(dimup, dimdw, prec, real_amp, nt_up, nt_dw) <
           std::tie(o.dimup, o.dimdw, o.prec, o.real_amp, o.nt_up, o.nt_dw);
  }
};
static std::map<TuneKey, int> g_tuned;
static std::mutex g_tune_mu;
constexpr int64_t kTuneMinDim = (int64_t)1 << 17;  // one vector is timed from here up

// Times one H x v (launch_hxv on a zero vector, as the sector's mode is
// set) for every variant that pick(var) admits and sets, and returns the
// fastest: a warm-up launch, then the best of three.
template <typename Pick>
static int time_variants(Sector &sec, int nv, Pick pick, const char *what) {
  void *v = dev_alloc_async(sec.vec_bytes(), sec.stream);
  void *hv = dev_alloc_async(sec.vec_bytes(), sec.stream);
  VCA_HIP(hipMemsetAsync(v, 0, sec.vec_bytes(), sec.stream));
  hipEvent_t e0, e1;
  VCA_HIP(hipEventCreate(&e0));
  VCA_HIP(hipEventCreate(&e1));
  int best = 0;
  float best_ms = std::numeric_limits<float>::max();
  const char *verbose = std::getenv("VCA_TUNE_VERBOSE");
  for (int var = 0; var < nv; ++var) {
    if (!pick(var)) continue;
    launch_hxv(sec, v, v, hv, false);
    float ms_min = std::numeric_limits<float>::max();
    for (int r = 0; r < 3; ++r) {
      VCA_HIP(hipEventRecord(e0, sec.stream));
      launch_hxv(sec, v, v, hv, false);
      VCA_HIP(hipEventRecord(e1, sec.stream));
      VCA_HIP(hipEventSynchronize(e1));
      float ms = 0;
      VCA_HIP(hipEventElapsedTime(&ms, e0, e1));
      ms_min = std::min(ms_min, ms);
    }
    if (ms_min < best_ms) {
      best_ms = ms_min;
      best = var;
    }
    if (verbose && verbose[0] == '1')
      std::fprintf(stderr, "%s tune dim %dx%d prec %d variant %d: %.2f us\n", what, sec.up.dim,
                   sec.dw.dim, sec.precision, var, ms_min * 1e3);
  }
  VCA_HIP(hipEventDestroy(e0));
  VCA_HIP(hipEventDestroy(e1));
  dev_free_async(v, sec.stream);
  dev_free_async(hv, sec.stream);
  return best;
}

static void hxv_tune(Sector &sec) {
  const char *force = std::getenv("VCA_HXV_VARIANT");
  if (force && *force) {
    sec.hxv_variant = std::atoi(force);
    return;
  }
  const char *t = std::getenv("VCA_TUNE");
  const int nv = hxv_num_variants(sec);
  if ((t && t[0] == '0') || nv <= 1) return;
  // below ~1e5 states every launch shape is bound by launch latency: the
  // default shape, no timing runs (they would be noise next to concurrent
  // sector solves anyway)
  if (sec.dim < kTuneMinDim) return;
  const TuneKey key{sec.up.dim, sec.dw.dim, sec.precision,
                    (int)(sec.up.real_amp && sec.dw.real_amp), sec.up.nterms, sec.dw.nterms};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      sec.hxv_variant = it->second;
      return;
    }
  }
  const int best = time_variants(sec, nv, [&](int var) {
    if (!hxv_variant_ok(sec, var)) return false;
    sec.hxv_variant = var;
    return true;
  }, "hxv");
  sec.hxv_variant = best;
  if (sec.tbuf && hxv_scratch_vectors(sec) == 0) {  // timed, not picked
    VCA_HIP(hipStreamSynchronize(sec.stream));
    VCA_HIP(hipFree(sec.tbuf));
    sec.tbuf = nullptr;
    sec.tbuf_cap = 0;
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[key] = best;
}

// The real-vector direct form (real vectors below the transposed form's
// size) has launch shapes of its own (rows per lane, adjacent row pairs),
// timed the same way with the sector in real mode.
static std::map<TuneKey, int> g_tuned_real;

void tune_real_direct(Sector &sec) {
  if (sec.real_variant >= 0 || !sec.real_vec || !real_variant_ok(sec, 0)) return;
  const char *force = std::getenv("VCA_REAL_VARIANT");  // a shape this sector can run
  if (force && *force) {
    const int var = std::atoi(force);
    if (!real_variant_ok(sec, var))
      fail(VCA_ERR_ARG, "VCA_REAL_VARIANT=%d does not fit sector (%d,%d)", var, sec.nup, sec.ndw);
    sec.real_variant = var;
    return;
  }
  const char *t = std::getenv("VCA_TUNE");
  if ((t && t[0] == '0') || sec.dim < kTuneMinDim) return;
  const TuneKey key{sec.up.dim, sec.dw.dim, sec.precision, 1, sec.up.nterms, sec.dw.nterms};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned_real.find(key);
    if (it != g_tuned_real.end()) {
      sec.real_variant = it->second;
      return;
    }
  }
  const int best = time_variants(sec, real_num_variants(), [&](int var) {
    if (!real_variant_ok(sec, var)) return false;
    sec.real_variant = var;
    return true;
  }, "real hxv");
  sec.real_variant = best;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned_real[key] = best;
}

// Launch shape for nb stacked vectors (the batched steps of the excitation
// chains), timed with the batch itself, for sectors below the size where
// one vector is timed (hxv_tune): there one vector is launch-bound while a
// batch fills the chip, so the choices differ (the 2x5 cluster's 190-chain
// batches: 212 us against 373 us with the default shape).  Above that size
// the single-vector choice carries over (on the 3x4 cluster's 276-chain
// batches both pick the same shape).  Once per (sector shape, nb);
// sec.buf[0] (zeroed) is the input, sec.buf[2] the output.
static std::map<std::pair<TuneKey, int>, int> g_tuned_batch;

static void hxv_tune_batch(Sector &sec, int nb) {
  // real-vector mode picks among the real direct form's shapes
  const bool real = sec.real_vec;
  const char *force = std::getenv(real ? "VCA_REAL_VARIANT" : "VCA_HXV_VARIANT");
  const char *t = std::getenv("VCA_TUNE");
  const int nv = real ? real_num_variants() : hxv_num_variants(sec);
  if ((force && *force) || (t && t[0] == '0') || nv <= 1 || nb <= 1) return;
  // real vectors: timed with the batch at every size (the direct form's
  // best shape for one vector is not the best for a batch of chains)
  if (real ? !real_variant_ok(sec, 0) : sec.dim >= kTuneMinDim) return;
  int &cur = real ? sec.real_variant : sec.hxv_variant;
  const TuneKey tk{sec.up.dim,   sec.dw.dim,   sec.precision,
                   real ? 2 : (int)(sec.up.real_amp && sec.dw.real_amp),
                   sec.up.nterms, sec.dw.nterms};
  const auto key = std::make_pair(tk, nb);
  const int orig = cur;
  int best = -1;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned_batch.find(key);
    if (it != g_tuned_batch.end()) best = it->second;
  }
  if (best < 0) {
    VCA_HIP(hipMemsetAsync(sec.buf[0], 0, sec.vec_bytes() * nb, sec.stream));
    hipEvent_t e0, e1;
    VCA_HIP(hipEventCreate(&e0));
    VCA_HIP(hipEventCreate(&e1));
    float best_ms = std::numeric_limits<float>::max();
    for (int var = 0; var < nv; ++var) {
      if (!(real ? real_variant_ok(sec, var) : hxv_variant_ok(sec, var))) continue;
      cur = var;
      launch_hxv(sec, sec.buf[0], sec.buf[0], sec.buf[2], false, nb);
      float ms_min = std::numeric_limits<float>::max();
      for (int r = 0; r < 2; ++r) {
        VCA_HIP(hipEventRecord(e0, sec.stream));
        launch_hxv(sec, sec.buf[0], sec.buf[0], sec.buf[2], false, nb);
        VCA_HIP(hipEventRecord(e1, sec.stream));
        VCA_HIP(hipEventSynchronize(e1));
        float ms = 0;
        VCA_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms_min = std::min(ms_min, ms);
      }
      if (ms_min < best_ms) {
        best_ms = ms_min;
        best = var;
      }
      const char *verbose = std::getenv("VCA_TUNE_VERBOSE");
      if (verbose && verbose[0] == '1')
        std::fprintf(stderr, "hxv batch tune dim %dx%d x %d prec %d real %d variant %d: %.2f us\n",
                     sec.up.dim, sec.dw.dim, nb, sec.precision, (int)real, var, ms_min * 1e3);
    }
    VCA_HIP(hipEventDestroy(e0));
    VCA_HIP(hipEventDestroy(e1));
    cur = orig;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned_batch[key] = best;
  }
  if (best != cur) {
    clear_graphs(sec);
    cur = best;
  }
}

// ------------------------------------------------------------ sectors ---
// stream: a stream of the solver's for the sector's work (not destroyed
// with the sector), or null for a stream of its own.
static std::unique_ptr<Sector> make_sector(Solver &s, int nup, int ndw, int prec,
                                           hipStream_t stream = nullptr) {
  if (!s.md_valid) fail(VCA_ERR_STATE, "Hloc not set");
  if (nup < 0 || ndw < 0 || nup > s.ns || ndw > s.ns)
    fail(VCA_ERR_ARG, "sector (%d,%d) out of range for Ns=%d", nup, ndw, s.ns);
  auto sec = std::make_unique<Sector>();
  sec->solver = &s;
  sec->nup = nup;
  sec->ndw = ndw;
  sec->precision = prec;
  if (stream) {
    sec->stream = stream;
    sec->own_stream = false;
  } else {
    VCA_HIP(hipStreamCreate(&sec->stream));
  }
  build_spin_tables(*sec, sec->up, s.md.spin[0], s.ns, nup, s.d_binom, s.md.e_const);
  build_spin_tables(*sec, sec->dw, s.md.spin[1], s.ns, ndw, s.d_binom, 0.0);
  sec->dim = (int64_t)sec->up.dim * sec->dw.dim;
  hxv_tune(*sec);
  return sec;
}

// Lanczos workspace for nb chains of up to cap steps: three vector buffers
// [nb][dim], one LanczosDev per chain with its alpha/beta/nrm/y arrays.
static void ensure_lanczos(Sector &sec, int cap, int nb = 1) {
  if (sec.nb_cap < nb || sec.lz_cap < cap) clear_graphs(sec);
  if (sec.nb_cap < nb) {
    VCA_HIP(hipStreamSynchronize(sec.stream));
    for (int i = 0; i < 3; ++i)
      if (sec.buf[i]) VCA_HIP(hipFree(sec.buf[i]));
    for (int i = 0; i < 3; ++i) VCA_HIP(hipMalloc(&sec.buf[i], sec.vec_bytes() * nb));
    sec.nb_cap = nb;
    sec.lz_cap = 0;  // per-chain arrays follow the new chain count
  }
  if (!sec.xvec) VCA_HIP(hipMalloc(&sec.xvec, sec.vec_bytes()));
  if (sec.lz_cap < cap) {
    VCA_HIP(hipStreamSynchronize(sec.stream));
    for (void *q : {(void *)sec.lz_arrays, (void *)sec.lz, (void *)sec.d_ops})
      if (q) VCA_HIP(hipFree(q));
    const size_t n = (size_t)cap + 4;
    const int nbc = sec.nb_cap;
    VCA_HIP(hipMalloc(&sec.lz_arrays, sizeof(double) * 4 * n * nbc));
    VCA_HIP(hipMemsetAsync(sec.lz_arrays, 0, sizeof(double) * 4 * n * nbc, sec.stream));
    VCA_HIP(hipMalloc(&sec.lz, sizeof(LanczosDev) * nbc));
    VCA_HIP(hipMalloc(&sec.d_ops, sizeof(ExcOp) * nbc));
    sec.lzh_all.assign(nbc, LanczosDev{});
    for (int c = 0; c < nbc; ++c) {
      double *base = sec.lz_arrays + (size_t)c * 4 * n;
      sec.lzh_all[c].alpha = base;
      sec.lzh_all[c].beta = base + n;
      sec.lzh_all[c].nrm = base + 2 * n;
      sec.lzh_all[c].y = base + 3 * n;
    }
    sec.lzh = sec.lzh_all[0];
    VCA_HIP(hipMemcpyAsync(sec.lz, sec.lzh_all.data(), sizeof(LanczosDev) * nbc,
                           hipMemcpyHostToDevice, sec.stream));
    VCA_HIP(hipStreamSynchronize(sec.stream));
    sec.lz_cap = cap;
  }
}

// A chunk of Lanczos steps recorded once into a HIP graph and replayed:
// the kernels take the step index and the buffer parity from device
// memory, so the same graph serves every chunk of the run.  The environment
// variable VCA_NO_GRAPHS=1 launches every step eagerly instead.
static hipGraphExec_t step_graph(Sector &sec, int n, int nb, bool accx) {
  for (auto &g : sec.graphs)
    if (g.nsteps == n && g.nb == nb && g.accx == accx && g.real == sec.real_vec) return g.exec;
  hipGraph_t graph;
  VCA_HIP(hipStreamBeginCapture(sec.stream, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < n; ++i) {
    launch_hxv(sec, sec.buf[0], sec.buf[1], sec.buf[2], true, nb);
    launch_lanczos_update(sec, accx, nb);
  }
  VCA_HIP(hipStreamEndCapture(sec.stream, &graph));
  hipGraphExec_t ex;
  VCA_HIP(hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0));
  VCA_HIP(hipGraphDestroy(graph));
  sec.graphs.push_back(Sector::StepGraph{n, nb, accx, sec.real_vec, ex});
  return ex;
}

// Set on the worker threads of a concurrent sector sweep: while one thread
// captures a graph, the HIP runtime refuses device allocations on the
// other threads (hipMallocAsync returns "operation not permitted when
// stream is capturing" even with thread-local capture), so sweep workers
// launch their Lanczos steps eagerly; their launch latency is hidden by the
// other sectors running meanwhile.
static thread_local bool t_no_graphs = false;

static bool graphs_enabled() {
  if (t_no_graphs) return false;
  const char *e = std::getenv("VCA_NO_GRAPHS");
  return !(e && e[0] == '1');
}

static void run_steps(Sector &sec, int nsteps, bool accx, Solver *s, int nb = 1) {
  const int chunk = 16;
  ensure_step_partials(sec, nb);
  int done = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s) {
    VCA_HIP(hipEventCreate(&e0));
    VCA_HIP(hipEventCreate(&e1));
    VCA_HIP(hipEventRecord(e0, sec.stream));
  }
  if (nsteps >= chunk && graphs_enabled()) {
    hipGraphExec_t ex = step_graph(sec, chunk, nb, accx);
    for (; done + chunk <= nsteps; done += chunk) VCA_HIP(hipGraphLaunch(ex, sec.stream));
  }
  for (; done < nsteps; ++done) {
    launch_hxv(sec, sec.buf[0], sec.buf[1], sec.buf[2], true, nb);
    launch_lanczos_update(sec, accx, nb);
  }
  if (s) {
    VCA_HIP(hipEventRecord(e1, sec.stream));
    VCA_HIP(hipEventSynchronize(e1));
    float ms = 0;
    VCA_HIP(hipEventElapsedTime(&ms, e0, e1));
    VCA_HIP(hipEventDestroy(e0));
    VCA_HIP(hipEventDestroy(e1));
    std::lock_guard<std::mutex> lk(s->res_mu);
    s->res.hxv_seconds += ms * 1e-3;
    s->res.hxv_count += (int64_t)nsteps * nb;
    s->res.hxv_states += (double)nsteps * (double)sec.dim * nb;
  }
}

static void fetch(Sector &sec, const double *dsrc, int i0, int n, double *dst) {
  if (n <= 0) return;
  VCA_HIP(hipMemcpyAsync(dst + i0, dsrc + i0, sizeof(double) * n, hipMemcpyDeviceToHost,
                         sec.stream));
}

struct GsOut {
  double e0;
  int m;
  bool converged;
};

// Real-vector ground state (Sector::real_vec) for real Hamiltonians on the
// large sectors (those of the transposed H x v, where it halves seconds of
// work); smaller ones keep the complex Lanczos (and its start vector).
// VCA_REAL_GS=0 keeps complex vectors everywhere.
static bool use_real_gs(const Sector &sec) {
  const char *e = std::getenv("VCA_REAL_GS");
  return !(e && e[0] == '0') && sec.dim >= tform_min_dim() && real_mode_ok(sec);
}

// Plain Lanczos ground state on sec (its precision), start = device vector
// in that precision or null for the hash start vector.  The converged
// eigenvector is left normalised in sec.xvec.  rel_tol: stop when
// |dE| < tol*max(1,|E|) instead of |dE| < tol.  stall > 0: also stop once
// |dE| has not reached a new minimum for `stall` steps (after 20 steps),
// the noise floor of an FP32 run.  real: real-vector mode (start, if given,
// holds dim reals); the eigenvector in sec.xvec is then expanded to complex
// unless keep_real.
// VCA_GS_RESIDUAL=r: the FP64 ground-state Lanczos also requires the Ritz
// residual ||H x - E x|| (beta_{m+1} |y_m|, y the lowest eigenvector of T_m)
// to be at most r max(1, |E|) before it stops.  The |dE| < lanc_tolerance rule
// alone (the reference's, VCA_DIAG.f90:149-185) fixes E to ~tol but the vector
// only to ~sqrt(tol / gap), and G' is first order in the vector.  Unset: off.
static double gs_residual_tol() {
  const char *e = std::getenv("VCA_GS_RESIDUAL");
  if (!e || !*e) return 0.0;
  char *end = nullptr;
  const double r = std::strtod(e, &end);
  if (end == e || *end != '\0' || !(r > 0.0) || !std::isfinite(r))
    fail(VCA_ERR_ARG, "VCA_GS_RESIDUAL=%s: a positive number", e);
  return r;
}

static GsOut gs_lanczos(Sector &sec, double tol, int niter, const void *start, bool rel_tol,
                        Solver *s, int stall = 0, bool real = false, bool keep_real = false,
                        double rtol = 0.0) {
  niter = (int)std::min<int64_t>(niter, sec.dim);
  ensure_lanczos(sec, niter + 2);
  RealVecScope real_mode(sec, real);
  const size_t rsz = sec.precision == VCA_FP64 ? sizeof(double) : sizeof(float);
  auto init = [&]() {
    if (start)
      VCA_HIP(hipMemcpyAsync(sec.buf[0], start,
                             real ? (size_t)(2 * vec_elems(sec)) * rsz : sec.vec_bytes(),
                             hipMemcpyDeviceToDevice, sec.stream));
    else
      launch_start_vector(sec, 5001);
    if (real && (sec.dim & 1))  // the pad real of the last complex pair
      for (void *b : {sec.buf[0], sec.buf[1], sec.buf[2]})
        VCA_HIP(hipMemsetAsync(static_cast<char *>(b) + (size_t)sec.dim * rsz, 0, rsz,
                               sec.stream));
    launch_norm0(sec);
  };
  init();
  std::vector<double> al(niter + 4, 0.0), be(niter + 4, 0.0);
  // steps between convergence checks: every step past convergence in the
  // last chunk is wasted, so large sectors (ms per step, a check costs
  // ~0.1 ms) check every 4 steps; small ones every 16 (launch-bound steps)
  const int chunk = sec.dim >= (int64_t)1 << 22 ? 4 : 16;
  int k0 = 0, m = 0;
  double e_old = std::numeric_limits<double>::infinity(), e_m = 0;
  double best_de = std::numeric_limits<double>::infinity();
  int since_best = 0;
  bool conv = false;
  while (k0 < niter && !conv) {
    const int n = std::min(chunk, niter - k0);
    run_steps(sec, n, false, s);
    fetch(sec, sec.lzh.alpha, k0, n, al.data());
    fetch(sec, sec.lzh.beta, k0, n + 1, be.data());
    VCA_HIP(hipStreamSynchronize(sec.stream));
    for (int j = k0; j < k0 + n; ++j) {
      const int mm = j + 1;
      const double e = tridiag_lowest(mm, al.data(), be.data());
      const double thr = rel_tol ? tol * std::max(1.0, std::fabs(e)) : tol;
      const double de = std::fabs(e - e_old);
      if (de < best_de) {
        best_de = de;
        since_best = 0;
      } else {
        ++since_best;
      }
      bool stop = de < thr || be[j + 1] == 0.0 || (stall > 0 && mm >= 20 && since_best >= stall);
      if (stop && rtol > 0.0 && be[j + 1] != 0.0) {
        // y_m = the first component of the reversed tridiagonal's eigenvector
        std::vector<double> dr(mm), er(mm + 1, 0.0), w(mm), z0(mm);
        for (int i = 0; i < mm; ++i) dr[i] = al[mm - 1 - i];
        for (int i = 0; i + 1 < mm; ++i) er[i + 1] = be[mm - 1 - i];
        tridiag_eig_first(mm, dr.data(), er.data(), w.data(), z0.data());
        stop = be[j + 1] * std::fabs(z0[0]) <= rtol * std::max(1.0, std::fabs(e));
      }
      if (stop) {
        m = mm;
        e_m = e;
        conv = true;
        break;
      }
      e_old = e;
    }
    k0 += n;
  }
  if (!conv) {
    m = niter;
    e_m = tridiag_lowest(m, al.data(), be.data());
  }
  // second pass: x = sum_k y_k v_k
  std::vector<double> w(m), z((size_t)m * m);
  tridiag_eig(m, al.data(), be.data(), w.data(), z.data());
  VCA_HIP(hipMemcpyAsync(sec.lzh.y, z.data(), sizeof(double) * m, hipMemcpyHostToDevice,
                         sec.stream));
  init();
  VCA_HIP(hipMemsetAsync(sec.xvec, 0, sec.vec_bytes(), sec.stream));
  run_steps(sec, m, true, s);
  launch_scale_x(sec);
  if (real && !keep_real) {  // dim reals -> dim complex numbers (via buf[2])
    launch_expand_real(sec, sec.xvec, sec.buf[2]);
    VCA_HIP(hipMemcpyAsync(sec.xvec, sec.buf[2], sec.vec_bytes(), hipMemcpyDeviceToDevice,
                           sec.stream));
  }
  VCA_HIP(hipStreamSynchronize(sec.stream));
  return GsOut{e_m, m, conv};
}

// MIXED: FP32 Lanczos, then an FP64 Lanczos restarted from the FP32 ground
// state.  Returns the FP64 result; the eigenvector ends in sec64.xvec.
static GsOut gs_mixed(Solver &s, Sector &sec32, Sector &sec64, double tol, int niter,
                      double *e32) {
  // FP32 to its noise floor (a near-FP32-accurate vector), so that the FP64
  // restart starts within ~1e-6 of the ground state and needs few steps
  // real Hamiltonian: both runs on real vectors, the FP32 eigenvector
  // handed over as reals (converted as complex pairs)
  const bool real = use_real_gs(sec32) && use_real_gs(sec64);
  GsOut g32 = gs_lanczos(sec32, 1e-12, niter, nullptr, true, &s, 8, real, real);
  if (e32) *e32 = g32.e0;
  void *start64 = nullptr;
  VCA_HIP(hipMalloc(&start64, sec64.vec_bytes()));
  launch_convert(sec32.xvec, VCA_MIXED, start64, VCA_FP64,
                 real ? (sec32.dim + 1) / 2 : sec32.dim, sec32.stream);
  VCA_HIP(hipStreamSynchronize(sec32.stream));
  GsOut g64 = gs_lanczos(sec64, tol, niter, start64, false, &s, 0, real, false,
                         gs_residual_tol());
  VCA_HIP(hipFree(start64));
  return g64;
}

// Lanczos tridiagonalisation from the start vector already in buf[0] with
// its norm in nrm[0] (launch_excite or launch_norm0), nlanc steps.
static void tridiag_run(Sector &sec, int nlanc, double *alpha, double *beta, double *norm0,
                        Solver *s) {
  run_steps(sec, nlanc, false, s);
  fetch(sec, sec.lzh.alpha, 0, nlanc, alpha);
  fetch(sec, sec.lzh.beta, 0, nlanc, beta);
  fetch(sec, sec.lzh.nrm, 0, 1, norm0);
  VCA_HIP(hipStreamSynchronize(sec.stream));
}

// --------------------------------------------------------- dense path ---
static HostSpinTable host_spin_table(const SpinModel &m, int ns, int npart, double e_add) {
  Binom bn;
  HostSpinTable t;
  t.npart = npart;
  t.dim = (int)bn(ns, npart);
  const int nterms = (int)m.terms.size();
  t.maxnnz = std::max(1, std::min(nterms, npart * (ns - npart)));
  t.map.resize(t.dim);
  t.energy.resize(t.dim);
  t.cnt.resize(t.dim);
  t.ent.assign((size_t)t.dim * t.maxnnz, 0u);
  std::vector<int2> ab(std::max(1, nterms));
  for (int i = 0; i < nterms; ++i) ab[i] = make_int2(m.terms[i].a, m.terms[i].b);
  std::vector<double> pair((size_t)kMaxNs * kMaxNs);
  for (int a = 0; a < kMaxNs; ++a)
    for (int b = 0; b < kMaxNs; ++b) pair[(size_t)a * kMaxNs + b] = m.pair[a][b];
  for (int i = 0; i < t.dim; ++i) {
    const uint32_t str = unrank_string(ns, npart, (uint32_t)i, bn.c);
    t.map[i] = str;
    t.energy[i] = spin_energy(str, ns, m.eps, pair.data()) + e_add;
    t.cnt[i] = row_hops(str, nterms, ab.data(), bn.c, t.ent.data() + i, t.dim);
  }
  return t;
}

static double cross_host(const ModelData &md, uint32_t su, uint32_t sd) {
  double e = 0;
  for (int io = 0; io < md.norb; ++io)
    for (int jo = 0; jo < md.norb; ++jo)
      e += md.wx[io][jo] * (double)__builtin_popcount((su >> io) & (sd >> jo) & md.site_mask);
  return e;
}

static std::vector<zc> dense_h(const Solver &s, int nup, int ndw) {
  const ModelData &md = s.md;
  HostSpinTable tu = host_spin_table(md.spin[0], s.ns, nup, md.e_const);
  HostSpinTable td = host_spin_table(md.spin[1], s.ns, ndw, 0.0);
  const int64_t du = tu.dim, dd = td.dim, dim = du * dd;
  std::vector<zc> h((size_t)(dim * dim), zc(0.0));
  auto amp = [&](const SpinModel &m, uint32_t e) {
    const uint32_t ai = e >> kColBits;
    const Term &t = m.terms[ai >> 1];
    zc v(t.re, t.im);
    return (ai & 1u) ? -v : v;
  };
  for (int64_t idw = 0; idw < dd; ++idw)
    for (int64_t iup = 0; iup < du; ++iup) {
      const int64_t i = iup + idw * du;
      h[i + i * dim] += tu.energy[iup] + td.energy[idw] + cross_host(md, tu.map[iup], td.map[idw]);
      for (int q = 0; q < tu.cnt[iup]; ++q) {
        const uint32_t e = tu.ent[(size_t)q * du + iup];
        const int64_t j = (int64_t)(e & kColMask) + idw * du;
        h[i + j * dim] += amp(md.spin[0], e);
      }
      for (int q = 0; q < td.cnt[idw]; ++q) {
        const uint32_t e = td.ent[(size_t)q * dd + idw];
        const int64_t j = iup + (int64_t)(e & kColMask) * du;
        h[i + j * dim] += amp(md.spin[1], e);
      }
    }
  return h;
}

// --------------------------------------------------------- GS search ----
static bool finite_t(const vca_params &p) { return p.lanc_nstates_total > 1; }

// Eigenpairs of one sector (VCA_DIAG.f90:98-212): dense eigh when
// dim <= lanc_dim_threshold or Neigen == dim; else the GPU plain Lanczos
// (LANC_METHOD=lanczos, one state, FP64 or MIXED) or the GPU thick-restart
// Lanczos (ARPACK, FP64).  Eigenvectors are device complex128 vectors.
// The stream of the sweep worker running on this thread (null elsewhere).
static thread_local hipStream_t t_worker_stream = nullptr;

static void solve_sector(Solver &s, int nup, int ndw, int neigen, std::vector<double> &evals,
                         std::vector<void *> &evecs) {
  const vca_params &p = s.p;
  Binom bn;
  const int64_t dim = (int64_t)bn(s.ns, nup) * bn(s.ns, ndw);
  evals.clear();
  evecs.clear();
  const bool lanc_solve = !(neigen == dim || dim <= p.lanc_dim_threshold);
  if (!lanc_solve) {
    std::vector<zc> h = dense_h(s, nup, ndw);
    std::vector<double> w(dim);
    herm_eig((int)dim, h.data(), w.data());
    // copies on a stream other than the null stream, which would wait for
    // the sector solves running concurrently: the sweep's shared stream, or
    // one of their own
    hipStream_t st = t_worker_stream ? t_worker_stream : s.dense_stream;
    if (!st) VCA_HIP(hipStreamCreate(&st));
    for (int i = 0; i < neigen; ++i) {
      evals.push_back(w[i]);
      void *d = dev_alloc_async(sizeof(double2) * dim, st);
      VCA_HIP(hipMemcpyAsync(d, h.data() + (size_t)i * dim, sizeof(double2) * dim,
                             hipMemcpyHostToDevice, st));
      evecs.push_back(d);
    }
    VCA_HIP(hipStreamSynchronize(st));
    if (st != s.dense_stream && st != t_worker_stream) VCA_HIP(hipStreamDestroy(st));
    return;
  }
  if (p.lanc_method != VCA_LANC_LANCZOS) {
    auto s64 = make_sector(s, nup, ndw, VCA_FP64, t_worker_stream);
    const int ncv = (int)std::min<int64_t>(
        dim, (int64_t)p.lanc_ncv_factor * std::max(neigen, p.lanc_nstates_sector) + p.lanc_ncv_add);
    const int maxiter = (int)std::min<int64_t>(dim, p.lanc_niter);
    const EigsOut o =
        eigs_lanczos(*s64, neigen, ncv, p.lanc_tolerance, maxiter, nullptr, evals, evecs);
    if (o.nconv < neigen) {
      for (void *d : evecs) dev_free_async(d, s64->stream);
      evecs.clear();
      fail(VCA_ERR_NOCONV, "sector (%d,%d): %d of %d eigenpairs converged in %d restarts", nup,
           ndw, o.nconv, neigen, maxiter);
    }
    return;
  }
  GsOut g{};
  void *d = dev_alloc_async(sizeof(double2) * dim, s.stream);
  VCA_HIP(hipStreamSynchronize(s.stream));
  if (p.precision == VCA_MIXED) {
    auto s32 = make_sector(s, nup, ndw, VCA_MIXED, t_worker_stream);
    auto s64 = make_sector(s, nup, ndw, VCA_FP64, t_worker_stream);
    double e32;
    g = gs_mixed(s, *s32, *s64, p.lanc_tolerance, p.lanc_niter, &e32);
    {
      std::lock_guard<std::mutex> lk(s.res_mu);
      s.res.e0_mixed = e32;
    }
    VCA_HIP(hipMemcpyAsync(d, s64->xvec, sizeof(double2) * dim, hipMemcpyDeviceToDevice,
                           s64->stream));
    VCA_HIP(hipStreamSynchronize(s64->stream));
  } else {
    auto s64 = make_sector(s, nup, ndw, VCA_FP64, t_worker_stream);
    g = gs_lanczos(*s64, p.lanc_tolerance, p.lanc_niter, nullptr, false, &s, 0,
                   use_real_gs(*s64), false, gs_residual_tol());
    VCA_HIP(hipMemcpyAsync(d, s64->xvec, sizeof(double2) * dim, hipMemcpyDeviceToDevice,
                           s64->stream));
    VCA_HIP(hipStreamSynchronize(s64->stream));
  }
  evals.push_back(g.e0);
  evecs.push_back(d);
}

// The sector solves of one sweep of the diagonalisation are independent
// (VCA_DIAG.f90:83-256 fixes each sector's neigen before the sweep): worker
// threads solve them concurrently, each sector on streams of its own, while
// the caller takes the results in list order -- the order the state list is
// built in.  Small sectors are bound by the host round trips of their
// Lanczos steps, so overlapping them fills the GPU.  Workers claim sectors
// in list order, at most `nthreads` ahead of the caller, and within a
// device-memory budget (half the free HBM at the start of the sweep, the
// sector at the head of the list always admitted), so large sectors do not
// run out of memory together.  VCA_SECTOR_THREADS sets the worker count
// (default 16; 1: the sectors one after another on the calling thread).
// Runs fn(0..n-1) on up to 16 host threads (static slices); fn must only
// write per-index state.
template <typename Fn>
static void parallel_for(int n, Fn &&fn) {
  const int nt = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
  const int nth = std::min(nt, n);
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

// VCA_PROGRESS=1: one line on stderr per finished stage and chain batch,
// with the seconds since the solve began (long sectors run for minutes).
static bool progress_on() {
  static const bool on = [] {
    const char *e = std::getenv("VCA_PROGRESS");
    return e && e[0] == '1';
  }();
  return on;
}

static void progress(double t0, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static void progress(double t0, const char *fmt, ...) {
  if (!progress_on()) return;
  const double t =
      std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  std::fprintf(stderr, "vca [%8.2f s] ", t - t0);
  va_list ap;
  va_start(ap, fmt);
  std::vfprintf(stderr, fmt, ap);
  va_end(ap);
  std::fputc('\n', stderr);
  std::fflush(stderr);
}

// Thick-restart sectors below this size are solved together, their Lanczos
// steps in shared launches (eigs_lanczos_batch); VCA_EIGS_BATCH_MAX_DIM
// overrides (0: every sector on its own).
static int64_t eigs_batch_max_dim() {
  const char *e = std::getenv("VCA_EIGS_BATCH_MAX_DIM");
  return e && *e ? std::atoll(e) : ((int64_t)1 << 17);
}

class SectorSweep {
 public:
  SectorSweep(Solver &s, const std::vector<std::pair<int, int>> &list,
              const std::vector<int> &neig)
      : s_(s), list_(list), neig_(neig), slots_(list.size()), in_batch_(list.size(), false) {
    const char *env = std::getenv("VCA_SECTOR_THREADS");
    int nt = env && *env ? std::atoi(env) : 16;
    nt = std::max(1, std::min<int>(nt, (int)list.size()));
    nthreads_ = nt;
    if (nt <= 1) return;
    size_t free_b = 0, total_b = 0;
    VCA_HIP(hipMemGetInfo(&free_b, &total_b));
    budget_ = (int64_t)(free_b / 2);
    // the small thick-restart sectors: one batched solve on a thread of its own
    if (s.p.lanc_method != VCA_LANC_LANCZOS) {
      const int64_t cap = eigs_batch_max_dim();
      Binom bn;
      for (size_t k = 0; k < list.size(); ++k) {
        const int64_t dim = (int64_t)bn(s.ns, list[k].first) * bn(s.ns, list[k].second);
        const bool lanc_solve = !(neig[k] == dim || dim <= s.p.lanc_dim_threshold);
        if (lanc_solve && dim < cap) batch_.push_back(k);
      }
      if (batch_.size() < 2) batch_.clear();
      for (size_t k : batch_) in_batch_[k] = true;
    }
    if (!s.batch_stream) VCA_HIP(hipStreamCreate(&s.batch_stream));
    if (!s.dense_stream) VCA_HIP(hipStreamCreate(&s.dense_stream));
    while ((int)s.worker_streams.size() < nt) {
      hipStream_t st;
      VCA_HIP(hipStreamCreate(&st));
      s.worker_streams.push_back(st);
    }
    if (!batch_.empty()) batch_thread_ = std::thread([this]() { run_batch(); });
    for (int t = 0; t < nt; ++t) workers_.emplace_back([this, t]() { work(t); });
  }
  ~SectorSweep() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      abort_ = true;
    }
    cv_.notify_all();
    if (batch_thread_.joinable()) batch_thread_.join();
    for (auto &t : workers_) t.join();
    for (auto &sl : slots_)
      for (void *d : sl.evecs)
        if (d) (void)hipFreeAsync(d, s_.stream);
  }
  // Results of list entry k (k = 0, 1, ... in order); rethrows its error.
  void take(size_t k, std::vector<double> &evals, std::vector<void *> &evecs) {
    if (nthreads_ <= 1) {
      solve_sector(s_, list_[k].first, list_[k].second, neig_[k], evals, evecs);
      return;
    }
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&]() { return slots_[k].done; });
    Slot &sl = slots_[k];
    if (sl.err) {
      set_error("%s", sl.msg.c_str());  // the message of the worker's thread
      std::rethrow_exception(sl.err);
    }
    evals = std::move(sl.evals);
    evecs = std::move(sl.evecs);
    sl.evecs.clear();
    used_ -= sl.bytes;
    taken_ = k + 1;
    lk.unlock();
    cv_.notify_all();
  }

 private:
  struct Slot {
    std::vector<double> evals;
    std::vector<void *> evecs;
    std::exception_ptr err;
    std::string msg;
    bool done = false;
    int64_t bytes = 0;
  };
  // solved by the dense eigensolver (VCA_DIAG.f90:194)
  bool dense(size_t k) const {
    Binom bn;
    const int64_t dim = (int64_t)bn(s_.ns, list_[k].first) * bn(s_.ns, list_[k].second);
    return neig_[k] == dim || dim <= s_.p.lanc_dim_threshold;
  }
  // device bytes a sector solve holds: the thick-restart basis, rotation
  // target and eigenvectors, or the Lanczos buffers
  int64_t need(size_t k) const {
    Binom bn;
    const int64_t dim = (int64_t)bn(s_.ns, list_[k].first) * bn(s_.ns, list_[k].second);
    const int64_t m = (int64_t)s_.p.lanc_ncv_factor *
                          std::max(neig_[k], s_.p.lanc_nstates_sector) + s_.p.lanc_ncv_add;
    const int64_t vecs = s_.p.lanc_method == VCA_LANC_LANCZOS ? 8 : 2 * m + 1 + neig_[k];
    return vecs * dim * (int64_t)sizeof(double2);
  }
  // The batched sectors, in chunks within half the sweep's budget: their
  // memory is counted as the workers' is, but taken without waiting (the
  // caller releases memory in list order, and may be waiting for a batched
  // sector itself).
  void run_batch() {
    t_no_graphs = true;
    size_t b0 = 0;
    while (b0 < batch_.size()) {
      int64_t bytes = 0;
      size_t b1 = b0;
      while (b1 < batch_.size() && (b1 == b0 || bytes + need(batch_[b1]) <= budget_ / 2))
        bytes += need(batch_[b1++]);
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (abort_) return;
        used_ += bytes;
        for (size_t q = b0; q < b1; ++q) slots_[batch_[q]].bytes = need(batch_[q]);
      }
      std::vector<std::unique_ptr<Sector>> secs;
      std::vector<EigsJob> jobs(b1 - b0);
      int code = VCA_OK;
      std::string msg;
      try {
        // the sectors' tables, built on host threads (each on its own stream)
        secs.resize(b1 - b0);
        std::vector<std::exception_ptr> errs(b1 - b0);
        std::vector<std::string> emsg(b1 - b0);
        parallel_for((int)(b1 - b0), [&](int q) {
          const size_t k = batch_[b0 + q];
          try {
            secs[q] = make_sector(s_, list_[k].first, list_[k].second, VCA_FP64,
                                  s_.batch_stream);
          } catch (...) {
            errs[q] = std::current_exception();
            emsg[q] = g_err;
          }
        });
        for (size_t q = 0; q < errs.size(); ++q)
          if (errs[q]) {
            g_err = emsg[q];
            std::rethrow_exception(errs[q]);
          }
        for (size_t q = b0; q < b1; ++q) {
          const size_t k = batch_[q];
          const Sector &sec = *secs[q - b0];
          EigsJob &jb = jobs[q - b0];
          jb.sec = secs[q - b0].get();
          jb.nev = neig_[k];
          jb.ncv = (int)std::min<int64_t>(
              sec.dim, (int64_t)s_.p.lanc_ncv_factor *
                               std::max(neig_[k], s_.p.lanc_nstates_sector) +
                           s_.p.lanc_ncv_add);
          jb.maxiter = (int)std::min<int64_t>(sec.dim, s_.p.lanc_niter);
          jb.tol = s_.p.lanc_tolerance;
        }
        progress(s_.t_start, "eigs batch: %zu sectors built", b1 - b0);
        eigs_lanczos_batch(jobs, sector_cross(*secs[0]),
                           s_.md.spin[0].real_amp && s_.md.spin[1].real_amp, s_.batch_stream);
        progress(s_.t_start, "eigs batch: %zu sectors solved", b1 - b0);
      } catch (...) {
        code = VCA_ERR_HIP;
        try {
          throw;
        } catch (const Error &e) {
          code = e.code;
        } catch (...) {
        }
        msg = g_err;
      }
      for (size_t q = b0; q < b1; ++q) {
        const size_t k = batch_[q];
        EigsJob &jb = jobs[q - b0];
        Slot &sl = slots_[k];
        int c = code == VCA_OK ? jb.code : code;
        std::string m = code == VCA_OK ? jb.msg : msg;
        if (c == VCA_OK && jb.out.nconv < jb.nev) {
          c = VCA_ERR_NOCONV;
          char buf[256];
          std::snprintf(buf, sizeof buf, "sector (%d,%d): %d of %d eigenpairs converged in %d "
                        "restarts", list_[k].first, list_[k].second, jb.out.nconv, jb.nev,
                        jb.maxiter);
          m = buf;
        }
        if (c != VCA_OK) {
          for (void *d : jb.evecs) dev_free_async(d, s_.batch_stream);
          jb.evecs.clear();
          sl.err = std::make_exception_ptr(Error{c});
          sl.msg = m;
        } else {
          sl.evals = std::move(jb.evals);
          sl.evecs = std::move(jb.evecs);
        }
      }
      if (s_.batch_stream) (void)hipStreamSynchronize(s_.batch_stream);
      secs.clear();
      {
        std::lock_guard<std::mutex> lk(mu_);
        for (size_t q = b0; q < b1; ++q) slots_[batch_[q]].done = true;
      }
      cv_.notify_all();
      b0 = b1;
    }
  }

  void work(int t) {
    t_no_graphs = true;
    t_worker_stream = s_.worker_streams[t];
    for (;;) {
      size_t k;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&]() {
          while (next_ < list_.size() && in_batch_[next_]) ++next_;  // the batch's
          // dense sectors (a few hundred states) are not held to the window
          return abort_ || next_ >= list_.size() || next_ < taken_ + (size_t)nthreads_ ||
                 dense(next_);
        });
        if (abort_ || next_ >= list_.size()) return;
        k = next_++;
        const int64_t b = need(k);
        cv_.wait(lk, [&]() { return abort_ || used_ + b <= budget_ || k == taken_; });
        if (abort_) return;
        used_ += b;
        slots_[k].bytes = b;
      }
      Slot &sl = slots_[k];
      try {
        solve_sector(s_, list_[k].first, list_[k].second, neig_[k], sl.evals, sl.evecs);
      } catch (...) {
        sl.err = std::current_exception();
        sl.msg = g_err;
      }
      {
        std::lock_guard<std::mutex> lk(mu_);
        sl.done = true;
      }
      cv_.notify_all();
    }
  }

  Solver &s_;
  const std::vector<std::pair<int, int>> &list_;
  const std::vector<int> &neig_;
  std::vector<Slot> slots_;
  std::vector<bool> in_batch_;  // solved by the batch thread
  std::vector<size_t> batch_;
  std::thread batch_thread_;
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_;
  int nthreads_ = 1;
  size_t next_ = 0, taken_ = 0;
  int64_t used_ = 0, budget_ = 0;
  bool abort_ = false;
};

// es_insert_state (VCA_EIGENSPACE.f90:160-212): ascending energies, a new
// state goes before the first state of equal or higher energy.
static void es_insert(std::vector<KeptState> &list, const KeptState &st) {
  size_t pos = 0;
  while (pos < list.size() && !(st.e <= list[pos].e)) ++pos;
  list.insert(list.begin() + (long)pos, st);
}

static void es_pop(std::vector<KeptState> &list, hipStream_t st) {
  dev_free_async(list.back().d_vec, st);
  list.pop_back();
}

// es_add_state with a size limit (VCA_EIGENSPACE.f90:133-155): below the
// limit insert; at the limit replace the highest state if e is lower.
// Returns false when the state was not kept (the caller frees its vector).
static bool es_add_sized(std::vector<KeptState> &list, const KeptState &st, int size,
                         hipStream_t st_free) {
  if ((int)list.size() < size) {
    es_insert(list, st);
    return true;
  }
  if (!list.empty() && st.e < list.back().e) {
    es_pop(list, st_free);
    es_insert(list, st);
    return true;
  }
  return false;
}

// vca_diag_c (VCA_DIAG.f90:49-267) with vca_post_diag (:333-474): the
// sector scan, T = 0 ground-state selection or, at finite T, the
// size-limited state list with the spectrum annealing loop that adapts
// the states per sector and the list size until exp(-beta (Emax - Emin))
// < cutoff.
static void diagonalize(Solver &s) {
  const vca_params &p = s.p;
  const bool ft = finite_t(p);
  (void)gs_residual_tol();  // a bad value fails here, not in a sector worker
  free_states(s);
  std::vector<std::pair<int, int>> list;
  if (s.sectors.empty()) {
    for (int nup = 0; nup <= s.ns; ++nup)
      for (int ndw = 0; ndw <= s.ns; ++ndw) list.push_back({nup, ndw});
  } else {
    for (size_t i = 0; i + 1 < s.sectors.size(); i += 2)
      list.push_back({s.sectors[i], s.sectors[i + 1]});
  }
  Binom bn;
  // even state counts at finite T (VCA_SETUP.f90:165-176)
  int nst_sector = p.lanc_nstates_sector, nst_total = p.lanc_nstates_total;
  if (ft) {
    if (nst_sector % 2) ++nst_sector;
    if (nst_total % 2) ++nst_total;
  }
  std::map<std::pair<int, int>, int> neigen_sector;  // VCA_SETUP.f90:313
  for (int nup = 0; nup <= s.ns; ++nup)
    for (int ndw = 0; ndw <= s.ns; ++ndw) {
      const int64_t dim = (int64_t)bn(s.ns, nup) * bn(s.ns, ndw);
      neigen_sector[{nup, ndw}] = (int)std::min<int64_t>(dim, nst_sector);
    }
  bool converged = false;
  int iter = 0;
  double zeta = 1.0;
  while (!converged && iter < std::max(1, p.lanc_niter_spectrum)) {
    ++iter;
    free_states(s);
    double oldzero = 1000.0;
    std::vector<int> neig(list.size());
    for (size_t k = 0; k < list.size(); ++k) {
      const auto [nup, ndw] = list[k];
      const int64_t dim = (int64_t)bn(s.ns, nup) * bn(s.ns, ndw);
      neig[k] = p.lanc_method == VCA_LANC_LANCZOS
                    ? 1
                    : (int)std::min<int64_t>(dim, neigen_sector[{nup, ndw}]);
    }
    progress(s.t_start, "sweep: %zu sectors", list.size());
    SectorSweep sweep(s, list, neig);
    for (size_t k = 0; k < list.size(); ++k) {
      const auto [nup, ndw] = list[k];
      std::vector<double> evals;
      std::vector<void *> evecs;
      sweep.take(k, evals, evecs);
      for (size_t i = 0; i < evals.size(); ++i) {
        const KeptState st{evals[i], nup, ndw, evecs[i]};
        bool kept = false;
        if (ft) {
          kept = es_add_sized(s.states, st, nst_total, s.stream);
        } else {
          // T = 0 selection (VCA_DIAG.f90:224-235)
          const double en = evals[i];
          if (en < oldzero - 10.0 * p.gs_threshold) {
            oldzero = en;
            free_states(s);
            es_insert(s.states, st);
            kept = true;
          } else if (std::fabs(en - oldzero) <= p.gs_threshold) {
            oldzero = std::min(oldzero, en);
            es_insert(s.states, st);
            kept = true;
          }
        }
        if (kept) evecs[i] = nullptr;
      }
      for (void *d : evecs) dev_free_async(d, s.stream);
    }
    progress(s.t_start, "sweep: every sector taken");
    if (s.states.empty()) fail(VCA_ERR_NOCONV, "no ground state found");
    const double egs = s.states.front().e, emax = s.states.back().e;
    converged = ft ? std::exp(-p.beta * (emax - egs)) < p.cutoff : true;
    // vca_post_diag
    if (ft) {
      zeta = 0.0;
      for (auto &st : s.states) zeta += std::exp(-p.beta * (st.e - egs));
      std::map<std::pair<int, int>, int> count;
      for (auto &st : s.states) count[{st.nup, st.ndw}] += 1;
      for (auto &kv : neigen_sector) {
        const int c = count.count(kv.first) ? count[kv.first] : 0;
        int &ne = kv.second;
        ne += c > 0 ? 1 : -1;
        if (ne > c) ne = c + 1;
        if (ne <= 0) ne = 1;
      }
      if (std::exp(-p.beta * (emax - egs)) > p.cutoff) {
        nst_total += p.lanc_nstates_step;
      } else {
        while (s.states.size() > 1 &&
               std::exp(-p.beta * (s.states.back().e - egs)) <= p.cutoff)
          es_pop(s.states, s.stream);
        nst_total = std::max((int)s.states.size(), p.lanc_nstates_step) + p.lanc_nstates_step;
      }
    } else {
      zeta = (double)s.states.size();
    }
  }
  s.emin = s.states.front().e;
  s.zeta = zeta;
  s.res.nstates = (int)s.states.size();
  s.res.zeta = zeta;
  s.res.finite_t = ft ? 1 : 0;
  s.res.spectrum_iters = iter;
  if (ft) {
    double z = 0;
    for (auto &st : s.states) z += std::exp(-p.beta * (st.e - s.emin));
    s.res.omegaprime = s.emin - std::log(z) / p.beta;  // VCA_MAIN.f90:186-190
  } else {
    s.res.omegaprime = s.emin;
  }
}

// ----------------------------------------------------------- GF chains --
// One excitation chain of build_gf_normal (VCA_GF_NORMAL.f90:37-104):
// element, channel (1..4), pole sign, the spin the operators act on and
// the operators themselves.
struct ChainSpec {
  ElemKey key;
  int ichan, isign, spin;
  ExcOp op;
  int jnup, jndw;
};

// Chains of one kept state in the reference's order: for each spin,
// orbital and site, the diagonal channels c^+_i (1) and c_i (2), then for
// every other (site, orbital) the mixed channels (c^+_i + c^+_j) (1),
// (c_i + c_j) (2), (c^+_i + i c^+_j) (3), (c_i - i c_j) (4)
// (VCA_GF_NORMAL.f90:121-318, 561-958; 2 channels when gf_symmetric).
static std::vector<ChainSpec> chain_list(const Solver &s, const KeptState &st) {
  const vca_params &p = s.p;
  std::vector<ChainSpec> out;
  auto mk = [&](ElemKey key, int ichan, int isign, int spin, int pos1, int pos2, int dagger,
                double re2, double im2) {
    ChainSpec c{};
    c.key = key;
    c.ichan = ichan;
    c.isign = isign;
    c.spin = spin;
    c.op.nops = pos2 < 0 ? 1 : 2;
    c.op.pos[0] = pos1;
    c.op.dagger[0] = dagger;
    c.op.cre[0] = 1.0;
    c.op.cim[0] = 0.0;
    if (pos2 >= 0) {
      c.op.pos[1] = pos2;
      c.op.dagger[1] = dagger;
      c.op.cre[1] = re2;
      c.op.cim[1] = im2;
    }
    const int d = dagger ? 1 : -1;
    c.jnup = st.nup + (spin == 1 ? d : 0);
    c.jndw = st.ndw + (spin == 2 ? d : 0);
    out.push_back(c);
  };
  for (int ispin = 1; ispin <= p.nspin; ++ispin)
    for (int iorb = 1; iorb <= p.norb; ++iorb)
      for (int isite = 1; isite <= p.nlat; ++isite) {
        const int is = iorb - 1 + (isite - 1) * p.norb;
        mk({isite, isite, ispin, iorb, iorb}, 1, +1, ispin, is, -1, 1, 0, 0);
        mk({isite, isite, ispin, iorb, iorb}, 2, -1, ispin, is, -1, 0, 0, 0);
        for (int jsite = 1; jsite <= p.nlat; ++jsite)
          for (int jorb = 1; jorb <= p.norb; ++jorb) {
            if (isite == jsite && iorb == jorb) continue;
            const int js = jorb - 1 + (jsite - 1) * p.norb;
            const ElemKey k{isite, jsite, ispin, iorb, jorb};
            mk(k, 1, +1, ispin, is, js, 1, 1.0, 0.0);
            mk(k, 2, -1, ispin, is, js, 0, 1.0, 0.0);
            if (!p.gf_symmetric) {
              mk(k, 3, +1, ispin, is, js, 1, 0.0, 1.0);
              mk(k, 4, -1, ispin, is, js, 0, 0.0, -1.0);
            }
          }
      }
  return out;
}

// VCA_CHAIN_BATCH=n: at most n chains per launch (1: one chain at a time,
// the unbatched form); unset or 0: as many as half the free memory holds.
static int chain_batch_cap() {
  const char *e = std::getenv("VCA_CHAIN_BATCH");
  const int n = e ? std::atoi(e) : 0;
  return n > 0 ? n : 512;
}

// Chains of the same target sector run together: one launch per Lanczos
// step covers a batch of chains (grid.y = chain), their vectors stacked
// [chain][dim] in HBM, each chain with its own alpha/beta in device memory.
// A kept state whose imaginary part, after its global phase is removed, is
// below this fraction of its norm runs its chains on real vectors.  The
// ground-state Lanczos stops at |dE| < tol, which fixes the vector to
// ~sqrt(tol) (1e-6 at the default 1e-12); a complex combination of
// degenerate states is O(1) away.
constexpr double kRealStateTol = 1e-4;

// Target sectors below this size run their chains as one chain set (all
// such sectors of a kept state together, chains.hip); larger ones per
// sector with their tuned H x v.  VCA_CHAIN_SET_MAX_DIM overrides (0: off).
static int64_t chain_set_max_dim() {
  const char *e = std::getenv("VCA_CHAIN_SET_MAX_DIM");
  return e && *e ? std::atoll(e) : ((int64_t)1 << 17);
}

// Precision of the excitation chains: the solver's, or VCA_CHAIN_PRECISION
// = fp64 / mixed (a MIXED solve with FP64 chains keeps its FP32 ground-state
// Lanczos and FP32 log-dets, and its G' the accuracy of FP64 chains).
static int chain_precision(int prec) {
  const char *e = std::getenv("VCA_CHAIN_PRECISION");
  if (!e || !*e) return prec;
  if (std::strcmp(e, "fp64") == 0) return VCA_FP64;
  if (std::strcmp(e, "mixed") == 0) return VCA_MIXED;
  fail(VCA_ERR_ARG, "VCA_CHAIN_PRECISION=%s: fp64 or mixed", e);
}

static void build_gf(Solver &s) {
  const vca_params &p = s.p;
  s.gf.clear();
  s.max_exc = -std::numeric_limits<double>::infinity();
  const int prec = chain_precision(p.precision);
  std::map<std::array<int, 2>, std::unique_ptr<Sector>> cache;
  auto get_sector = [&](int nup, int ndw) -> Sector & {
    auto &slot = cache[{nup, ndw}];
    if (!slot) slot = make_sector(s, nup, ndw, prec);
    return *slot;
  };
  const double chan_re[4] = {1, 1, 0, 0}, chan_im[4] = {0, 0, -1, -1};
  Binom bn;
  for (size_t ist = 0; ist < s.states.size(); ++ist) {
    const KeptState &st = s.states[ist];
    const int dimup_i = (int)bn(s.ns, st.nup), dimdw_i = (int)bn(s.ns, st.ndw);
    const int64_t dim_i = (int64_t)dimup_i * dimdw_i;
    void *psi = st.d_vec;
    void *psi32 = nullptr;
    if (prec == VCA_MIXED) {
      VCA_HIP(hipMalloc(&psi32, sizeof(float2) * dim_i));
      launch_convert(st.d_vec, VCA_FP64, psi32, VCA_MIXED, dim_i, s.stream);
      psi = psi32;
    }
    VCA_HIP(hipStreamSynchronize(s.stream));
    // Real Hamiltonian: the chains run on real vectors (k_hxv_buf_real or the
    // one-plane transposed form), from the ground state with its global
    // phase removed -- G' does not depend on it.  Kept complex when the
    // state is not a real vector times a phase (a complex combination of
    // degenerate states) or VCA_REAL_CHAINS=0.
    void *psi_r = nullptr;
    const char *rc_env = std::getenv("VCA_REAL_CHAINS");
    if (!(rc_env && rc_env[0] == '0') && s.md.spin[0].real_amp && s.md.spin[1].real_amp) {
      const size_t rsz = prec == VCA_FP64 ? sizeof(double) : sizeof(float);
      VCA_HIP(hipMalloc(&psi_r, rsz * (size_t)dim_i));
      const double imf = real_vector_of(st.d_vec, dim_i, psi_r, prec, s.stream);
      if (!(imf < kRealStateTol)) {
        VCA_HIP(hipFree(psi_r));
        psi_r = nullptr;
      }
    }
    const std::vector<ChainSpec> specs = chain_list(s, st);
    std::vector<GfChan> res(specs.size());
    const int64_t cs_max = chain_set_max_dim();
    auto target_ok = [&](const ChainSpec &cs) {
      return cs.jnup >= 0 && cs.jndw >= 0 && cs.jnup <= s.ns && cs.jndw <= s.ns;
    };
    // Channels 3 and 4, (c^+_i + i c^+_j) and (c_i - i c_j), of a real
    // Hamiltonian and a real state: their start vector is a + i b (a - i b)
    // with a, b the real start vectors of the diagonal chains of i and j in
    // the same channel and target sector, and its spectral measure is the
    // sum of theirs (the cross terms cancel: (z - H)^-1 is symmetric), so
    // the chain's n-step tridiagonal is lanczos_of_sum of those two chains
    // (equal to the chain run from a + i b up to rounding).  Taken so where
    // the diagonal chains run on real vectors and the target sector is
    // larger than the chain (no exhausted Krylov space);
    // VCA_DERIVED_CHANNELS=0 runs every chain.
    std::vector<std::array<int, 2>> derived(specs.size(), {-1, -1});
    std::vector<char> keep(specs.size(), 0);  // diagonal chains the derived ones read
    {
      const char *dc = std::getenv("VCA_DERIVED_CHANNELS");
      if (psi_r && !(dc && dc[0] == '0')) {
        std::map<std::array<int, 3>, int> diag;  // (spin, channel, position) -> chain
        for (size_t c = 0; c < specs.size(); ++c)
          if (specs[c].op.nops == 1) diag[{specs[c].spin, specs[c].ichan, specs[c].op.pos[0]}] = (int)c;
        for (size_t c = 0; c < specs.size(); ++c) {
          const ChainSpec &cs = specs[c];
          if (cs.ichan < 3 || cs.op.nops != 2 || !target_ok(cs)) continue;
          const auto ia = diag.find({cs.spin, cs.ichan - 2, cs.op.pos[0]});
          const auto ib = diag.find({cs.spin, cs.ichan - 2, cs.op.pos[1]});
          if (ia == diag.end() || ib == diag.end()) continue;
          Sector &js = get_sector(cs.jnup, cs.jndw);
          if (js.dim <= (int64_t)p.lanc_ngfiter) continue;
          if (js.dim >= cs_max && !real_mode_ok(js)) continue;  // complex diagonal chains
          derived[c] = {ia->second, ib->second};
          keep[ia->second] = keep[ib->second] = 1;
        }
      }
    }
    // the diagonal chains' quadratures (Ritz values, first components)
    std::vector<std::vector<double>> kept_x(specs.size()), kept_z(specs.size());
    std::vector<double> kept_nrm(specs.size(), 0.0);
    std::map<std::array<int, 2>, std::vector<int>> groups;
    for (size_t c = 0; c < specs.size(); ++c) {
      const ChainSpec &cs = specs[c];
      if (!target_ok(cs) || derived[c][0] >= 0) continue;
      groups[{cs.jnup, cs.jndw}].push_back((int)c);
    }
    // pesoBZ (VCA_GF_NORMAL.f90:987-994): Boltzmann factor at finite T
    double boltz = 1.0;
    if (finite_t(p)) {
      const double x = p.beta * (st.e - s.emin);
      boltz = x < 200.0 ? std::exp(-x) : 0.0;
    }
    // add_to_lanczos_gf_normal (VCA_GF_NORMAL.f90:970-1022) for chain c:
    // the tridiagonal eigensolve of its nlanc alpha/beta and the poles and
    // weights of its channel; returns the largest excitation energy
    auto post_chain = [&](int c, int nlanc, const double *al, const double *be,
                          double nrm0) -> double {
      const ChainSpec &cs = specs[c];
      const double norm2 = nrm0 * nrm0;
      double cmax = -std::numeric_limits<double>::infinity();
      if (!(norm2 > 0.0)) return cmax;  // O|psi> = 0: empty channel
      std::vector<double> w(nlanc), z(nlanc);
      tridiag_eig_first(nlanc, al, be, w.data(), z.data());
      if (keep[c]) {
        kept_x[c] = w;
        kept_z[c] = z;
        kept_nrm[c] = nrm0;
      }
      const zc peso_bz =
          zc(chan_re[cs.ichan - 1], chan_im[cs.ichan - 1]) * norm2 * boltz / s.zeta;
      GfChan &out = res[c];
      out.w.resize(nlanc);
      out.p.resize(nlanc);
      for (int j = 0; j < nlanc; ++j) {
        const double de = w[j] - st.e;
        cmax = std::max(cmax, de);
        out.w[j] = peso_bz * z[j] * z[j];
        out.p[j] = cs.isign * de;
      }
      return cmax;
    };
    // chains on small target sectors: one chain set, within half the free
    // memory per run
    std::vector<ChainJob> jobs;
    std::vector<int> job_spec;
    const bool real_ham = psi_r != nullptr;
    for (auto it = groups.begin(); it != groups.end();) {
      Sector &js = get_sector(it->first[0], it->first[1]);
      if (js.dim >= cs_max) {
        ++it;
        continue;
      }
      const int nlanc = (int)std::min<int64_t>(js.dim, p.lanc_ngfiter);
      for (int c : it->second) {
        bool real_op = real_ham;
        for (int o = 0; o < specs[c].op.nops; ++o) real_op = real_op && specs[c].op.cim[o] == 0.0;
        jobs.push_back(ChainJob{&js, real_op, nlanc, specs[c].op, specs[c].spin});
        job_spec.push_back(c);
      }
      it = groups.erase(it);
    }
    for (size_t j0 = 0; j0 < jobs.size();) {
      size_t free_b = 0, total_b = 0;
      VCA_HIP(hipMemGetInfo(&free_b, &total_b));
      const int64_t budget = (int64_t)(free_b / 2);
      int64_t used = 0;
      size_t j1 = j0;
      while (j1 < jobs.size()) {
        const int64_t b = chain_set_bytes(*jobs[j1].js, jobs[j1].real, prec);
        if (j1 > j0 && used + b > budget) break;
        used += b;
        ++j1;
      }
      const std::vector<ChainJob> part(jobs.begin() + j0, jobs.begin() + j1);
      std::vector<ChainOut> outs;
      const ChainSetStats cst = run_chain_set(part, prec, dimup_i, dimdw_i, psi, psi_r,
                                              sector_cross(*part[0].js), s.md.spin[0].real_amp &&
                                              s.md.spin[1].real_amp, s.stream, outs);
      std::vector<double> cmax(part.size());
      parallel_for((int)part.size(), [&](int q) {
        cmax[q] = post_chain(job_spec[j0 + q], part[q].nsteps, outs[q].alpha.data(),
                             outs[q].beta.data(), outs[q].nrm0);
      });
      for (double c : cmax) s.max_exc = std::max(s.max_exc, c);
      s.res.nchains += (int)part.size();
      s.res.hxv_seconds += cst.seconds;
      s.res.hxv_count += cst.hxv_count;
      s.res.hxv_states += cst.hxv_states;
      progress(s.t_start, "state %zu chain set: chains %zu-%zu of %zu, %d steps, %lld launches, "
               "%d segments in column items, %.3f s", ist, j0 + 1, j1, jobs.size(), cst.steps,
               (long long)cst.launches, cst.column_items, cst.seconds);
      j0 = j1;
    }
    for (auto &g : groups) {
      Sector &js = get_sector(g.first[0], g.first[1]);
      const int nlanc = (int)std::min<int64_t>(js.dim, p.lanc_ngfiter);
      size_t free_b = 0, total_b = 0;
      VCA_HIP(hipMemGetInfo(&free_b, &total_b));
      // three Lanczos vectors per chain, plus the H x v shape's scratch
      const size_t per_chain = (size_t)(3 + hxv_scratch_vectors(js)) * js.vec_bytes();
      const size_t have = (size_t)js.nb_cap * 3 * js.vec_bytes() +
                          (size_t)js.tbuf_cap * js.vec_bytes();
      const size_t budget = (free_b + have) / 2;
      int nb_max = (int)std::max<size_t>(1, budget / per_chain);
      nb_max = std::min(nb_max, chain_batch_cap());
      // the chains whose operator has real coefficients run on real vectors
      // (real Hamiltonian, a real state); the others (the c_i - i c_j
      // channels) on complex ones, in batches of their own
      std::vector<int> part[2];  // [0] complex, [1] real vectors
      const bool real_ok = psi_r && real_mode_ok(js);
      for (int c : g.second) {
        bool real_op = real_ok;
        for (int o = 0; o < specs[c].op.nops; ++o) real_op = real_op && specs[c].op.cim[o] == 0.0;
        part[real_op ? 1 : 0].push_back(c);
      }
      for (int pr = 0; pr < 2; ++pr) {
      const std::vector<int> &idx = part[pr];
      for (size_t b0 = 0; b0 < idx.size(); b0 += nb_max) {
        const int nb = (int)std::min<size_t>(nb_max, idx.size() - b0);
        ensure_lanczos(js, nlanc + 2, std::max(nb, js.nb_cap));
        std::vector<ExcOp> ops(nb);
        for (int c = 0; c < nb; ++c) ops[c] = specs[idx[b0 + c]].op;
        VCA_HIP(hipMemcpyAsync(js.d_ops, ops.data(), sizeof(ExcOp) * nb, hipMemcpyHostToDevice,
                               js.stream));
        RealVecScope real_mode(js, pr == 1);
        hxv_tune_batch(js, nb);
        if (js.real_vec && (js.dim & 1))  // the pad reals of Hv and u_{k-1} stay 0
          for (int i = 1; i < 3; ++i)
            VCA_HIP(hipMemsetAsync(js.buf[i], 0, js.vec_bytes() * nb, js.stream));
        launch_excite(js, dimup_i, dimdw_i, js.real_vec ? psi_r : psi, specs[idx[b0]].spin,
                      js.d_ops, nb);
        run_steps(js, nlanc, false, &s, nb);
        std::vector<double> al((size_t)nb * nlanc), be((size_t)nb * nlanc), nrm0(nb);
        for (int c = 0; c < nb; ++c) {
          VCA_HIP(hipMemcpyAsync(al.data() + (size_t)c * nlanc, js.lzh_all[c].alpha,
                                 sizeof(double) * nlanc, hipMemcpyDeviceToHost, js.stream));
          VCA_HIP(hipMemcpyAsync(be.data() + (size_t)c * nlanc, js.lzh_all[c].beta,
                                 sizeof(double) * nlanc, hipMemcpyDeviceToHost, js.stream));
          VCA_HIP(hipMemcpyAsync(&nrm0[c], js.lzh_all[c].nrm, sizeof(double),
                                 hipMemcpyDeviceToHost, js.stream));
        }
        VCA_HIP(hipStreamSynchronize(js.stream));
        // the chains' tridiagonal eigensolves run on host threads
        std::vector<double> cmax(nb);
        parallel_for(nb, [&](int c) {
          cmax[c] = post_chain(idx[b0 + c], nlanc, al.data() + (size_t)c * nlanc,
                               be.data() + (size_t)c * nlanc, nrm0[c]);
        });
        for (int c = 0; c < nb; ++c) s.max_exc = std::max(s.max_exc, cmax[c]);
        s.res.nchains += nb;
        progress(s.t_start, "state %zu sector (%d,%d) %s chains %zu-%zu of %zu, %d steps",
                 ist, g.first[0], g.first[1], pr ? "real" : "complex", b0 + 1, b0 + nb,
                 idx.size(), nlanc);
      }
      }  // the complex and the real-vector part
    }
    std::vector<int> dlist;
    for (size_t c = 0; c < specs.size(); ++c)
      if (derived[c][0] >= 0) dlist.push_back((int)c);
    if (!dlist.empty()) {
      std::vector<double> cmax(dlist.size(), -std::numeric_limits<double>::infinity());
      parallel_for((int)dlist.size(), [&](int q) {
        const int c = dlist[q], a = derived[c][0], b = derived[c][1];
        // an empty diagonal chain (norm 0) keeps no quadrature; the two
        // chains of one target sector have the same length
        const size_t na = kept_x[a].size(), nb = kept_x[b].size();
        const int n = (int)(na && nb ? std::min(na, nb) : std::max(na, nb));
        if (n == 0) return;  // both empty: so is the channel
        std::vector<double> al(n), be(n);
        const double nrm = lanczos_of_quadratures(n, kept_x[a].data(), kept_z[a].data(),
                                                  kept_nrm[a], kept_x[b].data(),
                                                  kept_z[b].data(), kept_nrm[b], al.data(),
                                                  be.data());
        cmax[q] = post_chain(c, n, al.data(), be.data(), nrm);
      });
      for (double c : cmax) s.max_exc = std::max(s.max_exc, c);
      s.res.nchains_derived += (int)dlist.size();
      progress(s.t_start, "state %zu: %zu c_i - i c_j channels from their diagonal chains", ist,
               dlist.size());
    }
    for (size_t c = 0; c < specs.size(); ++c) {
      auto &el = s.gf[specs[c].key];
      el.resize(s.states.size());
      el[ist].push_back(std::move(res[c]));
    }
    if (psi32) VCA_HIP(hipFree(psi32));
    if (psi_r) VCA_HIP(hipFree(psi_r));
  }
}

// -------------------------------------------------------------- Omega ---
static zc hk_at(const Solver &s, int il, int jl, int si, int sj, int io, int jo, int64_t k) {
  const int nl = s.p.nlat, nsp = s.p.nspin, no = s.p.norb;
  const size_t per_k = (size_t)nl * nl * nsp * nsp * no * no;
  return s.hk[idx6(il, jl, si, sj, io, jo, nl, nl, nsp, nsp, no) + per_k * (size_t)k];
}

static void prepare_omega(Solver &s) {
  const vca_params &p = s.p;
  const int nl = p.nlat, no = p.norb, nsp = p.nspin, nlb = p.nlat_bath, nob = p.norb_bath;
  const bool bath = nlb > 0 && nob > 0;
  // spin-block structure
  bool blockdiag = true;
  for (int si = 0; si < nsp && blockdiag; ++si)
    for (int sj = 0; sj < nsp && blockdiag; ++sj) {
      if (si == sj) continue;
      for (int il = 0; il < nl; ++il)
        for (int jl = 0; jl < nl; ++jl)
          for (int io = 0; io < no; ++io)
            for (int jo = 0; jo < no; ++jo) {
              if (s.hloc[idx6(il, jl, si, sj, io, jo, nl, nl, nsp, nsp, no)] != zc(0.0))
                blockdiag = false;
              for (int64_t k = 0; k < s.nk && blockdiag; ++k)
                if (hk_at(s, il, jl, si, sj, io, jo, k) != zc(0.0)) blockdiag = false;
            }
      if (bath) {
        for (int il = 0; il < nlb; ++il)
          for (int jl = 0; jl < nlb; ++jl)
            for (int io = 0; io < nob; ++io)
              for (int jo = 0; jo < nob; ++jo)
                if (s.bath_h[idx6(il, jl, si, sj, io, jo, nlb, nlb, nsp, nsp, nob)] != zc(0.0))
                  blockdiag = false;
        for (int il = 0; il < nl; ++il)
          for (int jl = 0; jl < nlb; ++jl)
            for (int io = 0; io < no; ++io)
              for (int jo = 0; jo < nob; ++jo)
                if (s.bath_v[idx6(il, jl, si, sj, io, jo, nl, nlb, nsp, nsp, no)] != zc(0.0))
                  blockdiag = false;
      }
    }
  OmegaHost h;
  h.nlat = nl;
  h.norb = no;
  h.nspin = nsp;
  h.nspin_blk = blockdiag ? nsp : 1;
  h.nb = blockdiag ? nl * no : nl * no * nsp;
  h.nbb = bath ? (blockdiag ? nlb * nob : nlb * nob * nsp) : 0;
  if (h.nb > 16 || h.nbb > 16)
    fail(VCA_ERR_ARG, "Omega: block size %d / bath block %d above 16 is not supported", h.nb,
         h.nbb);
  h.nk = (int)s.nk;
  h.chan4 = p.gf_symmetric ? 0.0 : 1.0;
  h.xmu = p.xmu;
  // pole store in element order (ispin, ilat, jlat, iorb, jorb)
  h.nelem = nsp * nl * nl * no * no;
  h.off.assign(h.nelem + 1, 0);
  for (int sp = 0; sp < nsp; ++sp)
    for (int il = 0; il < nl; ++il)
      for (int jl = 0; jl < nl; ++jl)
        for (int io = 0; io < no; ++io)
          for (int jo = 0; jo < no; ++jo) {
            const int e = (((sp * nl + il) * nl + jl) * no + io) * no + jo;
            auto it = s.gf.find({il + 1, jl + 1, sp + 1, io + 1, jo + 1});
            if (it != s.gf.end())
              for (auto &chans : it->second)
                for (auto &c : chans)
                  for (size_t j = 0; j < c.w.size(); ++j) {
                    h.w.push_back(make_double2(c.w[j].real(), c.w[j].imag()));
                    h.pole.push_back(c.p[j]);
                  }
            h.off[e + 1] = (int64_t)h.w.size();
          }
  // lso position inside a block: io + il*no (+ spin*nl*no when spin-mixed)
  auto lso = [&](int il, int io, int sp) { return io + il * no + (blockdiag ? 0 : sp * nl * no); };
  h.hk.assign((size_t)s.nk * h.nspin_blk * 256, make_double2(0, 0));
  h.hloc.assign((size_t)h.nspin_blk * h.nb * h.nb, make_double2(0, 0));
  for (int si = 0; si < nsp; ++si)
    for (int sj = 0; sj < nsp; ++sj) {
      if (blockdiag && si != sj) continue;
      const int sb = blockdiag ? si : 0;
      for (int il = 0; il < nl; ++il)
        for (int jl = 0; jl < nl; ++jl)
          for (int io = 0; io < no; ++io)
            for (int jo = 0; jo < no; ++jo) {
              const int r = lso(il, io, si), c = lso(jl, jo, sj);
              const zc hv = s.hloc[idx6(il, jl, si, sj, io, jo, nl, nl, nsp, nsp, no)];
              h.hloc[((size_t)sb * h.nb + r) * h.nb + c] = make_double2(hv.real(), hv.imag());
              for (int64_t k = 0; k < s.nk; ++k) {
                const zc v = hk_at(s, il, jl, si, sj, io, jo, k);
                h.hk[(((size_t)k * h.nspin_blk + sb) * 16 + r) * 16 + c] =
                    make_double2(v.real(), v.imag());
              }
            }
      if (bath) {
        if (sb == 0 && h.hb.empty()) {
          h.hb.assign((size_t)h.nspin_blk * h.nbb * h.nbb, make_double2(0, 0));
          h.vb.assign((size_t)h.nspin_blk * h.nb * h.nbb, make_double2(0, 0));
        }
        auto blso = [&](int ib, int ob, int sp) {
          return ob + ib * nob + (blockdiag ? 0 : sp * nlb * nob);
        };
        for (int ib = 0; ib < nlb; ++ib)
          for (int jb = 0; jb < nlb; ++jb)
            for (int ob = 0; ob < nob; ++ob)
              for (int pb = 0; pb < nob; ++pb) {
                const zc v = s.bath_h[idx6(ib, jb, si, sj, ob, pb, nlb, nlb, nsp, nsp, nob)];
                h.hb[((size_t)sb * h.nbb + blso(ib, ob, si)) * h.nbb + blso(jb, pb, sj)] =
                    make_double2(v.real(), v.imag());
              }
        for (int il = 0; il < nl; ++il)
          for (int jb = 0; jb < nlb; ++jb)
            for (int io = 0; io < no; ++io)
              for (int pb = 0; pb < nob; ++pb) {
                const zc v = s.bath_v[idx6(il, jb, si, sj, io, pb, nl, nlb, nsp, nsp, no)];
                h.vb[((size_t)sb * h.nb + lso(il, io, si)) * h.nbb + blso(jb, pb, sj)] =
                    make_double2(v.real(), v.imag());
              }
      }
    }
  s.res.npoles = (int64_t)h.w.size();
  if (!s.om) s.om = std::make_unique<OmegaDev>();
  s.om->upload(h, s.stream);
  s.om_ready = true;
}

static std::vector<double> integrand(Solver &s, const std::vector<double2> &z, int prec,
                                     std::vector<double> *fim = nullptr) {
  std::vector<double> f(z.size());
  if (fim) fim->assign(z.size(), 0.0);
  const int batch = fim ? 256 : 1024;
  for (size_t i = 0; i < z.size(); i += batch) {
    const int n = (int)std::min<size_t>(batch, z.size() - i);
    float ms = 0;
    s.om->integrand(n, z.data() + i, prec, f.data() + i, &ms, fim ? fim->data() + i : nullptr);
    s.res.omega_seconds += ms * 1e-3;
    s.res.omega_evals += (int64_t)n * s.nk;
  }
  return f;
}

// Gauss-Kronrod 7/15 abscissae and weights (Kronrod 1965, as tabulated in
// QUADPACK's qk15): x descending from the end point to the centre.
static const double kXgk[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                               0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                               0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                               0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
static const double kWgk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                               0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                               0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                               0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double kWg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                              0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

struct Panel {
  double a, b, val, err;
};

// The 15 Kronrod nodes of [a, b].
static void panel_nodes(double a, double b, double *t) {
  const double c = 0.5 * (a + b), hl = 0.5 * (b - a);
  for (int j = 0; j < 7; ++j) {
    t[j] = c - hl * kXgk[j];
    t[14 - j] = c + hl * kXgk[j];
  }
  t[7] = c;
}

static void panel_eval(Panel &pn, const double *g /*15 values at the nodes*/) {
  const double hl = 0.5 * (pn.b - pn.a);
  double rk = kWgk[7] * g[7], rg = kWg[3] * g[7];
  for (int j = 0; j < 7; ++j) {
    rk += kWgk[j] * (g[j] + g[14 - j]);
    if (j & 1) rg += kWg[j / 2] * (g[j] + g[14 - j]);
  }
  pn.val = rk * hl;
  pn.err = std::fabs((rk - rg) * hl);
}

// Adaptive Gauss-Kronrod 7/15 of g on [a, b]; eval(t) returns g at a batch
// of nodes.  Each round splits the largest-error panel and every panel
// above the average share of the tolerance, until the summed error
// estimate is below max(epsabs, epsrel |I|).
template <typename Eval>
static double adapt_gk15(double a, double b, double epsabs, double epsrel, Eval &&eval) {
  std::vector<Panel> panels;
  auto evaluate = [&](std::vector<Panel> &ps) {
    std::vector<double> t(ps.size() * 15);
    for (size_t i = 0; i < ps.size(); ++i) panel_nodes(ps[i].a, ps[i].b, t.data() + i * 15);
    const std::vector<double> g = eval(t);
    for (size_t i = 0; i < ps.size(); ++i) panel_eval(ps[i], g.data() + i * 15);
  };
  std::vector<Panel> init = {{a, b, 0, 0}};
  evaluate(init);
  panels = init;
  for (int round = 0; round < 200; ++round) {
    double tot = 0, err = 0;
    for (auto &pn : panels) {
      tot += pn.val;
      err += pn.err;
    }
    const double tol = std::max(epsabs, epsrel * std::fabs(tot));
    if (err <= tol || panels.size() > 20000) break;
    const double thr = tol / (double)panels.size();
    std::vector<Panel> keep, fresh;
    size_t imax = 0;
    for (size_t i = 1; i < panels.size(); ++i)
      if (panels[i].err > panels[imax].err) imax = i;
    for (size_t i = 0; i < panels.size(); ++i) {
      const Panel &pn = panels[i];
      if (i == imax || pn.err > thr) {
        const double m = 0.5 * (pn.a + pn.b);
        fresh.push_back({pn.a, m, 0, 0});
        fresh.push_back({m, pn.b, 0, 0});
      } else {
        keep.push_back(pn);
      }
    }
    evaluate(fresh);
    keep.insert(keep.end(), fresh.begin(), fresh.end());
    panels.swap(keep);
  }
  double tot = 0;
  for (auto &pn : panels) tot += pn.val;
  return tot;
}

// Quadrature tolerances for an integrand of precision prec.  FP32 log-dets
// carry ~1e-8 relative noise (the exact-identity LU), so an adaptive rule
// asked for 1e-12 would bisect to its panel limit chasing the noise (the
// 3x4 solve: 7.5 s of Omega instead of 0.04 s): in MIXED the relative
// target is floored at 5e-8 of the integral, which costs at most ~1e-9 of
// Omega itself.
static void quad_tol(const vca_params &p, int prec, double *epsabs, double *epsrel) {
  *epsabs = p.quad_epsabs;
  *epsrel = p.quad_epsrel;
  if (prec == VCA_MIXED) *epsrel = std::max(*epsrel, 5e-8);
}

// int_0^inf f(i x) dx through the QAGI map x = (1-t)/t, t in (0, 1]
// (QUADPACK qagi, called through SciFortran quad at VCA_OMEGA.f90:132).
static double quad_imag_axis(Solver &s, int prec) {
  double ea, er;
  quad_tol(s.p, prec, &ea, &er);
  return adapt_gk15(0.0, 1.0, ea, er, [&](const std::vector<double> &t) {
    std::vector<double2> z(t.size());
    for (size_t i = 0; i < t.size(); ++i) z[i] = make_double2(0.0, (1.0 - t[i]) / t[i]);
    std::vector<double> g = integrand(s, z, prec);
    for (size_t i = 0; i < t.size(); ++i) g[i] /= t[i] * t[i];
    return g;
  });
}

// frequency_integration_finite_t (VCA_OMEGA.f90:153-188): the Matsubara
// sum (2/beta) sum_{n=0}^{nmax} Re F(i w_n) plus the contour integral
// int_0^pi Re[(1/pi) w f(w) F(w)] dzeta on w = R e^{i zeta} (R = 2(nmax+1)
// pi/beta, f the Fermi function, integral_contour :191-204), times
// (3 - Nspin); F = sum_kmesh_complex.
static double omega_finite_t(Solver &s, int prec) {
  const vca_params &p = s.p;
  const double beta = p.beta;
  int nmax = (int)(2.0 * (std::fabs(s.max_exc) + p.bandwidth) * beta / M_PI);
  nmax = (nmax % 2 == 0) ? nmax / 2 : (nmax + 1) / 2;
  const double radius = 2.0 * (nmax + 1) * M_PI / beta;
  s.res.nmax = nmax;
  std::vector<double2> zm(nmax + 1);
  for (int n = 0; n <= nmax; ++n) zm[n] = make_double2(0.0, (2.0 * n + 1.0) * M_PI / beta);
  const std::vector<double> fm = integrand(s, zm, prec);
  double sum = 0;
  for (double v : fm) sum += v;
  sum *= 2.0 / beta;
  double ea, er;
  quad_tol(p, prec, &ea, &er);
  const double contour =
      adapt_gk15(0.0, M_PI, ea, er, [&](const std::vector<double> &t) {
        std::vector<double2> z(t.size());
        for (size_t i = 0; i < t.size(); ++i)
          z[i] = make_double2(radius * std::cos(t[i]), radius * std::sin(t[i]));
        std::vector<double> fim;
        const std::vector<double> fre = integrand(s, z, prec, &fim);
        std::vector<double> g(t.size());
        for (size_t i = 0; i < t.size(); ++i) {
          const zc w(z[i].x, z[i].y);
          zc fermi(0.0, 0.0);
          if ((w.real() - p.xmu) * beta < 100.0) fermi = 1.0 / (std::exp(beta * (w - p.xmu)) + 1.0);
          g[i] = ((1.0 / M_PI) * w * fermi * zc(fre[i], fim[i])).real();
        }
        return g;
      });
  return (3.0 - p.nspin) * (sum + contour);
}

// Precision of the Omega log-dets in vca_solve: the solver's, or
// VCA_OMEGA_PRECISION = fp64 / mixed.
static int omega_precision(int prec) {
  const char *e = std::getenv("VCA_OMEGA_PRECISION");
  if (!e || !*e) return prec;
  if (std::strcmp(e, "fp64") == 0) return VCA_FP64;
  if (std::strcmp(e, "mixed") == 0) return VCA_MIXED;
  fail(VCA_ERR_ARG, "VCA_OMEGA_PRECISION=%s: fp64 or mixed", e);
}

static void omega_stage(Solver &s) {
  const vca_params &p = s.p;
  prepare_omega(s);
  const int prec = omega_precision(p.precision);
  double integral;
  if (finite_t(p)) {
    integral = omega_finite_t(s, prec);
  } else if (p.omega_mode == VCA_OMEGA_MATSUBARA) {
    std::vector<double2> z(p.omega_nfreq);
    for (int n = 0; n < p.omega_nfreq; ++n)
      z[n] = make_double2(0.0, (2.0 * n + 1.0) * M_PI / p.beta);
    std::vector<double> f = integrand(s, z, prec);
    double sum = 0;
    for (double v : f) sum += v;
    integral = (3.0 - p.nspin) * 2.0 / p.beta * sum;
  } else {
    integral = (3.0 - p.nspin) * quad_imag_axis(s, prec) / M_PI;
  }
  s.res.omega_integral = integral;
  s.res.sft_potential = s.res.omegaprime - integral;
  s.res.max_exc = s.max_exc;
}

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace vca

// ================================================================ C ABI ==
using namespace vca;

// ---------------------------------------- no-allocation H x v scratch ---
// Device scratch of vca_hxv_device_noalloc: the transposed form's v^T for
// one vector, in a file-scope device array instead of the sector's lazily
// allocated Sector::tbuf.  The runtime places one copy per GPU when it loads
// this code object; its address on the current device is looked up once
// (hipGetSymbolAddress) by vca_init_solver, and host code only ever uses
// that looked-up pointer.  Every call on one GPU shares it, so calls that
// overlap on one GPU race; nothing in it needs to start at zero (the
// transpose writes every element before the gather pass reads it).
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static void *g_scratch_dev[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// The current device's address of `scratch`, looked up on its first use.
static void resolve_scratch() {
  int dev = 0;
  VCA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) fail(VCA_ERR_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[dev]) return;
  void *ptr = nullptr;
  VCA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) fail(VCA_ERR_HIP, "no device address for the scratch array");
  g_scratch_dev[dev] = ptr;
}

#define VCA_GUARD_BEGIN try {
#define VCA_GUARD_END                         \
  }                                           \
  catch (const Error &e) {                    \
    return e.code;                            \
  }                                           \
  catch (const std::bad_alloc &) {            \
    set_error("host allocation failed");      \
    return VCA_ERR_NOMEM;                     \
  }                                           \
  catch (...) {                               \
    set_error("unexpected exception");        \
    return VCA_ERR_ARG;                       \
  }                                           \
  return VCA_OK;

extern "C" {

void vca_default_params(vca_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->nlat = 1;
  p->norb = 1;
  p->nspin = 1;
  p->hfmode = 1;
  p->hfshift = 1;
  p->uloc[0] = 2.0;
  p->beta = 1000.0;
  p->lanc_method = VCA_LANC_ARPACK;
  p->lanc_nstates_sector = 6;
  p->lanc_niter = 512;
  p->lanc_ngfiter = 200;
  p->lanc_dim_threshold = 256;
  p->lanc_tolerance = 1e-12;
  p->gs_threshold = 1e-9;
  p->precision = VCA_FP64;
  p->omega_mode = VCA_OMEGA_QUAD;
  p->omega_nfreq = 4096;
  p->quad_epsabs = 1e-10;
  p->quad_epsrel = 1e-12;
  p->lanc_nstates_total = 1;
  p->lanc_nstates_step = 2;
  p->lanc_niter_spectrum = 10;
  p->lanc_ncv_factor = 3;
  p->lanc_ncv_add = 5;
  p->cutoff = 1e-9;
  p->bandwidth = 0.0;
  p->lmats = 5000;
  p->lreal = 5000;
  p->wini = -5.0;
  p->wfin = 5.0;
  p->eps = 0.01;
}

const char *vca_last_error(void) { return g_err.c_str(); }

int vca_init_solver(const vca_params *p, vca_solver **out) {
  VCA_GUARD_BEGIN
  if (!p || !out) fail(VCA_ERR_ARG, "null argument");
  *out = nullptr;
  if (p->nlat < 1 || p->norb < 1 || p->norb > 5 || (p->nspin != 1 && p->nspin != 2) ||
      p->nlat_bath < 0 || p->norb_bath < 0)
    fail(VCA_ERR_ARG, "bad dimensions nlat=%d norb=%d nspin=%d", p->nlat, p->norb, p->nspin);
  const int ns = p->nlat * p->norb + p->nlat_bath * p->norb_bath;
  if (ns > kMaxNs) fail(VCA_ERR_ARG, "Ns=%d above the supported %d", ns, kMaxNs);
  if (p->precision != VCA_FP64 && p->precision != VCA_MIXED)
    fail(VCA_ERR_ARG, "unknown precision %d", p->precision);
  auto s = std::make_unique<Solver>();
  s->p = *p;
  if (p->nsectors > 0) {
    if (!p->sectors) fail(VCA_ERR_ARG, "nsectors > 0 with a null sector list");
    s->sectors.assign(p->sectors, p->sectors + 2 * p->nsectors);
  }
  s->p.sectors = nullptr;
  s->ns = ns;
  s->nlso = p->nlat * p->nspin * p->norb;
  int ndev = 0;
  VCA_HIP(hipGetDeviceCount(&ndev));
  if (ndev < 1) fail(VCA_ERR_HIP, "no HIP device");
  VCA_HIP(hipSetDevice(0));
  resolve_scratch();
  VCA_HIP(hipStreamCreate(&s->stream));
  Binom bn;
  VCA_HIP(hipMalloc(&s->d_binom, sizeof(bn.c)));
  VCA_HIP(hipMemcpy(s->d_binom, bn.c, sizeof(bn.c), hipMemcpyHostToDevice));
  *out = reinterpret_cast<vca_solver *>(s.release());
  VCA_GUARD_END
}

void vca_finalize_solver(vca_solver *s) { delete reinterpret_cast<Solver *>(s); }

int vca_set_hcluster(vca_solver *sv, const double *hloc) {
  VCA_GUARD_BEGIN
  Solver *s = reinterpret_cast<Solver *>(sv);
  if (!s || !hloc) fail(VCA_ERR_ARG, "null argument");
  const size_t n = (size_t)s->nlso * s->nlso;
  s->hloc.assign(reinterpret_cast<const zc *>(hloc), reinterpret_cast<const zc *>(hloc) + n);
  s->has_hloc = true;
  if (s->p.nlat_bath * s->p.norb_bath == 0 || s->has_bath) derive_model(*s);
  VCA_GUARD_END
}

int vca_set_bath(vca_solver *sv, const double *bath_h, const double *bath_v) {
  VCA_GUARD_BEGIN
  Solver *s = reinterpret_cast<Solver *>(sv);
  if (!s || !bath_h || !bath_v) fail(VCA_ERR_ARG, "null argument");
  const vca_params &p = s->p;
  const int nbl = p.nlat_bath * p.nspin * p.norb_bath;
  const size_t nh = (size_t)nbl * nbl, nv = (size_t)s->nlso * nbl;
  if (nbl == 0) fail(VCA_ERR_ARG, "model has no bath");
  s->bath_h.assign(reinterpret_cast<const zc *>(bath_h), reinterpret_cast<const zc *>(bath_h) + nh);
  s->bath_v.assign(reinterpret_cast<const zc *>(bath_v), reinterpret_cast<const zc *>(bath_v) + nv);
  // VCA_BATH_SETUP.f90:127-136
  if (p.hfmode && !p.hfshift) {
    const int nlb = p.nlat_bath, nsp = p.nspin, nob = p.norb_bath;
    for (int il = 0; il < nlb; ++il)
      for (int io = 0; io < nob; ++io)
        for (int sp = 0; sp < nsp; ++sp)
          s->bath_h[idx6(il, il, sp, sp, io, io, nlb, nlb, nsp, nsp, nob)] -=
              0.5 * p.uloc[io] + 0.5 * p.ust * (p.norb - 1) + 0.5 * (p.ust - p.jh) * (p.norb - 1);
  }
  s->has_bath = true;
  if (s->has_hloc) derive_model(*s);
  VCA_GUARD_END
}

int vca_solve(vca_solver *sv, const double *hloc, const double *hk, int64_t nk,
              const double *bath_h, const double *bath_v) {
  VCA_GUARD_BEGIN
  Solver *s = reinterpret_cast<Solver *>(sv);
  if (!s || !hloc || !hk || nk < 1) fail(VCA_ERR_ARG, "null argument or nk < 1");
  const bool bath = s->p.nlat_bath * s->p.norb_bath > 0;
  if (bath && (!bath_h || !bath_v)) fail(VCA_ERR_ARG, "model has a bath: bath_h/bath_v needed");
  s->res = vca_results{};
  s->solved = false;
  s->om_ready = false;
  if (bath) {
    int rc = vca_set_bath(sv, bath_h, bath_v);
    if (rc) return rc;
  }
  int rc = vca_set_hcluster(sv, hloc);
  if (rc) return rc;
  const size_t per_k = (size_t)s->nlso * s->nlso;
  s->hk.assign(reinterpret_cast<const zc *>(hk), reinterpret_cast<const zc *>(hk) + per_k * nk);
  s->nk = nk;
  double t0 = now();
  s->t_start = t0;
  diagonalize(*s);
  double t1 = now();
  progress(t0, "ground state: %zu kept states", s->states.size());
  build_gf(*s);
  double t2 = now();
  progress(t0, "excitation chains: %d", s->res.nchains);
  omega_stage(*s);
  double t3 = now();
  progress(t0, "Omega: %lld (k, z) evaluations", (long long)s->res.omega_evals);
  s->res.t_diag = t1 - t0;
  s->res.t_gf = t2 - t1;
  s->res.t_omega = t3 - t2;
  s->solved = true;
  VCA_GUARD_END
}

int vca_get_sft_potential(const vca_solver *sv, double *out) {
  VCA_GUARD_BEGIN
  const Solver *s = reinterpret_cast<const Solver *>(sv);
  if (!s || !out) fail(VCA_ERR_ARG, "null argument");
  if (!s->solved) fail(VCA_ERR_STATE, "vca_solve has not completed");
  *out = s->res.sft_potential;
  VCA_GUARD_END
}

int vca_get_results(const vca_solver *sv, vca_results *out) {
  VCA_GUARD_BEGIN
  const Solver *s = reinterpret_cast<const Solver *>(sv);
  if (!s || !out) fail(VCA_ERR_ARG, "null argument");
  *out = s->res;
  VCA_GUARD_END
}

int vca_sector_create(vca_solver *sv, int32_t nup, int32_t ndw, int32_t precision,
                      vca_sector **out) {
  VCA_GUARD_BEGIN
  Solver *s = reinterpret_cast<Solver *>(sv);
  if (!s || !out) fail(VCA_ERR_ARG, "null argument");
  if (precision != VCA_FP64 && precision != VCA_MIXED) fail(VCA_ERR_ARG, "bad precision");
  *out = reinterpret_cast<vca_sector *>(make_sector(*s, nup, ndw, precision).release());
  VCA_GUARD_END
}

void vca_sector_destroy(vca_sector *sec) { delete reinterpret_cast<Sector *>(sec); }

int vca_sector_dims(const vca_sector *sv, int64_t *dim, int32_t *dimup, int32_t *dimdw) {
  VCA_GUARD_BEGIN
  const Sector *sec = reinterpret_cast<const Sector *>(sv);
  if (!sec) fail(VCA_ERR_ARG, "null sector");
  if (dim) *dim = sec->dim;
  if (dimup) *dimup = sec->up.dim;
  if (dimdw) *dimdw = sec->dw.dim;
  VCA_GUARD_END
}

int vca_sector_variant(const vca_sector *sv, int32_t *hxv_variant, int32_t *real_variant) {
  VCA_GUARD_BEGIN
  const Sector *sec = reinterpret_cast<const Sector *>(sv);
  if (!sec) fail(VCA_ERR_ARG, "null sector");
  if (hxv_variant) *hxv_variant = sec->hxv_variant;
  if (real_variant) *real_variant = sec->real_variant;
  VCA_GUARD_END
}

int vca_hxv(vca_sector *sv, int64_t n, const double *v, double *hv) {
  VCA_GUARD_BEGIN
  Sector *sec = reinterpret_cast<Sector *>(sv);
  if (!sec || !v || !hv) fail(VCA_ERR_ARG, "null argument");
  if (n != sec->dim) fail(VCA_ERR_ARG, "vector length %lld != sector dimension %lld",
                          (long long)n, (long long)sec->dim);
  void *d64 = nullptr, *din = nullptr, *dout = nullptr;
  const size_t b64 = sizeof(double2) * n;
  VCA_HIP(hipMalloc(&d64, b64));
  VCA_HIP(hipMalloc(&din, sec->vec_bytes()));
  VCA_HIP(hipMalloc(&dout, sec->vec_bytes()));
  VCA_HIP(hipMemcpyAsync(d64, v, b64, hipMemcpyHostToDevice, sec->stream));
  launch_convert(d64, VCA_FP64, din, sec->precision, n, sec->stream);
  launch_hxv(*sec, din, din, dout, false);
  launch_convert(dout, sec->precision, d64, VCA_FP64, n, sec->stream);
  VCA_HIP(hipMemcpyAsync(hv, d64, b64, hipMemcpyDeviceToHost, sec->stream));
  VCA_HIP(hipStreamSynchronize(sec->stream));
  VCA_HIP(hipFree(d64));
  VCA_HIP(hipFree(din));
  VCA_HIP(hipFree(dout));
  VCA_GUARD_END
}

int vca_hxv_device(vca_sector *sv, const void *v, void *hv, int32_t repeat, float *ms) {
  VCA_GUARD_BEGIN
  Sector *sec = reinterpret_cast<Sector *>(sv);
  if (!sec || !v || !hv) fail(VCA_ERR_ARG, "null argument");
  if (repeat < 1) repeat = 1;
  hipEvent_t e0, e1;
  VCA_HIP(hipEventCreate(&e0));
  VCA_HIP(hipEventCreate(&e1));
  VCA_HIP(hipEventRecord(e0, sec->stream));
  for (int r = 0; r < repeat; ++r) launch_hxv(*sec, v, v, hv, false);
  VCA_HIP(hipEventRecord(e1, sec->stream));
  VCA_HIP(hipEventSynchronize(e1));
  if (ms) {
    float t = 0;
    VCA_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t / repeat;
  }
  VCA_HIP(hipEventDestroy(e0));
  VCA_HIP(hipEventDestroy(e1));
  VCA_GUARD_END
}

int vca_device_alloc(int64_t bytes, void **ptr) {
  VCA_GUARD_BEGIN
  if (!ptr || bytes < 0) fail(VCA_ERR_ARG, "bad argument");
  *ptr = nullptr;
  if (bytes > 0) VCA_HIP(hipMalloc(ptr, (size_t)bytes));
  VCA_GUARD_END
}

int vca_device_synchronize(void) {
  VCA_GUARD_BEGIN
  VCA_HIP(hipDeviceSynchronize());
  VCA_GUARD_END
}

int vca_device_free(void *ptr) {
  VCA_GUARD_BEGIN
  if (ptr) VCA_HIP(hipFree(ptr));
  VCA_GUARD_END
}

int vca_device_copy(void *dst, const void *src, int64_t bytes, int32_t kind) {
  VCA_GUARD_BEGIN
  if (!dst || !src || bytes < 0) fail(VCA_ERR_ARG, "bad argument");
  hipMemcpyKind k;
  switch (kind) {
    case VCA_COPY_H2D: k = hipMemcpyHostToDevice; break;
    case VCA_COPY_D2H: k = hipMemcpyDeviceToHost; break;
    case VCA_COPY_D2D: k = hipMemcpyDeviceToDevice; break;
    default: fail(VCA_ERR_ARG, "unknown copy kind %d", kind);
  }
  VCA_HIP(hipMemcpy(dst, src, (size_t)bytes, k));
  VCA_GUARD_END
}

int vca_hxv_device_noalloc(vca_sector *sv, const void *v, void *hv, void *stream) {
  VCA_GUARD_BEGIN
  Sector *sec = reinterpret_cast<Sector *>(sv);
  if (!sec || !v || !hv) fail(VCA_ERR_ARG, "null argument");
  int dev = 0;
  VCA_HIP(hipGetDevice(&dev));
  void *ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (dev >= 0 && dev < kMaxDevices) ptr = g_scratch_dev[dev];
  }
  if (!ptr) fail(VCA_ERR_STATE, "no scratch on device %d (vca_init_solver sets it up)", dev);
  const size_t need = (size_t)hxv_scratch_vectors(*sec) * sec->vec_bytes();
  if (need > kNoAllocScratchBytes)
    fail(VCA_ERR_ARG,
         "sector (%d,%d) needs %zu bytes of scratch; the no-allocation H x v holds %zu "
         "(use vca_hxv_device)", sec->nup, sec->ndw, need, kNoAllocScratchBytes);
  // launch on the caller's stream (e.g. one being captured), the sector's
  // otherwise
  const hipStream_t own = sec->stream;
  if (stream) sec->stream = static_cast<hipStream_t>(stream);
  try {
    launch_hxv(*sec, v, v, hv, false, 1, need ? ptr : nullptr);
  } catch (...) {
    sec->stream = own;
    throw;
  }
  sec->stream = own;
  VCA_GUARD_END
}

int vca_hxv_device_batch(vca_sector *sv, int32_t nb, const void *v, void *hv, int32_t repeat,
                         float *ms) {
  VCA_GUARD_BEGIN
  Sector *sec = reinterpret_cast<Sector *>(sv);
  if (!sec || !v || !hv || nb < 1) fail(VCA_ERR_ARG, "bad argument");
  if (repeat < 1) repeat = 1;
  hipEvent_t e0, e1;
  VCA_HIP(hipEventCreate(&e0));
  VCA_HIP(hipEventCreate(&e1));
  VCA_HIP(hipEventRecord(e0, sec->stream));
  for (int r = 0; r < repeat; ++r) launch_hxv(*sec, v, v, hv, false, nb);
  VCA_HIP(hipEventRecord(e1, sec->stream));
  VCA_HIP(hipEventSynchronize(e1));
  if (ms) {
    float t = 0;
    VCA_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t / repeat;
  }
  VCA_HIP(hipEventDestroy(e0));
  VCA_HIP(hipEventDestroy(e1));
  VCA_GUARD_END
}

int vca_hxv_device_real(vca_sector *sv, const void *v, void *hv, int32_t repeat, float *ms) {
  VCA_GUARD_BEGIN
  Sector *sec = reinterpret_cast<Sector *>(sv);
  if (!sec || !v || !hv) fail(VCA_ERR_ARG, "null argument");
  if (!real_mode_ok(*sec))
    fail(VCA_ERR_ARG,
         "sector (%d,%d): the real-vector H x v needs real amplitudes and a sector large "
         "enough for the transposed form",
         sec->nup, sec->ndw);
  if (repeat < 1) repeat = 1;
  RealVecScope real_mode(*sec, true);
  tune_real_direct(*sec);
  hipEvent_t e0, e1;
  VCA_HIP(hipEventCreate(&e0));
  VCA_HIP(hipEventCreate(&e1));
  VCA_HIP(hipEventRecord(e0, sec->stream));
  for (int r = 0; r < repeat; ++r) launch_hxv(*sec, v, v, hv, false);
  VCA_HIP(hipEventRecord(e1, sec->stream));
  VCA_HIP(hipEventSynchronize(e1));
  if (ms) {
    float t = 0;
    VCA_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t / repeat;
  }
  VCA_HIP(hipEventDestroy(e0));
  VCA_HIP(hipEventDestroy(e1));
  VCA_GUARD_END
}

int vca_sector_spin_coo(const vca_sector *sv, int32_t spin, int64_t *nnz, int32_t *rows,
                        int32_t *cols, double *amps) {
  VCA_GUARD_BEGIN
  const Sector *sec = reinterpret_cast<const Sector *>(sv);
  if (!sec || !nnz || (spin != 1 && spin != 2)) fail(VCA_ERR_ARG, "bad argument");
  const DevSpin &sp = spin == 1 ? sec->up : sec->dw;
  const SpinModel &m = sec->solver->md.spin[spin - 1];
  std::vector<int32_t> cnt(sp.dim);
  std::vector<uint32_t> ent((size_t)sp.dim * sp.maxnnz);
  VCA_HIP(hipMemcpy(cnt.data(), sp.cnt, sizeof(int32_t) * sp.dim, hipMemcpyDeviceToHost));
  VCA_HIP(hipMemcpy(ent.data(), sp.ent, sizeof(uint32_t) * ent.size(), hipMemcpyDeviceToHost));
  int64_t n = 0;
  for (int i = 0; i < sp.dim; ++i)
    for (int q = 0; q < cnt[i]; ++q) {
      if (rows) {
        const uint32_t e = ent[(size_t)q * sp.dim + i];
        const uint32_t ai = e >> kColBits;
        const Term &t = m.terms[ai >> 1];
        const double sg = (ai & 1u) ? -1.0 : 1.0;
        rows[n] = i;
        cols[n] = (int32_t)(e & kColMask);
        amps[2 * n] = sg * t.re;
        amps[2 * n + 1] = sg * t.im;
      }
      ++n;
    }
  *nnz = n;
  VCA_GUARD_END
}

int vca_lanczos_gs(vca_sector *sv, double tol, int32_t niter, double *e0, double *vec,
                   int32_t *iters) {
  VCA_GUARD_BEGIN
  Sector *sec = reinterpret_cast<Sector *>(sv);
  if (!sec || !e0) fail(VCA_ERR_ARG, "null argument");
  GsOut g{};
  std::unique_ptr<Sector> twin;
  Sector *res_sec = sec;
  if (sec->precision == VCA_MIXED) {
    twin = make_sector(*sec->solver, sec->nup, sec->ndw, VCA_FP64);
    g = gs_mixed(*sec->solver, *sec, *twin, tol, niter, nullptr);
    res_sec = twin.get();
  } else {
    g = gs_lanczos(*sec, tol, niter, nullptr, false, nullptr, 0, use_real_gs(*sec));
  }
  *e0 = g.e0;
  if (iters) *iters = g.m;
  if (vec)
    VCA_HIP(hipMemcpy(vec, res_sec->xvec, sizeof(double2) * sec->dim, hipMemcpyDeviceToHost));
  VCA_GUARD_END
}

int vca_lanczos_eigs(vca_sector *sv, int32_t nev, int32_t ncv, double tol, int32_t maxiter,
                     double *evals, double *evecs, int32_t *nconv) {
  VCA_GUARD_BEGIN
  Sector *sec = reinterpret_cast<Sector *>(sv);
  if (!sec || !evals) fail(VCA_ERR_ARG, "null argument");
  std::unique_ptr<Sector> twin;
  Sector *s64 = sec;
  if (sec->precision != VCA_FP64) {
    twin = make_sector(*sec->solver, sec->nup, sec->ndw, VCA_FP64);
    s64 = twin.get();
  }
  std::vector<double> ev;
  std::vector<void *> vecs;
  const EigsOut o = eigs_lanczos(*s64, nev, ncv, tol, maxiter, nullptr, ev, vecs);
  for (int i = 0; i < nev; ++i) evals[i] = ev[i];
  for (int i = 0; i < nev; ++i) {
    if (evecs)
      VCA_HIP(hipMemcpy(evecs + (size_t)2 * i * sec->dim, vecs[i], sizeof(double2) * sec->dim,
                        hipMemcpyDeviceToHost));
    dev_free_async(vecs[i], s64->stream);
  }
  if (nconv) *nconv = o.nconv;
  VCA_GUARD_END
}

int vca_lanczos_of_sum(int32_t nlanc, const double *a1, const double *b1, double n1,
                       const double *a2, const double *b2, double n2, double *alpha,
                       double *beta, double *nrm) {
  VCA_GUARD_BEGIN
  if (nlanc < 1 || !a1 || !b1 || !a2 || !b2 || !alpha || !beta || !nrm || !(n1 >= 0.0) ||
      !(n2 >= 0.0))
    fail(VCA_ERR_ARG, "bad argument");
  *nrm = lanczos_of_sum(nlanc, a1, b1, n1, a2, b2, n2, alpha, beta);
  VCA_GUARD_END
}

int vca_lanczos_tridiag(vca_sector *sv, const double *v0, int32_t nlanc, double *alpha,
                        double *beta) {
  VCA_GUARD_BEGIN
  Sector *sec = reinterpret_cast<Sector *>(sv);
  if (!sec || !v0 || !alpha || !beta || nlanc < 1) fail(VCA_ERR_ARG, "bad argument");
  ensure_lanczos(*sec, nlanc + 2);
  // a real start vector on a real Hamiltonian: the chain on real vectors,
  // as the excitation chains of such models run (VCA_REAL_CHAINS=0: complex)
  bool real_v0 = true;
  for (int64_t i = 0; i < sec->dim && real_v0; ++i) real_v0 = v0[2 * i + 1] == 0.0;
  const char *rc_env = std::getenv("VCA_REAL_CHAINS");
  RealVecScope real_mode(*sec, real_v0 && !(rc_env && rc_env[0] == '0') && real_mode_ok(*sec));
  void *d64 = nullptr;
  VCA_HIP(hipMalloc(&d64, sizeof(double2) * sec->dim));
  if (sec->real_vec) {
    // dim reals, the pad of an odd dim 0, converted as complex pairs
    std::vector<double> re(2 * (size_t)vec_elems(*sec), 0.0);
    for (int64_t i = 0; i < sec->dim; ++i) re[i] = v0[2 * i];
    VCA_HIP(hipMemcpy(d64, re.data(), sizeof(double) * re.size(), hipMemcpyHostToDevice));
    if (sec->dim & 1)
      for (int i = 1; i < 3; ++i)
        VCA_HIP(hipMemsetAsync(sec->buf[i], 0, sec->vec_bytes(), sec->stream));
  } else {
    VCA_HIP(hipMemcpyAsync(d64, v0, sizeof(double2) * sec->dim, hipMemcpyHostToDevice,
                           sec->stream));
  }
  launch_convert(d64, VCA_FP64, sec->buf[0], sec->precision, vec_elems(*sec), sec->stream);
  launch_norm0(*sec);
  double nrm0;
  tridiag_run(*sec, nlanc, alpha, beta, &nrm0, nullptr);
  VCA_HIP(hipFree(d64));
  VCA_GUARD_END
}

int vca_gf_poles(const vca_solver *sv, int32_t ilat, int32_t jlat, int32_t ispin, int32_t iorb,
                 int32_t jorb, int32_t istate, int32_t ichan, int32_t *n, double *weights,
                 double *poles) {
  VCA_GUARD_BEGIN
  const Solver *s = reinterpret_cast<const Solver *>(sv);
  if (!s || !n) fail(VCA_ERR_ARG, "null argument");
  if (!s->solved && !s->gf_loaded) fail(VCA_ERR_STATE, "no Green's function: vca_solve or vca_read_gfprime first");
  auto it = s->gf.find({ilat, jlat, ispin, iorb, jorb});
  if (it == s->gf.end()) fail(VCA_ERR_ARG, "no such element");
  if (istate < 1 || istate > (int)it->second.size()) fail(VCA_ERR_ARG, "bad state index");
  const auto &chans = it->second[istate - 1];
  if (ichan < 1 || ichan > (int)chans.size()) fail(VCA_ERR_ARG, "bad channel index");
  const GfChan &c = chans[ichan - 1];
  *n = (int32_t)c.w.size();
  if (weights)
    for (size_t j = 0; j < c.w.size(); ++j) {
      weights[2 * j] = c.w[j].real();
      weights[2 * j + 1] = c.w[j].imag();
    }
  if (poles)
    for (size_t j = 0; j < c.p.size(); ++j) poles[j] = c.p[j];
  VCA_GUARD_END
}

int vca_gf_cluster(const vca_solver *sv, double zre, double zim, double *gfo) {
  VCA_GUARD_BEGIN
  const Solver *s = reinterpret_cast<const Solver *>(sv);
  if (!s || !gfo) fail(VCA_ERR_ARG, "null argument");
  if (!s->solved && !s->gf_loaded) fail(VCA_ERR_STATE, "no Green's function: vca_solve or vca_read_gfprime first");
  const vca_params &p = s->p;
  const int nl = p.nlat, no = p.norb, nsp = p.nspin;
  const zc z(zre, zim);
  std::vector<zc> g((size_t)s->nlso * s->nlso, zc(0.0));
  auto at = [&](int il, int jl, int sp, int io, int jo) -> zc & {
    return g[idx6(il, jl, sp, sp, io, jo, nl, nl, nsp, nsp, no)];
  };
  for (auto &kv : s->gf) {
    zc acc = 0;
    for (auto &chans : kv.second)
      for (auto &c : chans)
        for (size_t j = 0; j < c.w.size(); ++j) acc += c.w[j] / (z - c.p[j]);
    at(kv.first[0] - 1, kv.first[1] - 1, kv.first[2] - 1, kv.first[3] - 1, kv.first[4] - 1) = acc;
  }
  const double chan4 = p.gf_symmetric ? 0.0 : 1.0;
  const zc f(1.0, -chan4);
  for (int sp = 0; sp < nsp; ++sp)
    for (int io = 0; io < no; ++io)
      for (int jo = 0; jo < no; ++jo)
        for (int il = 0; il < nl; ++il)
          for (int jl = 0; jl < nl; ++jl) {
            if (il == jl && io == jo) continue;
            at(il, jl, sp, io, jo) =
                0.5 * (at(il, jl, sp, io, jo) - f * at(il, il, sp, io, io) - f * at(jl, jl, sp, jo, jo));
          }
  std::memcpy(gfo, g.data(), sizeof(zc) * g.size());
  VCA_GUARD_END
}

// save_gfprime / read_gfprime file name: file + ".restart" (".used" when
// used), VCA_AUX_FUNX.f90:620-622 (the reference's file_suffix is empty by
// default and is not an input of this path).
static std::string gfprime_name(const char *file, int32_t used) {
  if (!file || !*file) fail(VCA_ERR_ARG, "empty file name");
  return std::string(file) + (used ? ".used" : ".restart");
}

int vca_save_gfprime(const vca_solver *sv, const char *file, int32_t used,
                     int32_t use_formatted) {
  VCA_GUARD_BEGIN
  const Solver *s = reinterpret_cast<const Solver *>(sv);
  if (!s) fail(VCA_ERR_ARG, "null solver");
  if (!s->solved && !s->gf_loaded)
    fail(VCA_ERR_STATE, "vca_gf_cluster ERROR: impGmatrix not allocated!");
  const vca_params &p = s->p;
  std::vector<GfEntry> el;
  el.reserve((size_t)p.nlat * p.nlat * p.nspin * p.norb * p.norb);
  for (int il = 1; il <= p.nlat; ++il)
    for (int jl = 1; jl <= p.nlat; ++jl)
      for (int sp = 1; sp <= p.nspin; ++sp)
        for (int io = 1; io <= p.norb; ++io)
          for (int jo = 1; jo <= p.norb; ++jo) {
            GfEntry e;
            auto it = s->gf.find({il, jl, sp, io, jo});
            if (it != s->gf.end())
              for (const auto &chans : it->second) {
                e.emplace_back();
                for (const GfChan &c : chans) {
                  GfChanIO io_c;
                  io_c.w = c.w;
                  io_c.p.assign(c.p.begin(), c.p.end());
                  e.back().push_back(std::move(io_c));
                }
              }
            el.push_back(std::move(e));
          }
  gfprime_write(gfprime_name(file, used), use_formatted != 0, el);
  VCA_GUARD_END
}

int vca_read_gfprime(vca_solver *sv, const char *file, int32_t used, int32_t use_formatted) {
  VCA_GUARD_BEGIN
  Solver *s = reinterpret_cast<Solver *>(sv);
  if (!s) fail(VCA_ERR_ARG, "null solver");
  const vca_params &p = s->p;
  const std::string path = gfprime_name(file, used);
  std::vector<GfEntry> el = gfprime_read(
      path, use_formatted != 0, (size_t)p.nlat * p.nlat * p.nspin * p.norb * p.norb);
  std::map<ElemKey, std::vector<std::vector<GfChan>>> gf;
  double max_pole = 0.0;
  size_t k = 0;
  for (int il = 1; il <= p.nlat; ++il)
    for (int jl = 1; jl <= p.nlat; ++jl)
      for (int sp = 1; sp <= p.nspin; ++sp)
        for (int io = 1; io <= p.norb; ++io)
          for (int jo = 1; jo <= p.norb; ++jo) {
            auto &dst = gf[{il, jl, sp, io, jo}];
            for (const auto &st : el[k]) {
              dst.emplace_back();
              for (const GfChanIO &c : st) {
                GfChan g;
                g.w = c.w;
                g.p.resize(c.p.size());
                for (size_t j = 0; j < c.p.size(); ++j) {
                  if (c.p[j].imag() != 0.0)
                    fail(VCA_ERR_ARG, "%s: complex pole (%g,%g); the cluster poles are real",
                         path.c_str(), c.p[j].real(), c.p[j].imag());
                  g.p[j] = c.p[j].real();
                  max_pole = std::max(max_pole, std::fabs(g.p[j]));
                }
                dst.back().push_back(std::move(g));
              }
            }
            ++k;
          }
  s->gf = std::move(gf);
  s->gf_loaded = true;
  s->solved = false;
  s->om_ready = false;
  // The file keeps no excitation energies: the largest |pole| bounds the
  // largest excitation, which sets the finite-T Matsubara cutoff.
  s->max_exc = max_pole;
  VCA_GUARD_END
}

int vca_omega_restart(vca_solver *sv, const double *hloc, const double *hk, int64_t nk,
                      const double *bath_h, const double *bath_v, double omegaprime) {
  VCA_GUARD_BEGIN
  Solver *s = reinterpret_cast<Solver *>(sv);
  if (!s || !hloc || !hk || nk < 1) fail(VCA_ERR_ARG, "null argument or nk < 1");
  if (!s->solved && !s->gf_loaded)
    fail(VCA_ERR_STATE, "no Green's function: vca_solve or vca_read_gfprime first");
  const bool bath = s->p.nlat_bath * s->p.norb_bath > 0;
  if (bath && (!bath_h || !bath_v)) fail(VCA_ERR_ARG, "model has a bath: bath_h/bath_v needed");
  if (bath) {
    int rc = vca_set_bath(sv, bath_h, bath_v);
    if (rc) return rc;
  }
  int rc = vca_set_hcluster(sv, hloc);
  if (rc) return rc;
  const size_t per_k = (size_t)s->nlso * s->nlso;
  s->hk.assign(reinterpret_cast<const zc *>(hk), reinterpret_cast<const zc *>(hk) + per_k * nk);
  s->nk = nk;
  s->solved = false;
  s->om_ready = false;
  const double max_exc = s->max_exc;
  s->res = vca_results{};
  s->res.omegaprime = omegaprime;
  s->max_exc = max_exc;
  const double t0 = now();
  omega_stage(*s);
  s->res.t_omega = now() - t0;
  s->solved = true;
  VCA_GUARD_END
}

int vca_get_gf_grid(vca_solver *sv, int32_t axis, double *gimp, double *sigma, double *g0) {
  VCA_GUARD_BEGIN
  Solver *s = reinterpret_cast<Solver *>(sv);
  if (!s) fail(VCA_ERR_ARG, "null solver");
  if (axis != VCA_AXIS_MATSUBARA && axis != VCA_AXIS_REAL)
    fail(VCA_ERR_ARG, "axis must be VCA_AXIS_MATSUBARA or VCA_AXIS_REAL");
  if (!s->solved) fail(VCA_ERR_STATE, "vca_solve has not completed");
  const vca_params &p = s->p;
  const int L = axis == VCA_AXIS_MATSUBARA ? p.lmats : p.lreal;
  if (axis == VCA_AXIS_MATSUBARA && (L < 1 || !(p.beta > 0)))
    fail(VCA_ERR_ARG, "Matsubara grid needs LMATS >= 1 and BETA > 0");
  if (axis == VCA_AXIS_REAL && L < 2) fail(VCA_ERR_ARG, "real-axis grid needs LREAL >= 2");
  if (!s->om_ready) prepare_omega(*s);
  const float ms =
      s->om->grid_gf(axis, L, p.beta, p.wini, p.wfin, p.eps, reinterpret_cast<double2 *>(gimp),
                     reinterpret_cast<double2 *>(sigma), reinterpret_cast<double2 *>(g0));
  s->res.grid_seconds += ms * 1e-3;
  s->res.grid_pole_evals += (int64_t)s->om->host.w.size() * L;
  VCA_GUARD_END
}

int vca_omega_integrand(vca_solver *sv, int32_t nz, const double *z, int32_t precision,
                        double *f) {
  VCA_GUARD_BEGIN
  Solver *s = reinterpret_cast<Solver *>(sv);
  if (!s || !z || !f || nz < 1) fail(VCA_ERR_ARG, "bad argument");
  if (!s->solved) fail(VCA_ERR_STATE, "vca_solve has not completed");
  if (!s->om_ready) prepare_omega(*s);
  std::vector<double2> zz(nz);
  for (int i = 0; i < nz; ++i) zz[i] = make_double2(z[2 * i], z[2 * i + 1]);
  std::vector<double> out = integrand(*s, zz, precision);
  std::memcpy(f, out.data(), sizeof(double) * nz);
  VCA_GUARD_END
}

int vca_omega_integrand_complex(vca_solver *sv, int32_t nz, const double *z, int32_t precision,
                                double *f) {
  VCA_GUARD_BEGIN
  Solver *s = reinterpret_cast<Solver *>(sv);
  if (!s || !z || !f || nz < 1) fail(VCA_ERR_ARG, "bad argument");
  if (!s->solved) fail(VCA_ERR_STATE, "vca_solve has not completed");
  if (!s->om_ready) prepare_omega(*s);
  std::vector<double2> zz(nz);
  for (int i = 0; i < nz; ++i) zz[i] = make_double2(z[2 * i], z[2 * i + 1]);
  std::vector<double> fim;
  std::vector<double> out = integrand(*s, zz, precision, &fim);
  for (int i = 0; i < nz; ++i) {
    f[2 * i] = out[i];
    f[2 * i + 1] = fim[i];
  }
  VCA_GUARD_END
}

}  // extern "C"
