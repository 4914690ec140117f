/*
 * vca_gpu.h -- C ABI of the MI355X-native VCA solver core (libvca_gpu.so).
 *
 * Plain pointers and sizes only.  Complex numbers are complex128 stored as
 * interleaved (re, im) doubles, arrays with several indices are Fortran
 * column-major exactly as the reference passes them (so a Fortran caller
 * passes c_loc(array) of its own arrays).  Every entry point returns a
 * status code (VCA_OK or a negative error) and never aborts the process;
 * vca_last_error() gives the message of the last failure on the calling
 * thread.  Handles are independent: no hidden global solver state.
 *
 * Reference interfaces replaced (QcmPlab/VCA):
 *   vca_init_solver      VCA_MAIN.f90:51-76     (vca_init_solver_serial)
 *   vca_solve            VCA_MAIN.f90:124-215   (vca_solve_serial)
 *   vca_get_sft_potential VCA_IO/get_sft_potential.f90:1-4
 *   vca_set_hcluster     VCA_AUX_FUNX.f90:124-131 (set_Hcluster_nnn)
 *   vca_set_bath         VCA_BATH_SETUP.f90:112-138
 *   vca_sector_create /
 *   vca_sector_destroy   VCA_HAMILTONIAN.f90:38-186 (build_Hv_sector /
 *                        delete_Hv_sector)
 *   vca_hxv              VCA_VARS_GLOBAL.f90:83-89 (cc_sparse_HxV operator,
 *                        bound to directMatVec_main,
 *                        VCA_HAMILTONIAN_DIRECT_HxV.f90:42-78); the device
 *                        forms vca_hxv_device / vca_hxv_device_batch too
 *   vca_lanczos_gs       SciFortran sp_lanc_eigh as called at VCA_DIAG.f90:172-184
 *   vca_lanczos_eigs     SciFortran sp_eigh (ARPACK, which='SA') as called at
 *                        VCA_DIAG.f90:153-170
 *   vca_lanczos_tridiag  SciFortran sp_lanc_tridiag as called at
 *                        VCA_GF_NORMAL.f90:218-223
 *   vca_gf_cluster       VCA_IO/gf_cluster.f90:1-64 (vca_gf_cluster_scalar)
 *   vca_gf_poles         VCA_VARS_GLOBAL.f90:61-71 (GFmatrix weights/poles)
 *   vca_omega_integrand  VCA_OMEGA.f90:37-75     (sum_kmesh, batched in omega)
 *   vca_omega_integrand_complex VCA_OMEGA.f90:78-116 (sum_kmesh_complex)
 *   vca_get_gf_grid      VCA_IO/get_gimp_matsubara.f90, get_gimp_realaxis.f90,
 *                        get_sigma_matsubara.f90, get_sigma_realaxis.f90
 *                        (impGmats/impGreal of VCA_GF_NORMAL.f90:89-101,
 *                        1024-1031; impSmats/impSreal/impG0* of
 *                        build_sigma_normal, VCA_GF_NORMAL.f90:1045-1111)
 *   vca_save_gfprime     VCA_AUX_FUNX.f90:608-652 (save_gfprime) with the
 *                        GFmatrix defined I/O at :484-604
 *   vca_read_gfprime     VCA_AUX_FUNX.f90:654-697 (read_gfprime)
 *   vca_omega_restart    VCA_MAIN.f90:181-197 (the Omega stage of vca_solve)
 *                        on a pole store read back, the restart that
 *                        VCA_MAIN.f90:280-281 sketches
 */
#ifndef VCA_GPU_H
#define VCA_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define VCA_OK 0
#define VCA_ERR_ARG (-1)    /* bad argument or shape            */
#define VCA_ERR_HIP (-2)    /* HIP runtime error                */
#define VCA_ERR_NOMEM (-3)  /* device or host allocation failed */
#define VCA_ERR_STATE (-4)  /* call out of order                */
#define VCA_ERR_NOCONV (-5) /* iteration did not converge       */

/* precision of the H x v and Lanczos vectors */
#define VCA_FP64 0  /* complex128 everywhere (drop-in)                      */
#define VCA_MIXED 1 /* complex64 vectors and H x v, FP64 alpha/beta, FP64
                       refinement of the ground state                      */

/* ground-state method, reference LANC_METHOD (VCA_DIAG.f90:98-107) */
#define VCA_LANC_LANCZOS 0 /* one state per sector                        */
#define VCA_LANC_ARPACK 1  /* up to lanc_nstates_sector states per sector:
                              dense eigh up to lanc_dim_threshold, GPU
                              thick-restart Lanczos (FP64) above it        */

/* frequency integration of Omega */
#define VCA_OMEGA_QUAD 0      /* adaptive Gauss-Kronrod on [0,inf), T=0     */
#define VCA_OMEGA_MATSUBARA 1 /* (2/beta) sum over omega_nfreq Matsubara
                                 frequencies                               */

/* frequency axis of the grid outputs (VCA_GF_SHARED.f90:43-56) */
#define VCA_AXIS_MATSUBARA 0 /* z_n = i (2n - 1) pi / beta, n = 1..lmats     */
#define VCA_AXIS_REAL 1      /* z_n = wini + (n-1)(wfin - wini)/(lreal - 1)
                                + i eps, n = 1..lreal                      */

/* Input variables of the path (VCA_INPUT_VARS.f90:112-172 names in
 * comments).  vca_default_params fills the reference defaults. */
typedef struct vca_params {
  int32_t nlat;          /* NLAT      */
  int32_t norb;          /* NORB      */
  int32_t nspin;         /* NSPIN     */
  int32_t nlat_bath;     /* NLAT_BATH */
  int32_t norb_bath;     /* NORB_BATH */
  int32_t hfmode;        /* HFMODE    */
  int32_t hfshift;       /* HFSHIFT   */
  int32_t gf_symmetric;  /* VCA_GF_SYMMETRIC */
  double uloc[5];        /* ULOC      */
  double ust;            /* UST       */
  double jh;             /* JH        */
  double xmu;            /* XMU       */
  double beta;           /* BETA      */
  int32_t lanc_method;   /* LANC_METHOD: VCA_LANC_*            */
  int32_t lanc_nstates_sector; /* LANC_NSTATES_SECTOR          */
  int32_t lanc_niter;    /* LANC_NITER                         */
  int32_t lanc_ngfiter;  /* LANC_NGFITER                       */
  int32_t lanc_dim_threshold; /* LANC_DIM_THRESHOLD            */
  double lanc_tolerance; /* LANC_TOLERANCE                     */
  double gs_threshold;   /* GS_THRESHOLD                       */
  int32_t precision;     /* VCA_FP64 / VCA_MIXED               */
  int32_t omega_mode;    /* VCA_OMEGA_*                        */
  int32_t omega_nfreq;   /* Matsubara frequencies (grid mode)  */
  double quad_epsabs;    /* quadrature tolerances (quad mode)  */
  double quad_epsrel;
  int32_t nsectors;      /* 0: scan every sector; else the list below, the
                            equivalent of VCA_SECTORS=T with a
                            sectors_list.restart file (VCA_DIAG.f90:287-319) */
  const int32_t *sectors; /* nsectors pairs (nup, ndw)                     */
  /* finite temperature (lanc_nstates_total > 1) and the ARPACK basis */
  int32_t lanc_nstates_total;  /* LANC_NSTATES_TOTAL (1: T = 0)           */
  int32_t lanc_nstates_step;   /* LANC_NSTATES_STEP                       */
  int32_t lanc_niter_spectrum; /* LANC_NITER_SPECTRUM                     */
  int32_t lanc_ncv_factor;     /* LANC_NCV_FACTOR                         */
  int32_t lanc_ncv_add;        /* LANC_NCV_ADD                            */
  double cutoff;               /* CUTOFF                                  */
  double bandwidth;            /* bandwidth (set by the caller, as the
                                  reference drivers do)                   */
  /* frequency grids of the G / Sigma outputs (VCA_GF_SHARED.f90:43-56) */
  int32_t lmats;               /* LMATS: w_n = (2n-1) pi / beta, n=1..   */
  int32_t lreal;               /* LREAL: linspace(WINI, WFIN, LREAL)     */
  double wini, wfin;           /* WINI, WFIN                             */
  double eps;                  /* EPS: real-axis broadening              */
} vca_params;

/* Per-solve results and timings (seconds) in addition to the SFT potential. */
typedef struct vca_results {
  double sft_potential;  /* Omega = omegaprime - omega_integral           */
  double omegaprime;     /* E0 at T=0                                     */
  double omega_integral;
  double e0_mixed;       /* MIXED: E0 before FP64 refinement              */
  int32_t nstates;       /* ground states kept (zeta function)            */
  int32_t nchains;       /* excitation chains run                         */
  int64_t hxv_count;     /* Lanczos steps (H x v applications), all chains */
  double hxv_states;     /* sum over those steps of the sector dimension  */
  double hxv_seconds;    /* GPU time of those steps (H x v + vector update
                            kernels, events)                              */
  double t_diag, t_gf, t_omega;
  int64_t omega_evals;   /* (k, omega) log-det evaluations                */
  double omega_seconds;  /* GPU time in the k x omega kernels             */
  double zeta;           /* zeta_function (VCA_DIAG.f90:355-364)          */
  double max_exc;        /* largest excitation energy of the GF chains    */
  int32_t finite_t;      /* 1 when lanc_nstates_total > 1                 */
  int32_t nmax;          /* finite T: Matsubara terms summed (0..nmax)    */
  int32_t spectrum_iters;/* finite T: spectrum annealing iterations       */
  int32_t nchains_derived; /* c_i - i c_j channels of a real H and real
                              state taken from their two diagonal chains,
                              not run (counted apart from nchains)      */
  int64_t npoles;        /* Lehmann poles stored, all elements            */
  int64_t grid_pole_evals; /* pole x frequency terms of vca_get_gf_grid   */
  double grid_seconds;   /* GPU time of those pole sums (events)          */
} vca_results;

typedef struct vca_solver vca_solver;
typedef struct vca_sector vca_sector;

void vca_default_params(vca_params *p);
const char *vca_last_error(void);

/* VCA_MAIN.f90:51-76.  Fixes the dimensions; binds device 0. */
int vca_init_solver(const vca_params *p, vca_solver **out);
void vca_finalize_solver(vca_solver *s);

/* VCA_AUX_FUNX.f90:124-131: Hloc complex128 [nlat,nlat,nspin,nspin,norb,norb] */
int vca_set_hcluster(vca_solver *s, const double *hloc);
/* VCA_BATH_SETUP.f90:112-138: bath_h [nlat_b,nlat_b,nspin,nspin,norb_b,norb_b],
 * bath_v [nlat,nlat_b,nspin,nspin,norb,norb_b]; applies the hfmode shift. */
int vca_set_bath(vca_solver *s, const double *bath_h, const double *bath_v);

/* VCA_MAIN.f90:124-215.  hk complex128 [nlat,nlat,nspin,nspin,norb,norb,nk];
 * bath_h/bath_v may be NULL when the model has no bath. */
int vca_solve(vca_solver *s, const double *hloc, const double *hk, int64_t nk,
              const double *bath_h, const double *bath_v);
/* VCA_IO/get_sft_potential.f90:1-4 */
int vca_get_sft_potential(const vca_solver *s, double *sft_potential);
int vca_get_results(const vca_solver *s, vca_results *out);

/* ---- operator level (the cc_sparse_HxV plugin point) ---------------- */
/* Sector (nup, ndw) of the current Hloc/bath; device tables built on GPU. */
int vca_sector_create(vca_solver *s, int32_t nup, int32_t ndw,
                      int32_t precision, vca_sector **out);
void vca_sector_destroy(vca_sector *sec);
int vca_sector_dims(const vca_sector *sec, int64_t *dim, int32_t *dimup,
                    int32_t *dimdw);
/* The launch shapes the autotuner picked for the sector's complex H x v
 * and for its real-vector direct form (-1: not tuned / not used).  No
 * counterpart in the reference; it names the kernels a measurement of
 * this sector ran, so stored profiler counters can be matched to them. */
int vca_sector_variant(const vca_sector *sec, int32_t *hxv_variant,
                       int32_t *real_variant);
/* Hv = H v, host complex128 arrays of length n = dim (state index
 * i = iup + idw*DimUp, 0-based).  In MIXED precision the product runs on
 * complex64 and is widened back. */
int vca_hxv(vca_sector *sec, int64_t n, const double *v, double *hv);
/* Same on device buffers already resident in HBM (complex128 for FP64,
 * complex64 for MIXED sectors), enqueued on the sector's stream;
 * repeat > 1 applies the product repeat times (for timing).  Returns the
 * average kernel time in *ms when ms is not NULL. */
int vca_hxv_device(vca_sector *sec, const void *v, void *hv, int32_t repeat,
                   float *ms);
/* Device buffers for the device-resident forms (for callers without a HIP
 * runtime of their own, e.g. Fortran).  Copies are synchronous. */
#define VCA_COPY_H2D 0
#define VCA_COPY_D2H 1
#define VCA_COPY_D2D 2
int vca_device_alloc(int64_t bytes, void **ptr);
int vca_device_free(void *ptr);
int vca_device_copy(void *dst, const void *src, int64_t bytes, int32_t kind);
/* Waits for all work the library queued on the device. */
int vca_device_synchronize(void);
/* No counterpart in the reference: a form of vca_hxv_device for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  It allocates nothing: the transposed
 * form's scratch vector (v^T) lives in a device array of 256 MiB that the
 * library holds on each GPU it runs on until the process ends, and the
 * direct forms need none.  A sector whose product needs more scratch than
 * that (complex128 sectors above 16,777,216 states on the transposed form)
 * is refused with VCA_ERR_ARG.  The product is enqueued on `stream` (a
 * hipStream_t; NULL: the sector's stream) and not waited for.  Every call on
 * one GPU shares the scratch: calls that overlap on one GPU race.  The
 * sector must have been created (and its launch shape tuned) beforehand. */
int vca_hxv_device_noalloc(vca_sector *sec, const void *v, void *hv, void *stream);
/* The same for nb vectors stacked [nb][dim] in v and hv: one launch per
 * application covers all nb products (the batched form the excitation
 * chains use); *ms is the average time of one application to all nb. */
int vca_hxv_device_batch(vca_sector *sec, int32_t nb, const void *v, void *hv,
                         int32_t repeat, float *ms);
/* Real form for real Hamiltonians (every hop amplitude real): v and hv are
 * dim reals on the device (double for FP64, float for MIXED sectors).  It
 * is the product the excitation chains with real coefficients, and the
 * ground state of large sectors, run on for such Hamiltonians.  Needs
 * DimUp >= 128 and the staged column to fit the LDS (VCA_ERR_ARG
 * otherwise).  No counterpart in the reference, whose vectors are complex
 * (cc_sparse_HxV, VCA_VARS_GLOBAL.f90:83-89); a real v gives the real part
 * of what vca_hxv_device gives for v + 0i. */
int vca_hxv_device_real(vca_sector *sec, const void *v, void *hv, int32_t repeat,
                        float *ms);
/* One-spin hop table in coordinate form (rows, cols: 0-based string
 * indices; amps complex128), spin 1 = up, 2 = down.  With rows == NULL only
 * the count is returned in *nnz.  Connectivity/sign check of the device
 * tables. */
int vca_sector_spin_coo(const vca_sector *sec, int32_t spin, int64_t *nnz,
                        int32_t *rows, int32_t *cols, double *amps);
/* Plain Lanczos ground state of the sector, FP64 result; vec (host
 * complex128, length dim) may be NULL. */
int vca_lanczos_gs(vca_sector *sec, double tol, int32_t niter, double *e0,
                   double *vec, int32_t *iters);
/* nev lowest eigenpairs (ARPACK semantics: ncv basis vectors, maxiter
 * restarts, a Ritz pair is converged when |beta| |y_ncv| <= tol *
 * max(eps^(2/3), |theta|)) by thick-restart Lanczos with full
 * reorthogonalisation.  FP64 (a MIXED sector is solved in FP64).
 * evals[nev] ascending; evecs (host complex128 [nev][dim]) may be NULL;
 * *nconv = number of converged pairs (nconv < nev is not an error here). */
int vca_lanczos_eigs(vca_sector *sec, int32_t nev, int32_t ncv, double tol,
                     int32_t maxiter, double *evals, double *evecs,
                     int32_t *nconv);
/* Lanczos tridiagonalisation from host start vector v0 (need not be
 * normalised), nlanc steps: alpha[nlanc], beta[nlanc] (beta[0] = 0). */
int vca_lanczos_tridiag(vca_sector *sec, const double *v0, int32_t nlanc,
                        double *alpha, double *beta);
/* Host only (no GPU call).  The nlanc-step tridiagonal of the sum of two
 * chains' spectral measures: chain i has start norm ni and coefficients
 * ai[nlanc], bi[nlanc] (bi[0] = 0) as vca_lanczos_tridiag returns them.
 * For a real H and real a, b this is the chain of the start vector a + i b
 * (or a - i b), which is how vca_solve takes the (c^+_i + i c^+_j) and
 * (c_i - i c_j) channels (VCA_GF_NORMAL.f90:121-318, 561-958) of a real
 * Hamiltonian from the c^+_i, c^+_j (c_i, c_j) chains instead of running
 * them.  No counterpart in the reference, which runs every chain.
 * alpha[nlanc], beta[nlanc] (beta[0] = 0); *nrm = sqrt(n1^2 + n2^2). */
int vca_lanczos_of_sum(int32_t nlanc, const double *a1, const double *b1, double n1,
                       const double *a2, const double *b2, double n2, double *alpha,
                       double *beta, double *nrm);

/* ---- after vca_solve: Green's function and Omega integrand ---------- */
/* Lehmann data of element (ilat,jlat,ispin,iorb,jorb) (1-based), state and
 * channel (1-based): *n poles; weights complex128[n], poles double[n]. */
int vca_gf_poles(const vca_solver *s, int32_t ilat, int32_t jlat,
                 int32_t ispin, int32_t iorb, int32_t jorb, int32_t istate,
                 int32_t ichan, int32_t *n, double *weights, double *poles);
/* G'(z) -> gf complex128 [nlat,nlat,nspin,nspin,norb,norb] */
int vca_gf_cluster(const vca_solver *s, double zre, double zim, double *gf);
/* f(z_m) = (1/Nk) sum_k log|det(1 + (Delta + Hloc - H_k) G'(z_m))| for nz
 * complex128 frequencies z (imaginary axis or any z off the real axis). */
int vca_omega_integrand(vca_solver *s, int32_t nz, const double *z,
                        int32_t precision, double *f);
/* VCA_OMEGA.f90:78-116 (sum_kmesh_complex): the complex mean
 * (1/Nk) sum_k log(det(...)) with the principal logarithm of each
 * determinant; f complex128[nz]. */
int vca_omega_integrand_complex(vca_solver *s, int32_t nz, const double *z,
                                int32_t precision, double *f);
/* The cluster Green's function G, the self-energy Sigma = G0^-1 - G^-1 and
 * G0 = ((z + xmu) - Hloc - Delta(z))^-1 on the Matsubara or real-axis grid
 * (lmats or lreal points of vca_params), each complex128
 * [nlat,nlat,nspin,nspin,norb,norb,L] with the first index fastest.  Any of
 * gimp, sigma, g0 may be NULL.  The pole sums and the per-frequency
 * inversions run on the GPU. */
int vca_get_gf_grid(vca_solver *s, int32_t axis, double *gimp, double *sigma,
                    double *g0);
/* Writes the pole store to file + ".restart" (".used" when used != 0) in
 * the reference's layout: one entry per (ilat, jlat, ispin, iorb, jorb),
 * jorb fastest, each Nstates, per state Nchan, per channel Nexc, poles,
 * weights (complex128).  use_formatted = 0: Fortran sequential unformatted,
 * one record per entry with 4-byte markers; 1: list-directed text. */
int vca_save_gfprime(const vca_solver *s, const char *file, int32_t used,
                     int32_t use_formatted);
/* Reads such a file (also those the reference writes) into the pole store:
 * vca_gf_cluster, vca_gf_poles and vca_save_gfprime then use it; the Omega
 * integrands and the grid outputs need vca_omega_restart (or vca_solve). */
int vca_read_gfprime(vca_solver *s, const char *file, int32_t used,
                     int32_t use_formatted);
/* Omega = omegaprime - integral from the current pole store without
 * diagonalising: sets Hloc, H(k) and the bath as vca_solve does and runs its
 * frequency integration; results as after vca_solve.  After a file read the
 * finite-T Matsubara cutoff takes the largest |pole| as the largest
 * excitation (the file keeps no excitation energies). */
int vca_omega_restart(vca_solver *s, const double *hloc, const double *hk, int64_t nk,
                      const double *bath_h, const double *bath_v, double omegaprime);

#ifdef __cplusplus
}
#endif
#endif /* VCA_GPU_H */
