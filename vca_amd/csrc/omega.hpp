// omega.hpp -- device state of the Omega integrand (see omega.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "vca_internal.hpp"

namespace vca {

// Host-side description of one Omega problem (after the GF chains).
struct OmegaHost {
  int nlat = 0, norb = 0, nspin = 0;
  int nb = 0;         // block size (Nlat*Norb, or Nlat*Nspin*Norb if spin-mixed)
  int nspin_blk = 1;  // number of blocks
  int nbb = 0;        // bath block size, 0 without bath
  int nk = 0;
  double chan4 = 1.0, xmu = 0.0;
  int nelem = 0;      // nspin*nlat*nlat*norb*norb pole-sum elements
  std::vector<int64_t> off;   // [nelem+1] CSR offsets into w/pole
  std::vector<double2> w;     // Lehmann weights
  std::vector<double> pole;   // Lehmann poles
  std::vector<double2> hk;    // [nk][nspin_blk][16][16], zero padded
  std::vector<double2> hloc;  // [nspin_blk][nb][nb]
  std::vector<double2> hb;    // [nspin_blk][nbb][nbb]
  std::vector<double2> vb;    // [nspin_blk][nb][nbb]
};

struct OmegaDev {
  OmegaHost host;
  hipStream_t stream = nullptr;
  int64_t *d_off = nullptr;
  double2 *d_w = nullptr;
  double *d_pole = nullptr;
  double2 *d_hk64 = nullptr;
  float2 *d_hk32 = nullptr;
  double2 *d_hloc = nullptr, *d_hb = nullptr, *d_vb = nullptr;
  double2 *d_z = nullptr, *d_graw = nullptr, *d_g64 = nullptr, *d_p64 = nullptr;
  float2 *d_g32 = nullptr, *d_p32 = nullptr;
  double *d_part = nullptr, *d_f = nullptr;
  double2 *d_phase = nullptr;  // [zcap][nspin_blk][nk] unit phases (complex log mode)
  double *d_fim = nullptr;
  int zcap = 0;
  int nbt = 16;  // rows per matrix in the log-det kernel (8 or 16)

  ~OmegaDev();
  void release();
  void upload(const OmegaHost &h, hipStream_t st);
  void ensure_z(int nz);
  // f[m] for the nz frequencies z (host array); FP64 or MIXED k x z
  // log-dets; optional time of the log-det kernel in ms.  With fim, also
  // the mean principal argument of det (the imaginary part of the complex
  // logarithm of sum_kmesh_complex).
  void integrand(int nz, const double2 *z, int precision, double *f, float *ms_logdet,
                 double *fim = nullptr);
  // G, Sigma and G0 on L points of the Matsubara (axis 0) or real (axis 1)
  // grid, written to the host arrays that are not null, each
  // [Nlat][Nlat][Nspin][Nspin][Norb][Norb][L] with the first index fastest.
  // Returns the GPU time (ms) of the pole-sum kernel.
  float grid_gf(int axis, int L, double beta, double wini, double wfin, double eps, double2 *g,
                double2 *sigma, double2 *g0);
};

}  // namespace vca
