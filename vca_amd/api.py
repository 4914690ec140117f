"""Python mirror of the reference's public interface for the solver core.

Reference names (QcmPlab/VCA):
  vca_init_solver        VCA_MAIN.f90:51-76
  vca_solve              VCA_MAIN.f90:124-215
  vca_get_sft_potential  VCA_IO/get_sft_potential.f90:1-4
  vca_gf_cluster         VCA_IO/gf_cluster.f90:1-64
  operator (cc_sparse_HxV, spHtimesV_p)  VCA_VARS_GLOBAL.f90:83-89,137

Arrays are complex128 with the reference's shapes and Fortran order:
Hloc [Nlat,Nlat,Nspin,Nspin,Norb,Norb], Hk [...,Nk], bath_h
[Nlat_bath,Nlat_bath,Nspin,Nspin,Norb_bath,Norb_bath], bath_v
[Nlat,Nlat_bath,Nspin,Nspin,Norb,Norb_bath].  Errors raise VcaError where
the reference would `stop`.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (VCA_FP64, VCA_MIXED, VCA_LANC_LANCZOS, VCA_LANC_ARPACK,  # noqa: F401
                   VCA_OMEGA_QUAD, VCA_OMEGA_MATSUBARA, VCA_AXIS_MATSUBARA, VCA_AXIS_REAL,
                   VcaError, check)


@dataclass
class VcaParams:
    """Input variables of the path (VCA_INPUT_VARS.f90:112-172); the
    defaults are the reference's."""
    nlat: int = 1
    norb: int = 1
    nspin: int = 1
    nlat_bath: int = 0
    norb_bath: int = 0
    uloc: tuple = (2.0, 0.0, 0.0, 0.0, 0.0)
    ust: float = 0.0
    jh: float = 0.0
    xmu: float = 0.0
    beta: float = 1000.0
    hfmode: bool = True
    hfshift: bool = True
    gf_symmetric: bool = False
    lanc_method: str = "arpack"      # "arpack" | "lanczos"
    lanc_nstates_sector: int = 6
    lanc_niter: int = 512
    lanc_ngfiter: int = 200
    lanc_dim_threshold: int = 256
    lanc_tolerance: float = 1e-12
    gs_threshold: float = 1e-9
    precision: str = "fp64"          # "fp64" | "mixed"
    omega_mode: str = "quad"         # "quad" | "matsubara"
    omega_nfreq: int = 4096
    quad_epsabs: float = 1e-10
    quad_epsrel: float = 1e-12
    sectors: list | None = None      # [(nup, ndw), ...] or None for all
    lanc_nstates_total: int = 1      # > 1: finite temperature
    lanc_nstates_step: int = 2
    lanc_niter_spectrum: int = 10
    lanc_ncv_factor: int = 3
    lanc_ncv_add: int = 5
    cutoff: float = 1e-9
    bandwidth: float = 0.0
    lmats: int = 5000
    lreal: int = 5000
    wini: float = -5.0
    wfin: float = 5.0
    eps: float = 0.01
    _keep: list = field(default_factory=list, repr=False)

    @property
    def ns(self):
        return self.nlat * self.norb + self.nlat_bath * self.norb_bath

    @property
    def nlso(self):
        return self.nlat * self.nspin * self.norb

    def to_c(self) -> _lib.CParams:
        c = _lib.CParams()
        _lib.load().vca_default_params(ctypes.byref(c))
        c.nlat, c.norb, c.nspin = self.nlat, self.norb, self.nspin
        c.nlat_bath, c.norb_bath = self.nlat_bath, self.norb_bath
        c.hfmode, c.hfshift = int(self.hfmode), int(self.hfshift)
        c.gf_symmetric = int(self.gf_symmetric)
        u = (list(self.uloc) + [0.0] * 5)[:5]
        for i in range(5):
            c.uloc[i] = float(u[i])
        c.ust, c.jh, c.xmu, c.beta = self.ust, self.jh, self.xmu, self.beta
        c.lanc_method = {"arpack": VCA_LANC_ARPACK, "lanczos": VCA_LANC_LANCZOS}[self.lanc_method]
        c.lanc_nstates_sector = self.lanc_nstates_sector
        c.lanc_niter, c.lanc_ngfiter = self.lanc_niter, self.lanc_ngfiter
        c.lanc_dim_threshold = self.lanc_dim_threshold
        c.lanc_tolerance, c.gs_threshold = self.lanc_tolerance, self.gs_threshold
        c.precision = {"fp64": VCA_FP64, "mixed": VCA_MIXED}[self.precision]
        c.omega_mode = {"quad": VCA_OMEGA_QUAD, "matsubara": VCA_OMEGA_MATSUBARA}[self.omega_mode]
        c.omega_nfreq = self.omega_nfreq
        c.quad_epsabs, c.quad_epsrel = self.quad_epsabs, self.quad_epsrel
        c.lanc_nstates_total, c.lanc_nstates_step = self.lanc_nstates_total, self.lanc_nstates_step
        c.lanc_niter_spectrum = self.lanc_niter_spectrum
        c.lanc_ncv_factor, c.lanc_ncv_add = self.lanc_ncv_factor, self.lanc_ncv_add
        c.cutoff, c.bandwidth = self.cutoff, self.bandwidth
        c.lmats, c.lreal = self.lmats, self.lreal
        c.wini, c.wfin, c.eps = self.wini, self.wfin, self.eps
        if self.sectors:
            arr = (ctypes.c_int32 * (2 * len(self.sectors)))(
                *[int(x) for s in self.sectors for x in s])
            self._keep = [arr]
            c.nsectors = len(self.sectors)
            c.sectors = ctypes.cast(arr, ctypes.POINTER(ctypes.c_int32))
        return c


def _carr(a, shape):
    a = np.asfortranarray(np.asarray(a, dtype=np.complex128))
    if a.shape != tuple(shape):
        raise VcaError(-1, f"wrong shape {a.shape}, expected {tuple(shape)}")
    return a


class Sector:
    """One (Nup, Ndw) sector with device-resident hop tables: the operator
    the reference binds to spHtimesV_p (build_Hv_sector/delete_Hv_sector,
    VCA_HAMILTONIAN.f90:38-186)."""

    def __init__(self, solver: "VcaSolver", nup: int, ndw: int, precision="fp64"):
        self._lib = _lib.load()
        self.solver = solver
        h = ctypes.c_void_p()
        prec = {"fp64": VCA_FP64, "mixed": VCA_MIXED}[precision]
        check(self._lib.vca_sector_create(solver._h, nup, ndw, prec, ctypes.byref(h)))
        self._h = h
        dim, du, dd = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        check(self._lib.vca_sector_dims(h, ctypes.byref(dim), ctypes.byref(du), ctypes.byref(dd)))
        self.dim, self.dimup, self.dimdw = dim.value, du.value, dd.value

    def close(self):
        if self._h:
            self._lib.vca_sector_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def variant(self):
        """(complex H x v launch shape, real-vector direct shape) the
        autotuner picked; -1 where none."""
        a, b = ctypes.c_int32(), ctypes.c_int32()
        check(self._lib.vca_sector_variant(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def hxv(self, v):
        """Hv = H v (cc_sparse_HxV semantics: Nloc = dim)."""
        v = np.ascontiguousarray(v, dtype=np.complex128)
        if v.shape != (self.dim,):
            raise VcaError(-1, f"vector of length {v.size}, sector dimension {self.dim}")
        hv = np.empty_like(v)
        check(self._lib.vca_hxv(self._h, self.dim, v.ctypes.data, hv.ctypes.data))
        return hv

    def hxv_device(self, v_ptr, hv_ptr, repeat=1):
        ms = ctypes.c_float()
        check(self._lib.vca_hxv_device(self._h, ctypes.c_void_p(v_ptr), ctypes.c_void_p(hv_ptr),
                                       repeat, ctypes.byref(ms)))
        return ms.value

    def hxv_device_batch(self, nb, v_ptr, hv_ptr, repeat=1):
        """nb products at once on device buffers [nb][dim] (sector precision)."""
        ms = ctypes.c_float()
        check(self._lib.vca_hxv_device_batch(self._h, nb, ctypes.c_void_p(v_ptr),
                                             ctypes.c_void_p(hv_ptr), repeat, ctypes.byref(ms)))
        return ms.value

    def hxv_device_real(self, v_ptr, hv_ptr, repeat=1):
        """H v on device buffers of dim reals (double for FP64, float for
        MIXED), for real Hamiltonians on sectors with DimUp >= 128 (VcaError
        otherwise)."""
        ms = ctypes.c_float()
        check(self._lib.vca_hxv_device_real(self._h, ctypes.c_void_p(v_ptr),
                                            ctypes.c_void_p(hv_ptr), repeat, ctypes.byref(ms)))
        return ms.value

    def spin_coo(self, spin):
        nnz = ctypes.c_int64()
        check(self._lib.vca_sector_spin_coo(self._h, spin, ctypes.byref(nnz), None, None, None))
        n = nnz.value
        rows = np.zeros(n, np.int32)
        cols = np.zeros(n, np.int32)
        amps = np.zeros(n, np.complex128)
        check(self._lib.vca_sector_spin_coo(self._h, spin, ctypes.byref(nnz), rows.ctypes.data,
                                            cols.ctypes.data, amps.ctypes.data))
        return rows, cols, amps

    def lanczos_gs(self, tol=1e-12, niter=512, want_vec=True):
        e0 = ctypes.c_double()
        it = ctypes.c_int32()
        vec = np.zeros(self.dim, np.complex128) if want_vec else None
        check(self._lib.vca_lanczos_gs(self._h, tol, niter, ctypes.byref(e0),
                                       vec.ctypes.data if want_vec else None, ctypes.byref(it)))
        return e0.value, vec, it.value

    def lanczos_eigs(self, nev, ncv=None, tol=1e-12, maxiter=512, want_vecs=True):
        """nev lowest eigenpairs (SciFortran sp_eigh / ARPACK semantics,
        VCA_DIAG.f90:153-170); ncv defaults to the reference's Nblock rule
        3*nev + 5 (VCA_DIAG.f90:102).  Returns (evals, evecs[nev][dim] or
        None, nconv)."""
        if ncv is None:
            ncv = 3 * nev + 5
        evals = np.zeros(nev)
        vecs = np.zeros((nev, self.dim), np.complex128) if want_vecs else None
        nconv = ctypes.c_int32()
        check(self._lib.vca_lanczos_eigs(self._h, nev, ncv, tol, maxiter, evals.ctypes.data,
                                         vecs.ctypes.data if want_vecs else None,
                                         ctypes.byref(nconv)))
        return evals, vecs, nconv.value

    def lanczos_tridiag(self, v0, nlanc):
        v0 = np.ascontiguousarray(v0, dtype=np.complex128)
        alpha = np.zeros(nlanc)
        beta = np.zeros(nlanc)
        check(self._lib.vca_lanczos_tridiag(self._h, v0.ctypes.data, nlanc, alpha.ctypes.data,
                                            beta.ctypes.data))
        return alpha, beta


def lanczos_of_sum(a1, b1, n1, a2, b2, n2):
    """The n-step Lanczos tridiagonal of the sum of two chains' spectral
    measures (host only): for a real H and real a, b, the chain of a + i b
    from the chains of a and b (norms n1, n2; beta[0] = 0).  Returns
    (alpha, beta, norm)."""
    a1, b1, a2, b2 = (np.ascontiguousarray(x, dtype=np.float64) for x in (a1, b1, a2, b2))
    n = len(a1)
    if not (len(b1) == len(a2) == len(b2) == n):
        raise ValueError("the two chains need the same length")
    alpha = np.zeros(n)
    beta = np.zeros(n)
    nrm = ctypes.c_double()
    check(_lib.load().vca_lanczos_of_sum(n, a1.ctypes.data, b1.ctypes.data, float(n1),
                                         a2.ctypes.data, b2.ctypes.data, float(n2),
                                         alpha.ctypes.data, beta.ctypes.data,
                                         ctypes.byref(nrm)))
    return alpha, beta, nrm.value


def device_synchronize():
    """Waits for all work the library queued on the device."""
    check(_lib.load().vca_device_synchronize())


class DeviceArray:
    """A device buffer of the library (vca_device_alloc), with numpy
    upload/download; `ptr` is what the device-resident entry points take."""

    def __init__(self, shape, dtype):
        self._lib = _lib.load()
        self.shape = tuple(np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(self._lib.vca_device_alloc(self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(self._lib.vca_device_copy(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p),
                                        self.nbytes, 0))
        return self

    def download(self):
        out = np.empty(self.shape, self.dtype)
        check(self._lib.vca_device_copy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                                        self.nbytes, 1))
        return out

    def free(self):
        if self.ptr:
            self._lib.vca_device_free(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class VcaSolver:
    """Handle returned by vca_init_solver."""

    def __init__(self, params: VcaParams):
        self._lib = _lib.load()
        self.params = params
        self._cp = params.to_c()
        h = ctypes.c_void_p()
        check(self._lib.vca_init_solver(ctypes.byref(self._cp), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vca_finalize_solver(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- shapes ---------------------------------------------------------
    def _shape_hloc(self):
        p = self.params
        return (p.nlat, p.nlat, p.nspin, p.nspin, p.norb, p.norb)

    def _shape_bath(self):
        p = self.params
        return ((p.nlat_bath, p.nlat_bath, p.nspin, p.nspin, p.norb_bath, p.norb_bath),
                (p.nlat, p.nlat_bath, p.nspin, p.nspin, p.norb, p.norb_bath))

    # --- input ----------------------------------------------------------
    def vca_set_hcluster(self, hloc):
        a = _carr(hloc, self._shape_hloc())
        check(self._lib.vca_set_hcluster(self._h, a.ctypes.data))

    def vca_set_bath(self, bath_h, bath_v):
        sh, sv = self._shape_bath()
        bh, bv = _carr(bath_h, sh), _carr(bath_v, sv)
        check(self._lib.vca_set_bath(self._h, bh.ctypes.data, bv.ctypes.data))

    def sector(self, nup, ndw, precision="fp64") -> Sector:
        return Sector(self, nup, ndw, precision)

    # --- solve ----------------------------------------------------------
    def vca_solve(self, hloc, hk, bath_h=None, bath_v=None):
        a = _carr(hloc, self._shape_hloc())
        hk = np.asfortranarray(np.asarray(hk, dtype=np.complex128))
        if hk.ndim != 7 or hk.shape[:6] != self._shape_hloc():
            raise VcaError(-1, "STOP: wrong lattice matrix dimensions")
        bh = bv = None
        if bath_h is not None and bath_v is not None:
            sh, sv = self._shape_bath()
            bh, bv = _carr(bath_h, sh), _carr(bath_v, sv)
        check(self._lib.vca_solve(self._h, a.ctypes.data, hk.ctypes.data, hk.shape[6],
                                  bh.ctypes.data if bh is not None else None,
                                  bv.ctypes.data if bv is not None else None))
        return self.vca_get_sft_potential()

    def vca_get_sft_potential(self) -> float:
        out = ctypes.c_double()
        check(self._lib.vca_get_sft_potential(self._h, ctypes.byref(out)))
        return out.value

    def results(self) -> dict:
        r = _lib.CResults()
        check(self._lib.vca_get_results(self._h, ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in _lib.CResults._fields_}

    # --- Green's function and Omega integrand ---------------------------
    def gf_poles(self, ilat, jlat, ispin, iorb, jorb, istate, ichan):
        n = ctypes.c_int32()
        args = (self._h, ilat, jlat, ispin, iorb, jorb, istate, ichan, ctypes.byref(n))
        check(self._lib.vca_gf_poles(*args, None, None))
        w = np.zeros(n.value, np.complex128)
        p = np.zeros(n.value)
        check(self._lib.vca_gf_poles(*args, w.ctypes.data, p.ctypes.data))
        return w, p

    def vca_gf_cluster(self, z):
        g = np.zeros(self._shape_hloc(), dtype=np.complex128, order="F")
        check(self._lib.vca_gf_cluster(self._h, float(np.real(z)), float(np.imag(z)),
                                       g.ctypes.data))
        return g

    def omega_integrand(self, z, precision="fp64"):
        z = np.ascontiguousarray(np.atleast_1d(z), dtype=np.complex128)
        f = np.zeros(z.size)
        prec = {"fp64": VCA_FP64, "mixed": VCA_MIXED}[precision]
        check(self._lib.vca_omega_integrand(self._h, z.size, z.ctypes.data, prec, f.ctypes.data))
        return f

    def omega_integrand_complex(self, z, precision="fp64"):
        """sum_kmesh_complex (VCA_OMEGA.f90:78-116): mean over k of the
        principal complex log of each determinant."""
        z = np.ascontiguousarray(np.atleast_1d(z), dtype=np.complex128)
        f = np.zeros(z.size, np.complex128)
        prec = {"fp64": VCA_FP64, "mixed": VCA_MIXED}[precision]
        check(self._lib.vca_omega_integrand_complex(self._h, z.size, z.ctypes.data, prec,
                                                    f.ctypes.data))
        return f

    # --- pole file (checkpoint) -----------------------------------------
    def save_gfprime(self, file="gfprime", used=False, use_formatted=False):
        """save_gfprime, VCA_AUX_FUNX.f90:608-652: writes file.restart
        (file.used when used)."""
        check(self._lib.vca_save_gfprime(self._h, str(file).encode(), int(bool(used)),
                                         int(bool(use_formatted))))

    def read_gfprime(self, file="gfprime", used=False, use_formatted=False):
        """read_gfprime, VCA_AUX_FUNX.f90:654-697: replaces the pole store."""
        check(self._lib.vca_read_gfprime(self._h, str(file).encode(), int(bool(used)),
                                         int(bool(use_formatted))))

    def omega_restart(self, hloc, hk, omegaprime, bath_h=None, bath_v=None):
        """Omega = omegaprime - integral from the current pole store, without
        diagonalising (the Omega stage of vca_solve, VCA_MAIN.f90:181-197)."""
        a = _carr(hloc, self._shape_hloc())
        hk = np.asfortranarray(np.asarray(hk, dtype=np.complex128))
        if hk.ndim != 7 or hk.shape[:6] != self._shape_hloc():
            raise VcaError(-1, "STOP: wrong lattice matrix dimensions")
        bh = bv = None
        if bath_h is not None and bath_v is not None:
            sh, sv = self._shape_bath()
            bh, bv = _carr(bath_h, sh), _carr(bath_v, sv)
        check(self._lib.vca_omega_restart(self._h, a.ctypes.data, hk.ctypes.data, hk.shape[6],
                                          bh.ctypes.data if bh is not None else None,
                                          bv.ctypes.data if bv is not None else None,
                                          float(omegaprime)))
        return self.vca_get_sft_potential()

    # --- G, Sigma, G0 on the frequency grids -----------------------------
    def frequency_grid(self, axis):
        """wm (axis "matsubara") or wr (axis "realaxis") of allocate_grids,
        VCA_GF_SHARED.f90:43-56."""
        p = self.params
        if axis == "matsubara":
            return (np.pi / p.beta) * (2 * np.arange(1, p.lmats + 1) - 1).astype(np.float64)
        step = (p.wfin - p.wini) / float(p.lreal - 1)
        return p.wini + np.arange(p.lreal, dtype=np.float64) * step

    def gf_grid(self, axis, which=("g", "sigma", "g0")):
        """The arrays named in `which` ("g", "sigma", "g0") on the Matsubara
        or real-axis grid, each [Nlat,Nlat,Nspin,Nspin,Norb,Norb,L]."""
        p = self.params
        ax = {"matsubara": VCA_AXIS_MATSUBARA, "realaxis": VCA_AXIS_REAL}[axis]
        L = p.lmats if ax == VCA_AXIS_MATSUBARA else p.lreal
        out = {k: np.zeros(self._shape_hloc() + (L,), np.complex128, order="F") for k in which}
        ptr = [out[k].ctypes.data if k in out else None for k in ("g", "sigma", "g0")]
        check(self._lib.vca_get_gf_grid(self._h, ax, *ptr))
        return out

    def _grid_get(self, axis, which, ilat, jlat):
        a = self.gf_grid(axis, (which,))[which]
        if ilat is None:
            return a
        return np.asfortranarray(a[ilat - 1, jlat - 1])

    def vca_get_gimp_matsubara(self, ilat=None, jlat=None):
        """VCA_IO/get_gimp_matsubara.f90 (full, or the (ilat, jlat) slice)."""
        return self._grid_get("matsubara", "g", ilat, jlat)

    def vca_get_gimp_realaxis(self, ilat=None, jlat=None):
        """VCA_IO/get_gimp_realaxis.f90."""
        return self._grid_get("realaxis", "g", ilat, jlat)

    def vca_get_sigma_matsubara(self, ilat=None, jlat=None):
        """VCA_IO/get_sigma_matsubara.f90."""
        return self._grid_get("matsubara", "sigma", ilat, jlat)

    def vca_get_sigma_realaxis(self, ilat=None, jlat=None):
        """VCA_IO/get_sigma_realaxis.f90."""
        return self._grid_get("realaxis", "sigma", ilat, jlat)


def vca_init_solver(params: VcaParams, bath_h=None, bath_v=None) -> VcaSolver:
    """VCA_MAIN.f90:51-76 (the optional bath arrays only fix shapes there)."""
    s = VcaSolver(params)
    if bath_h is not None and bath_v is not None:
        s._shape_bath()
    return s
