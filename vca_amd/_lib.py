"""ctypes binding of libvca_gpu.so (the C ABI declared in include/vca_gpu.h).

The library is the product: there is no Python or CPU fallback.  Loading
fails loudly when the shared object has not been built.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvca_gpu.so")

VCA_OK = 0
VCA_ERR_ARG = -1
VCA_ERR_HIP = -2
VCA_ERR_NOMEM = -3
VCA_ERR_STATE = -4
VCA_ERR_NOCONV = -5
VCA_FP64 = 0
VCA_MIXED = 1
VCA_LANC_LANCZOS = 0
VCA_LANC_ARPACK = 1
VCA_OMEGA_QUAD = 0
VCA_OMEGA_MATSUBARA = 1
VCA_AXIS_MATSUBARA = 0
VCA_AXIS_REAL = 1


class CParams(ctypes.Structure):
    _fields_ = [
        ("nlat", ctypes.c_int32), ("norb", ctypes.c_int32), ("nspin", ctypes.c_int32),
        ("nlat_bath", ctypes.c_int32), ("norb_bath", ctypes.c_int32),
        ("hfmode", ctypes.c_int32), ("hfshift", ctypes.c_int32),
        ("gf_symmetric", ctypes.c_int32),
        ("uloc", ctypes.c_double * 5), ("ust", ctypes.c_double), ("jh", ctypes.c_double),
        ("xmu", ctypes.c_double), ("beta", ctypes.c_double),
        ("lanc_method", ctypes.c_int32), ("lanc_nstates_sector", ctypes.c_int32),
        ("lanc_niter", ctypes.c_int32), ("lanc_ngfiter", ctypes.c_int32),
        ("lanc_dim_threshold", ctypes.c_int32),
        ("lanc_tolerance", ctypes.c_double), ("gs_threshold", ctypes.c_double),
        ("precision", ctypes.c_int32), ("omega_mode", ctypes.c_int32),
        ("omega_nfreq", ctypes.c_int32),
        ("quad_epsabs", ctypes.c_double), ("quad_epsrel", ctypes.c_double),
        ("nsectors", ctypes.c_int32), ("sectors", ctypes.POINTER(ctypes.c_int32)),
        ("lanc_nstates_total", ctypes.c_int32), ("lanc_nstates_step", ctypes.c_int32),
        ("lanc_niter_spectrum", ctypes.c_int32), ("lanc_ncv_factor", ctypes.c_int32),
        ("lanc_ncv_add", ctypes.c_int32),
        ("cutoff", ctypes.c_double), ("bandwidth", ctypes.c_double),
        ("lmats", ctypes.c_int32), ("lreal", ctypes.c_int32),
        ("wini", ctypes.c_double), ("wfin", ctypes.c_double), ("eps", ctypes.c_double),
    ]


class CResults(ctypes.Structure):
    _fields_ = [
        ("sft_potential", ctypes.c_double), ("omegaprime", ctypes.c_double),
        ("omega_integral", ctypes.c_double), ("e0_mixed", ctypes.c_double),
        ("nstates", ctypes.c_int32), ("nchains", ctypes.c_int32),
        ("hxv_count", ctypes.c_int64), ("hxv_states", ctypes.c_double),
        ("hxv_seconds", ctypes.c_double),
        ("t_diag", ctypes.c_double), ("t_gf", ctypes.c_double), ("t_omega", ctypes.c_double),
        ("omega_evals", ctypes.c_int64), ("omega_seconds", ctypes.c_double),
        ("zeta", ctypes.c_double), ("max_exc", ctypes.c_double),
        ("finite_t", ctypes.c_int32), ("nmax", ctypes.c_int32),
        ("spectrum_iters", ctypes.c_int32), ("nchains_derived", ctypes.c_int32),
        ("npoles", ctypes.c_int64), ("grid_pole_evals", ctypes.c_int64),
        ("grid_seconds", ctypes.c_double),
    ]


# name -> (restype, argtypes)
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.c_int32
_SIGS = {
    "vca_default_params": (None, [ctypes.POINTER(CParams)]),
    "vca_last_error": (ctypes.c_char_p, []),
    "vca_init_solver": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(_P)]),
    "vca_finalize_solver": (None, [_P]),
    "vca_set_hcluster": (ctypes.c_int, [_P, _P]),
    "vca_set_bath": (ctypes.c_int, [_P, _P, _P]),
    "vca_solve": (ctypes.c_int, [_P, _P, _P, ctypes.c_int64, _P, _P]),
    "vca_get_sft_potential": (ctypes.c_int, [_P, _D]),
    "vca_get_results": (ctypes.c_int, [_P, ctypes.POINTER(CResults)]),
    "vca_sector_create": (ctypes.c_int, [_P, _I32, _I32, _I32, ctypes.POINTER(_P)]),
    "vca_sector_destroy": (None, [_P]),
    "vca_sector_variant": (ctypes.c_int, [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "vca_sector_dims": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "vca_hxv": (ctypes.c_int, [_P, ctypes.c_int64, _P, _P]),
    "vca_hxv_device": (ctypes.c_int, [_P, _P, _P, _I32, ctypes.POINTER(ctypes.c_float)]),
    "vca_hxv_device_noalloc": (ctypes.c_int, [_P, _P, _P, _P]),
    "vca_hxv_device_batch": (ctypes.c_int, [_P, _I32, _P, _P, _I32,
                                            ctypes.POINTER(ctypes.c_float)]),
    "vca_hxv_device_real": (ctypes.c_int, [_P, _P, _P, _I32, ctypes.POINTER(ctypes.c_float)]),
    "vca_device_alloc": (ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(_P)]),
    "vca_device_free": (ctypes.c_int, [_P]),
    "vca_device_synchronize": (ctypes.c_int, []),
    "vca_device_copy": (ctypes.c_int, [_P, _P, ctypes.c_int64, _I32]),
    "vca_sector_spin_coo": (ctypes.c_int, [_P, _I32, ctypes.POINTER(ctypes.c_int64), _P, _P, _P]),
    "vca_lanczos_gs": (ctypes.c_int, [_P, ctypes.c_double, _I32, _D, _P, ctypes.POINTER(_I32)]),
    "vca_lanczos_eigs": (ctypes.c_int, [_P, _I32, _I32, ctypes.c_double, _I32, _P, _P,
                                        ctypes.POINTER(_I32)]),
    "vca_lanczos_tridiag": (ctypes.c_int, [_P, _P, _I32, _P, _P]),
    "vca_lanczos_of_sum": (ctypes.c_int, [_I32, _P, _P, ctypes.c_double, _P, _P,
                                          ctypes.c_double, _P, _P, _D]),
    "vca_gf_poles": (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _I32, _I32, _I32,
                                    ctypes.POINTER(_I32), _P, _P]),
    "vca_gf_cluster": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double, _P]),
    "vca_omega_integrand": (ctypes.c_int, [_P, _I32, _P, _I32, _P]),
    "vca_omega_integrand_complex": (ctypes.c_int, [_P, _I32, _P, _I32, _P]),
    "vca_get_gf_grid": (ctypes.c_int, [_P, _I32, _P, _P, _P]),
    "vca_save_gfprime": (ctypes.c_int, [_P, ctypes.c_char_p, _I32, _I32]),
    "vca_read_gfprime": (ctypes.c_int, [_P, ctypes.c_char_p, _I32, _I32]),
    "vca_omega_restart": (ctypes.c_int, [_P, _P, _P, ctypes.c_int64, _P, _P, ctypes.c_double]),
}
EXPORTS = tuple(_SIGS)

_LIB = None


def load():
    """Load libvca_gpu.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m vca_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


class VcaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vca error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != VCA_OK:
        raise VcaError(rc, load().vca_last_error().decode(errors="replace"))
