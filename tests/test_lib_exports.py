"""CPU checks of the C-ABI library: it loads, and exports every entry point
the header declares (no compute call without a GPU)."""
import ctypes
import os
import re

from vca_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "vca_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vca_[a-z_0-9]+)\s*\(", text)))


def test_header_lists_the_abi():
    names = header_functions()
    assert "vca_solve" in names and "vca_hxv" in names and "vca_omega_integrand" in names
    assert sorted(_lib.EXPORTS) == names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in header_functions():
        assert hasattr(lib, name), name


def test_params_struct_layout_matches_defaults():
    lib = _lib.load()
    p = _lib.CParams()
    lib.vca_default_params(ctypes.byref(p))
    # reference defaults, VCA_INPUT_VARS.f90:112-172
    assert p.nlat == 1 and p.norb == 1 and p.nspin == 1
    assert p.hfmode == 1 and p.hfshift == 1
    assert p.uloc[0] == 2.0 and p.beta == 1000.0
    assert p.lanc_nstates_sector == 6 and p.lanc_niter == 512 and p.lanc_ngfiter == 200
    assert p.lanc_dim_threshold == 256 and p.lanc_tolerance == 1e-12 and p.gs_threshold == 1e-9
    assert ctypes.sizeof(_lib.CParams) >= 150


def test_errors_are_reported_not_aborted():
    lib = _lib.load()
    p = _lib.CParams()
    lib.vca_default_params(ctypes.byref(p))
    p.nspin = 3
    h = ctypes.c_void_p()
    rc = lib.vca_init_solver(ctypes.byref(p), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert b"nspin" in lib.vca_last_error()
    rc = lib.vca_get_sft_potential(None, None)
    assert rc == -1


def test_ctypes_structs_match_the_header(tmp_path):
    """sizeof(vca_params) / sizeof(vca_results) of include/vca_gpu.h equal
    the ctypes mirrors (compiled with the host C compiler)."""
    import ctypes
    import os
    import subprocess

    from vca_amd import _lib

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include "vca_gpu.h"\n'
                   'int main(void){printf("%zu %zu\\n", sizeof(vca_params), '
                   'sizeof(vca_results));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.check_call(["gcc", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)])
    a, b = subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()
    assert int(a) == ctypes.sizeof(_lib.CParams)
    assert int(b) == ctypes.sizeof(_lib.CResults)
