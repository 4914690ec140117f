"""The Fortran drop-in: module vca_gpu (iso_c_binding over the C ABI) and a
reference-style driver compiled with the image's Fortran compiler."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "vca_amd", "fortran")
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

needs_fc = pytest.mark.skipif(not os.path.exists(FC), reason="no Fortran compiler in the image")


def _build(tmp_path):
    mod = tmp_path / "mod"
    mod.mkdir()
    obj = tmp_path / "vca_gpu.o"
    subprocess.check_call([FC, "-c", "-fPIC", f"-J{mod}", os.path.join(FDIR, "vca_gpu.f90"),
                           "-o", str(obj)])
    exe = tmp_path / "example_chain4"
    libdir = os.path.join(ROOT, "vca_amd")
    subprocess.check_call([FC, f"-I{mod}", os.path.join(FDIR, "example_chain4.f90"), str(obj),
                           f"-L{libdir}", "-lvca_gpu", f"-Wl,-rpath,{libdir}", "-o", str(exe)])
    return exe


@needs_fc
def test_fortran_module_and_driver_compile(tmp_path):
    exe = _build(tmp_path)
    assert exe.exists()


@needs_fc
@pytest.mark.gpu
def test_fortran_driver_matches_python_mirror(tmp_path):
    from vca_amd import configs
    from vca_amd.api import VcaSolver

    exe = _build(tmp_path)
    print("running the Fortran driver", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [x for x in r.stdout.splitlines() if x.startswith("SFT_POTENTIAL")][0]
    sft_f = float(line.split()[1])
    p, hloc, hk, _, _ = configs.config_chain4(nk=16)
    s = VcaSolver(p)
    sft_py = s.vca_solve(hloc, hk)
    assert abs(sft_f - sft_py) < 1e-12 * abs(sft_py)
    assert (tmp_path / "SFT_potential.vca").exists()
    # grid getters: full Matsubara Sigma and the (1,2) slice of real-axis G

    def val(tag):
        x = [ln for ln in r.stdout.splitlines() if ln.startswith(tag + " ")][0].split()
        return complex(float(x[1]), float(x[2]))

    sm = s.vca_get_sigma_matsubara()
    gr = s.vca_get_gimp_realaxis(1, 2)
    for tag, ref in (("SIGMA_MATS_11_1", sm[0, 0, 0, 0, 0, 0, 0]),
                     ("SIGMA_MATS_23_7", sm[1, 2, 0, 0, 0, 0, 6]),
                     ("GREAL_12_2500", gr[0, 0, 0, 0, 2499])):
        assert abs(val(tag) - ref) < 1e-12 * max(1.0, abs(ref)), (tag, val(tag), ref)
    # save_gfprime from Fortran: the same records as the Python mirror's file
    from oracle import vca_ref as R
    from tests._models import to_oracle
    s.save_gfprime(tmp_path / "gfprime_py")
    gf_f = R.read_gfprime(to_oracle(p), tmp_path / "gfprime_f.restart")
    gf_p = R.read_gfprime(to_oracle(p), str(tmp_path / "gfprime_py") + ".restart")
    assert gf_f.elems.keys() == gf_p.elems.keys()
    for k in gf_p.elems:
        for a, b in zip(gf_f.elems[k], gf_p.elems[k]):
            for (w1, p1), (w2, p2) in zip(a, b):
                assert len(p1) == len(p2)
                assert np.allclose(p1, p2, rtol=0, atol=1e-11) and np.allclose(w1, w2, atol=1e-11)


@needs_fc
def test_fortran_struct_layouts_match_the_c_abi(tmp_path):
    """The bind(C) mirrors of vca_params and vca_results have the C sizes
    (the ctypes mirrors are the same structs)."""
    import ctypes

    from vca_amd import _lib

    mod = tmp_path / "mod"
    mod.mkdir()
    obj = tmp_path / "vca_gpu.o"
    subprocess.check_call([FC, "-c", "-fPIC", f"-J{mod}", os.path.join(FDIR, "vca_gpu.f90"),
                           "-o", str(obj)])
    src = tmp_path / "sizes.f90"
    src.write_text("program sizes\n  use iso_c_binding\n  use vca_gpu\n"
                   "  type(vca_params_t) :: p\n  type(vca_results_t) :: r\n"
                   "  print '(I0,1X,I0)', c_sizeof(p), c_sizeof(r)\nend program sizes\n")
    exe = tmp_path / "sizes"
    libdir = os.path.join(ROOT, "vca_amd")
    subprocess.check_call([FC, f"-I{mod}", str(src), str(obj), f"-L{libdir}", "-lvca_gpu",
                           f"-Wl,-rpath,{libdir}", "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60).stdout.split()
    assert int(out[0]) == ctypes.sizeof(_lib.CParams)
    assert int(out[1]) == ctypes.sizeof(_lib.CResults)


def _build_program(tmp_path, src_path, name):
    mod = tmp_path / "mod"
    mod.mkdir(exist_ok=True)
    obj = tmp_path / "vca_gpu.o"
    if not obj.exists():
        subprocess.check_call([FC, "-c", "-fPIC", f"-J{mod}",
                               os.path.join(FDIR, "vca_gpu.f90"), "-o", str(obj)])
    exe = tmp_path / name
    libdir = os.path.join(ROOT, "vca_amd")
    subprocess.check_call([FC, f"-I{mod}", str(src_path), str(obj), f"-L{libdir}", "-lvca_gpu",
                           f"-Wl,-rpath,{libdir}", "-o", str(exe)])
    return exe


@needs_fc
@pytest.mark.gpu
def test_fortran_operator_plugin_lanczos_matches_oracle(tmp_path):
    """A reference-style solver loop in Fortran (example_operator.f90)
    drives H x v through spHtimesV_p after build_Hv_sector, on a 3-site
    cluster with 3 bath levels, sector (3,2): the product of the start vector
    equals the oracle's stored-path H x v and 20 Lanczos coefficients equal
    the oracle's recurrence."""
    from oracle import vca_ref as R
    from . import _models as M

    p, hloc, hk, bh, bv = M.bath_model(nlat=3, nb=3, hfshift=False)
    om = M.oracle_model(p, hloc, bh, bv)
    du, dd = om.dims(3, 2)
    rng = np.random.default_rng(21)
    v0 = rng.normal(size=du * dd) + 1j * rng.normal(size=du * dd)
    with open(tmp_path / "operator_in.bin", "wb") as f:
        for a in (hloc, bh, bv, v0):
            f.write(np.asfortranarray(a, dtype=np.complex128).tobytes(order="F"))
    exe = _build_program(tmp_path, os.path.join(FDIR, "example_operator.f90"), "example_operator")
    print("running the Fortran operator loop", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    hv = np.fromfile(tmp_path / "operator_hv.bin", dtype=np.complex128)
    ref = om.hxv(3, 2, v0)
    assert np.linalg.norm(hv - ref) < 1e-12 * np.linalg.norm(ref)
    ab = np.array([[float(x) for x in ln.split()[2:4]] for ln in r.stdout.splitlines()
                   if ln.startswith("AB ")])
    ao, bo = R.lanczos_tridiag(lambda x: om.hxv(3, 2, x), v0 / np.linalg.norm(v0), 20)
    assert ab.shape == (20, 2)
    assert np.allclose(ab[:, 0], ao, rtol=1e-10, atol=1e-10)
    assert np.allclose(ab[:, 1], bo, rtol=1e-10, atol=1e-10)


INIT_NO_BATH = """program init_no_bath
  use vca_gpu
  implicit none
  Nlat = 2
  Nlat_bath = 2
  Norb_bath = 1
  call vca_init_solver()
  write(*, "(A,I0,1X,I0)") "NBATH ", Nlat_bath, Norb_bath
  call vca_finalize_solver()
end program init_no_bath
"""

INIT_BAD_BATH = """program init_bad_bath
  use vca_gpu
  implicit none
  complex(8) :: bh(1, 1, 1, 1, 1, 1), bv(2, 1, 1, 1, 1, 1)
  Nlat = 2
  Nlat_bath = 0
  Norb_bath = 1
  bh = 0
  bv = 0
  call vca_init_solver(bh, bv)
  write(*, "(A)") "NOT STOPPED"
end program init_bad_bath
"""


@needs_fc
def test_fortran_init_solver_stops_on_wrong_bath_dimensions(tmp_path):
    """VCA_MAIN.f90:55-56: bath arrays with Nlat_bath < 1 or Norb_bath < 1
    stop the program (before any device work)."""
    src = tmp_path / "init_bad_bath.f90"
    src.write_text(INIT_BAD_BATH)
    exe = _build_program(tmp_path, src, "init_bad_bath")
    r = subprocess.run([str(exe)], capture_output=True, text=True, cwd=tmp_path, timeout=60)
    out = r.stdout + r.stderr
    assert "wrong bath dimensions" in out and "NOT STOPPED" not in out


@needs_fc
@pytest.mark.gpu
def test_fortran_init_solver_without_bath_zeroes_the_bath(tmp_path):
    """VCA_MAIN.f90:59-62: without bath arrays the cluster has no bath,
    whatever NLAT_BATH / NORB_BATH were."""
    src = tmp_path / "init_no_bath.f90"
    src.write_text(INIT_NO_BATH)
    exe = _build_program(tmp_path, src, "init_no_bath")
    r = subprocess.run([str(exe)], capture_output=True, text=True, cwd=tmp_path, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Bath not present, setting Nbath to 0" in r.stdout
    assert "NBATH 0 0" in r.stdout
