"""The cluster Green's function pole file (save_gfprime / read_gfprime,
VCA_AUX_FUNX.f90:484-697).

Fixtures (tests/golden, made by make_gfprime_golden.py): an unformatted
file written by a Fortran program with the reference's defined-I/O record
structure (gfprime_writer.f90, built with the image's Fortran compiler) and
a formatted file in gfortran's list-directed layout, both with the contents
of golden_gf_contents().  CPU tests pin the oracle's reader and writer on
them; GPU tests hold the library's reader and writer to the same files and
to the oracle, bit for bit, and check the restart path (Omega from a pole
store read back equals Omega of the solve that wrote it).
"""
import os

import numpy as np
import pytest

from oracle import vca_ref as R
from vca_amd import configs

from . import _models as M
from .golden.make_gfprime_golden import golden_gf_contents

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FLANG = os.path.join(GOLDEN, "gfprime_flang.restart")
GFSTYLE = os.path.join(GOLDEN, "gfprime_gfortran_style.restart")


def _same(gd_elems, expected):
    assert set(gd_elems) == set(expected)
    for k, states in expected.items():
        got = gd_elems[k]
        assert len(got) == len(states), k
        for a, b in zip(got, states):
            assert len(a) == len(b), k
            for (w1, p1), (w2, p2) in zip(a, b):
                assert np.array_equal(np.asarray(w1, complex), np.asarray(w2, complex)), k
                assert np.array_equal(np.asarray(p1, complex), np.asarray(p2, complex)), k


@pytest.mark.parametrize("path,formatted", [(FLANG, False), (GFSTYLE, True)])
def test_oracle_reads_golden_files(path, formatted):
    gd = R.read_gfprime(R.Params(nlat=2), path, formatted=formatted)
    _same(gd.elems, golden_gf_contents())


def test_golden_unformatted_record_framing():
    """One record per element with 4-byte markers: 4 records, the first
    holding 2 states (4 + 2 * (4 + channels))."""
    data = open(FLANG, "rb").read()
    pos, n = 0, 0
    while pos < len(data):
        m1 = int.from_bytes(data[pos:pos + 4], "little", signed=True)
        m2 = int.from_bytes(data[pos + 4 + m1:pos + 8 + m1], "little", signed=True)
        assert m1 == m2 > 0
        pos += m1 + 8
        n += 1
    assert n == 4 and pos == len(data)


@pytest.mark.parametrize("formatted", [False, True])
def test_oracle_round_trip_of_a_solved_gf(tmp_path, formatted):
    p, hloc, hk, _, _ = configs.config_chain4(nk=2)
    po = M.to_oracle(p)
    r = R.vca_solve(po, hloc, hk, epsabs=1e-6, epsrel=1e-6)
    f = tmp_path / "gfprime.restart"
    R.save_gfprime(r["gf"], f, formatted=formatted)
    back = R.read_gfprime(po, f, formatted=formatted)
    expect = {k: [[(np.asarray(w, complex), np.asarray(pl, complex)) for (w, pl) in chans]
                  for chans in states] for k, states in r["gf"].elems.items()}
    _same(back.elems, expect)
    for z in (0.4j, 1.0 + 0.3j):
        assert np.array_equal(R.gf_cluster(back, z), R.gf_cluster(r["gf"], z))


def test_formatted_reader_is_list_directed(tmp_path):
    """Repeat counts, values split over lines and D exponents read as
    list-directed input does."""
    text = open(GFSTYLE).read().split("\n")[0]
    # element 1 folded over lines between its complex items, with D exponents
    items = text.split(") (")
    wrapped = ")\n(".join(items)
    lines = open(GFSTYLE).read().split("\n")
    lines[0] = wrapped.replace("E-002", "D-002")
    f = tmp_path / "x.restart"
    f.write_text("\n".join(lines))
    gd = R.read_gfprime(R.Params(nlat=2), f, formatted=True)
    _same(gd.elems, golden_gf_contents())
    # element 1: 2 states of 1 channel, the first with one pole (1, 0) of
    # weight (1, 0) given as 2*(1.0,0.0), the second empty
    rep = "           2 1 1 2*(1.0,0.0) \\n 1 0 \\n\n" + "\n".join(lines[1:])
    f.write_text(rep)
    gd = R.read_gfprime(R.Params(nlat=2), f, formatted=True)
    (w, pl), = gd.elems[(1, 1, 1, 1, 1)][0]
    assert np.array_equal(w, [1.0]) and np.array_equal(pl, [1.0])
    (w, pl), = gd.elems[(1, 1, 1, 1, 1)][1]
    assert len(w) == 0 and len(pl) == 0


# ----------------------------------------------------------------- GPU --
def _poles_of(s, key, nstates, nchan):
    il, jl, sp, io, jo = key
    out = []
    for st in range(1, nstates + 1):
        chans = []
        for ch in range(1, nchan + 1):
            w, pl = s.gf_poles(il, jl, sp, io, jo, st, ch)
            chans.append((w, pl.astype(complex)))
        out.append(chans)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("path,formatted", [(FLANG, False), (GFSTYLE, True)])
def test_library_reads_golden_files(tmp_path, path, formatted):
    from vca_amd.api import VcaParams, VcaSolver
    s = VcaSolver(VcaParams(nlat=2, lanc_method="lanczos"))
    stem = tmp_path / "g"
    os.symlink(path, str(stem) + ".restart")
    s.read_gfprime(stem, use_formatted=formatted)
    ex = golden_gf_contents()
    got = {k: _poles_of(s, k, 2, 2 if k[0] == k[1] else 4) for k in ex}
    _same(got, ex)
    # written back unformatted it is the Fortran runtime's file byte for byte
    s.save_gfprime(tmp_path / "h")
    assert open(str(tmp_path / "h") + ".restart", "rb").read() == open(FLANG, "rb").read()


@pytest.mark.gpu
@pytest.mark.parametrize("formatted", [False, True])
def test_library_writer_against_oracle_and_restart(tmp_path, formatted):
    from vca_amd.api import VcaSolver
    p, hloc, hk, _, _ = configs.config_chain4(nk=8)
    print(f"gfprime chain4 formatted={formatted}", flush=True)
    s = VcaSolver(p)
    om = s.vca_solve(hloc, hk)
    res = s.results()
    s.save_gfprime(tmp_path / "gfprime", use_formatted=formatted)
    gd = R.read_gfprime(M.to_oracle(p), str(tmp_path / "gfprime") + ".restart",
                        formatted=formatted)
    for k, states in gd.elems.items():
        _same({k: states}, {k: _poles_of(s, k, len(states), len(states[0]))})
    # a fresh solver reads it back: same G', same Omega without diagonalising
    s2 = VcaSolver(p)
    s2.read_gfprime(tmp_path / "gfprime", use_formatted=formatted)
    for z in (0.3j, 1.2 + 0.4j):
        assert np.array_equal(s2.vca_gf_cluster(z), s.vca_gf_cluster(z))
    om2 = s2.omega_restart(hloc, hk, res["omegaprime"])
    assert abs(om2 - om) < 1e-12 * abs(om), (om2, om)
    r2 = s2.results()
    assert r2["t_diag"] == 0.0 and r2["nchains"] == 0
    g = s2.gf_grid("matsubara", ("sigma",))["sigma"]
    assert np.array_equal(g, s.gf_grid("matsubara", ("sigma",))["sigma"])


@pytest.mark.gpu
def test_restart_at_finite_temperature(tmp_path):
    from vca_amd.api import VcaSolver
    p, hloc, hk, _, _ = M.chain(nlat=4, u=4.0, nk=8)
    p.beta, p.bandwidth, p.lanc_nstates_total = 20.0, 4.0, 2
    print("gfprime finite T", flush=True)
    s = VcaSolver(p)
    om = s.vca_solve(hloc, hk)
    res = s.results()
    s.save_gfprime(tmp_path / "ft")
    s2 = VcaSolver(p)
    s2.read_gfprime(tmp_path / "ft")
    om2 = s2.omega_restart(hloc, hk, res["omegaprime"])
    # the file has no excitation energies: the Matsubara cutoff comes from
    # the largest |pole|, so the sum/contour split moves, not the integral
    assert s2.results()["nmax"] >= res["nmax"]
    assert abs(om2 - om) < 1e-9 * abs(om), (om2, om)


@pytest.mark.gpu
def test_read_errors(tmp_path):
    from vca_amd.api import VcaError, VcaParams, VcaSolver
    s = VcaSolver(VcaParams(nlat=2, lanc_method="lanczos"))
    with pytest.raises(VcaError):
        s.read_gfprime(tmp_path / "missing")
    with pytest.raises(VcaError):
        s.save_gfprime(tmp_path / "nothing")  # no pole store yet
    # a complex pole is refused
    gd = R.read_gfprime(R.Params(nlat=2), FLANG)
    w, pl = gd.elems[(1, 1, 1, 1, 1)][0][0]
    gd.elems[(1, 1, 1, 1, 1)][0][0] = (w, pl + 0.5j)
    R.save_gfprime(gd, str(tmp_path / "cx") + ".restart")
    with pytest.raises(VcaError):
        s.read_gfprime(tmp_path / "cx")
    # a file for a smaller cluster ends early
    with pytest.raises(VcaError):
        VcaSolver(VcaParams(nlat=3, lanc_method="lanczos")).read_gfprime(
            str(FLANG)[:-len(".restart")])
