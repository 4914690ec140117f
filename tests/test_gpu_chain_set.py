"""GPU: the excitation chains of all small target sectors of a kept state run
as one chain set (two launches per Lanczos step over every chain of every
sector, 16-step graphs), against the same solve with one batch per target
sector (VCA_CHAIN_SET_MAX_DIM=0) and against the oracle.

The two paths run the same Lanczos recurrences with different kernels (the
partial sums of alpha and beta are taken over different blocks), so they
agree to rounding: chains of 80 steps have not lost orthogonality, and
rounding is not amplified (1e-10 of G''s largest element, 1e-12 in Omega).
Graph replays against eager launches of the chain set: bit for bit.
"""
import numpy as np
import pytest

from oracle import vca_ref as R
from vca_amd import configs
from vca_amd._lib import VcaError
from vca_amd.api import VcaSolver

from . import _models as M

pytestmark = pytest.mark.gpu

ZS = (0.3j, 1.7j, 0.5 + 0.2j, -2.0 + 1.0j)


def _say(*a):
    print(*a, flush=True)


def _case(name):
    if name == "chain10":  # Nspin=1, real H: real and complex segments
        p, hloc, hk, bh, bv = M.chain(nlat=10, u=4.0, nk=4)
        p.lanc_ngfiter = 80
    elif name == "chain10_mixed":
        p, hloc, hk, bh, bv = M.chain(nlat=10, u=4.0, nk=4)
        p.lanc_ngfiter = 80
        p.precision = "mixed"
    elif name == "multiorb":  # Nspin=2, complex hopping: four target sectors
        p, hloc, hk, bh, bv = M.multiorb()
    elif name == "bath":
        p, hloc, hk, bh, bv = M.bath_model(nlat=4, nb=4)
        p.lanc_ngfiter = 80
    elif name == "bhz":
        p, hloc, hk, bh, bv = M.bhz_small(nk=4)
    elif name == "finite_t":  # several kept states, Boltzmann weights
        p, hloc, hk, bh, bv = M.chain(nlat=6, u=4.0, nk=4)
        p.beta = 20.0
        p.lanc_nstates_total = 6
        p.lanc_method = "arpack"
    else:
        raise KeyError(name)
    return p, hloc, hk, bh, bv


def _solve(p, hloc, hk, bh, bv):
    s = VcaSolver(p)
    om = s.vca_solve(hloc, hk, bh, bv)
    return s, om


@pytest.mark.parametrize("name", ["chain10", "chain10_mixed", "multiorb", "bath", "bhz",
                                  "finite_t"])
def test_chain_set_matches_per_sector_batches(name, monkeypatch):
    p, hloc, hk, bh, bv = _case(name)
    out = {}
    for mode, cap in (("per_sector", "0"), ("chain_set", None)):
        if cap is None:
            monkeypatch.delenv("VCA_CHAIN_SET_MAX_DIM", raising=False)
        else:
            monkeypatch.setenv("VCA_CHAIN_SET_MAX_DIM", cap)
        _say(f"{name}: solve, {mode}")
        s, om = _solve(p, hloc, hk, bh, bv)
        res = s.results()
        out[mode] = (om, res, [s.vca_gf_cluster(z) for z in ZS])
    (om_a, ra, ga), (om_b, rb, gb) = out["per_sector"], out["chain_set"]
    # the chain set runs real-vector chains on any sector shape, the
    # per-sector batches only where the real H x v fits: the c_i - i c_j
    # channels taken from diagonal chains may differ between the two
    assert ra["nchains"] + ra["nchains_derived"] == rb["nchains"] + rb["nchains_derived"]
    if ra["nchains_derived"] == rb["nchains_derived"]:
        assert ra["hxv_count"] == rb["hxv_count"]
    assert ra["omegaprime"] == rb["omegaprime"]
    bar = 1e-10 if p.precision == "fp64" else 1e-5
    for z, x, y in zip(ZS, ga, gb):
        err = np.abs(x - y).max() / np.abs(x).max()
        _say(f"  G'({z}): {err:.2e}")
        assert err <= bar, (z, err)
    assert abs(om_a - om_b) <= (1e-12 if p.precision == "fp64" else 1e-8) * abs(om_a)


def test_chain_set_uneven_lengths_against_oracle():
    """The 4-site chain: every target sector is shorter than lanc_ngfiter,
    so each chain runs min(dim, 200) steps (dims 4 to 24 in one set), and
    the oracle's G' and Omega at the pipeline bars (1e-8)."""
    p, hloc, hk, _, _ = configs.config_chain4(nk=16)
    _say("chain4: solve with the chain set")
    s, om = _solve(p, hloc, hk, None, None)
    ref = R.vca_solve(M.to_oracle(p), hloc, hk, epsabs=1e-11, epsrel=1e-12)
    for z in ZS:
        g = s.vca_gf_cluster(z)
        go = R.gf_cluster(ref["gf"], z)
        assert np.abs(g - go).max() < 1e-8 * max(1, np.abs(go).max()), z
    assert abs(om - ref["sft_potential"]) < 1e-8 * abs(ref["sft_potential"])


def test_chain_set_graph_replays_match_eager(monkeypatch):
    """Chain-set steps replayed from the captured 16-step graph (80 steps:
    five replays) against the same steps launched one by one: G' and the
    stored poles bit for bit."""
    p, hloc, hk, _, _ = M.chain(nlat=10, u=4.0, nk=4)
    p.lanc_ngfiter = 80
    out = {}
    for ng in ("1", "0"):
        monkeypatch.setenv("VCA_NO_GRAPHS", ng)
        _say(f"chain10: chain set, VCA_NO_GRAPHS={ng}")
        s, om = _solve(p, hloc, hk, None, None)
        out[ng] = (om, [s.vca_gf_cluster(z) for z in ZS], s.gf_poles(1, 2, 1, 1, 1, 1, 3))
    assert out["1"][0] == out["0"][0]
    for x, y in zip(out["1"][1], out["0"][1]):
        assert np.array_equal(x, y)
    for x, y in zip(out["1"][2], out["0"][2]):
        assert np.array_equal(x, y)


def test_chain_set_row_and_column_items_agree(monkeypatch):
    """The chain set's H x v layouts -- row blocks, and groups of 1, 2, 4 or
    8 whole columns staged in LDS (the 10-site chain's target columns have
    210 rows) -- and the timed choice among them give the same solve to
    rounding (each row's sum is the same; the alpha partials are grouped
    differently)."""
    p, hloc, hk, _, _ = M.chain(nlat=10, u=4.0, nk=4)
    p.lanc_ngfiter = 80
    out = {}
    for mode in ("0", "1", "2", "4", "8", None):
        if mode is None:
            monkeypatch.delenv("VCA_CHAIN_SET_COLUMNS", raising=False)
        else:
            monkeypatch.setenv("VCA_CHAIN_SET_COLUMNS", mode)
        _say(f"chain10: chain set, VCA_CHAIN_SET_COLUMNS={mode}")
        s, om = _solve(p, hloc, hk, None, None)
        out[mode] = (om, [s.vca_gf_cluster(z) for z in ZS])
    for m in ("1", "2", "4", "8", None):
        assert abs(out[m][0] - out["0"][0]) <= 1e-12 * abs(out["0"][0])
        for x, y in zip(out[m][1], out["0"][1]):
            assert np.abs(x - y).max() <= 1e-10 * np.abs(y).max()


@pytest.mark.parametrize("bad", ["9", "-1", "x", "2x"])
def test_chain_set_layout_knob_refuses_bad_values(bad, monkeypatch):
    p, hloc, hk, _, _ = M.chain(nlat=10, u=4.0, nk=4)
    p.lanc_ngfiter = 16
    monkeypatch.setenv("VCA_CHAIN_SET_COLUMNS", bad)
    with pytest.raises(VcaError):
        _solve(p, hloc, hk, None, None)
