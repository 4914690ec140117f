"""GPU: the (c^+_i + i c^+_j) and (c_i - i c_j) channels of a real
Hamiltonian with a real ground state, taken from the two diagonal chains of
their target sector (vca_lanczos_of_sum) instead of run.

Against the oracle, which runs every chain (VCA_GF_NORMAL.f90:121-318,
561-958): 12-step chains on target sectors of 300 states (a chain and a
bath model), 4-step chains on the 24-state sectors of a 2-site, 2-orbital,
2-spin model (no loss of orthogonality in so few steps, so the derived
tridiagonal equals the run one to rounding): E0 1e-10, G'(z) and Omega
1e-8, and the poles and weights of the derived channels 1e-9.

Against the same solve with every chain run (VCA_DERIVED_CHANNELS=0),
80-step chains on 52,920-state sectors, chain set and per-sector batches,
FP64 and MIXED: G' 1e-10 (MIXED 1e-5) of its largest element, Omega 1e-12
(1e-8)."""
import numpy as np
import pytest

from oracle import vca_ref as R

from . import _models as M
from vca_amd.api import VcaSolver

pytestmark = pytest.mark.gpu

ZS = (0.3j, 1.7j, 0.5 + 0.2j, -2.0 + 1.0j)


def _say(*a):
    print(*a, flush=True)


ELEMS = {"chain6": [(1, 2, 1, 1, 1)], "bath3+3": [(1, 2, 1, 1, 1)],
         # two orbitals, two spins: channels between orbitals of one site
         # and between different sites and orbitals, in each spin
         "multiorb_real": [(1, 1, 1, 1, 2), (1, 2, 2, 2, 1), (2, 1, 1, 1, 1)]}


@pytest.mark.parametrize("name", ["chain6", "bath3+3", "multiorb_real"])
def test_derived_channels_against_the_oracle(name):
    if name == "chain6":
        p, hloc, hk, bh, bv = M.chain(nlat=6, u=4.0, nk=8)
        p.lanc_ngfiter = 12
    elif name == "bath3+3":
        p, hloc, hk, bh, bv = M.bath_model(nlat=3, nb=3)
        p.lanc_ngfiter = 12
    else:
        # the 2-site, 2-orbital model (Ust, Jh, spin-dependent levels) with
        # real hopping; 4-step chains on its 24-state target sectors
        p, hloc, hk, bh, bv = M.multiorb()
        hloc = np.asfortranarray(hloc.real.astype(np.complex128))
        hk = np.asfortranarray(hk.real.astype(np.complex128))
        p.lanc_ngfiter = 4
    _say(f"{name}: solve with derived channels")
    s = VcaSolver(p)
    om = s.vca_solve(hloc, hk, bh, bv)
    res = s.results()
    assert res["nchains_derived"] > 0
    ref = R.vca_solve(M.to_oracle(p), hloc, hk, bh, bv, epsabs=1e-11, epsrel=1e-12)
    assert abs(res["omegaprime"] - ref["e0"]) < 1e-10 * abs(ref["e0"])
    for z in ZS:
        g = s.vca_gf_cluster(z)
        go = R.gf_cluster(ref["gf"], z)
        assert np.abs(g - go).max() < 1e-8 * max(1, np.abs(go).max()), z
    assert abs(om - ref["sft_potential"]) < 1e-8 * abs(ref["sft_potential"])
    # the derived channels' Lehmann data
    for key in ELEMS[name]:
        ch_ref = ref["gf"].elems[key][0]
        for ichan in (3, 4):
            w, poles = s.gf_poles(*key, 1, ichan)
            wo, po = ch_ref[ichan - 1]
            assert len(poles) == len(po) == p.lanc_ngfiter, (key, ichan)
            o, oo = np.argsort(poles), np.argsort(po)
            assert np.abs(poles[o] - po[oo]).max() < 1e-9 * np.abs(po).max(), (key, ichan)
            assert np.abs(w[o] - wo[oo]).max() < 1e-9 * np.abs(wo).max(), (key, ichan)


@pytest.mark.parametrize("cap", ["0", None])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_derived_channels_match_run_chains(prec, cap, monkeypatch):
    p, hloc, hk, _, _ = M.chain(nlat=10, u=4.0, nk=4)
    p.lanc_ngfiter = 80
    p.precision = prec
    if cap is None:
        monkeypatch.delenv("VCA_CHAIN_SET_MAX_DIM", raising=False)
    else:
        monkeypatch.setenv("VCA_CHAIN_SET_MAX_DIM", cap)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("VCA_DERIVED_CHANNELS", mode)
        _say(f"chain10 {prec} chain-set cap {cap}: VCA_DERIVED_CHANNELS={mode}")
        s = VcaSolver(p)
        om = s.vca_solve(hloc, hk)
        out[mode] = (om, s.results(), [s.vca_gf_cluster(z) for z in ZS])
    (om_r, rr, gr), (om_d, rd, gd) = out["0"], out["1"]
    assert rr["nchains_derived"] == 0 and rd["nchains_derived"] > 0
    assert rr["nchains"] == rd["nchains"] + rd["nchains_derived"]
    assert rd["hxv_count"] < rr["hxv_count"]
    bar = 1e-10 if prec == "fp64" else 1e-5
    for z, x, y in zip(ZS, gd, gr):
        err = np.abs(x - y).max() / np.abs(y).max()
        _say(f"  G'({z}): {err:.2e}")
        assert err <= bar, (z, err)
    assert abs(om_d - om_r) <= (1e-12 if prec == "fp64" else 1e-8) * abs(om_r)
