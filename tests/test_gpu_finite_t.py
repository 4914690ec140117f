"""GPU parity of the ARPACK-equivalent eigensolver and the finite-T path.

Bars: eigenvalues within 1e-10 of dense eigh, residuals |Hx - e x| below
1e-8, orthonormal eigenvectors; complex Omega integrand within 1e-9;
Omega within 1e-8 of the oracle; at U=0 the finite-T Omega within 2e-8 of
the closed form (the CUTOFF truncation of the kept spectrum bounds it).
"""
import numpy as np
import pytest

from oracle import vca_ref as R
from vca_amd import configs
from vca_amd.api import VcaParams, VcaSolver

from . import _models as M

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _solver(p, hloc, bath_h=None, bath_v=None):
    s = VcaSolver(p)
    if bath_h is not None:
        s.vca_set_bath(bath_h, bath_v)
    s.vca_set_hcluster(hloc)
    return s


def _bhz2x2():
    p = VcaParams(nlat=4, norb=2, nspin=2, uloc=(2.0, 2.0, 0, 0, 0), lanc_method="lanczos")
    hloc, hk = configs.bhz(2, 2, 2.0, 0.3, nk=(2, 2))
    return p, hloc, hk, None, None


# name -> (model, sector, lowest levels non-degenerate)
EIG_CASES = {
    "chain8_43": (lambda: M.chain(nlat=8, u=4.0), (4, 3), True),
    "bath3+3_33": (lambda: M.bath_model(), (3, 3), True),
    "bhz2x2_43": (_bhz2x2, (4, 3), True),
    "chain8_u0_44": (lambda: M.chain(nlat=8, u=0.0), (4, 4), False),
}


@pytest.mark.parametrize("name", sorted(EIG_CASES))
def test_eigs_against_arpack_and_dense(name):
    """The GPU thick-restart solver returns ARPACK's set: the oracle's
    eigsh (ARPACK itself, with the reference's which/ncv/tol) -- also where
    a degenerate level is only partly reached from one start vector (U=0
    chain: 2 of the 4 copies of a level, as ARPACK) -- and, for
    non-degenerate spectra, dense eigh."""
    mk, sector, nondegenerate = EIG_CASES[name]
    p, hloc, hk, bh, bv = mk()
    om = M.oracle_model(p, hloc, bh, bv)
    s = _solver(p, hloc, bh, bv)
    sec = s.sector(*sector)
    nev = 6
    _say(f"eigs {name} {sector} dim {sec.dim} nev {nev}")
    ev, vecs, nconv = sec.lanczos_eigs(nev, tol=1e-12, maxiter=512)
    assert nconv == nev
    om.p.lanc_method = "arpack"
    ea, _ = R.solve_sector(om, *sector, nev)
    scale = max(1.0, np.abs(ea).max())
    assert np.allclose(ev, ea, rtol=0, atol=1e-10 * scale), (ev, ea)
    if nondegenerate:
        ed = np.linalg.eigvalsh(om.dense_h(*sector))[:nev]
        assert np.allclose(ev, ed, rtol=0, atol=1e-10 * scale), (ev, ed)
    for i in range(nev):
        r = om.hxv(*sector, vecs[i]) - ev[i] * vecs[i]
        assert np.linalg.norm(r) < 1e-8, (i, np.linalg.norm(r))
    gram = vecs.conj() @ vecs.T
    assert np.abs(gram - np.eye(nev)).max() < 1e-10
    sec.close()


def _oracle_solve(p, hloc, hk, bh=None, bv=None):
    return R.vca_solve(M.to_oracle(p), hloc, hk, bh, bv, epsabs=1e-11, epsrel=1e-12)


def test_arpack_mode_solve_matches_oracle():
    """LANC_METHOD=arpack with Lanczos sectors (threshold lowered so every
    sector above 40 states goes through the thick-restart solver)."""
    p, hloc, hk, _, _ = M.chain(nlat=6, u=4.0, nk=8)
    p.lanc_method = "arpack"
    p.lanc_dim_threshold = 40
    _say("arpack-mode vca_solve chain6")
    s = VcaSolver(p)
    omega = s.vca_solve(hloc, hk)
    ref = _oracle_solve(p, hloc, hk)
    res = s.results()
    assert res["nstates"] == len(ref["states"])
    assert abs(res["omegaprime"] - ref["e0"]) < 1e-10 * abs(ref["e0"])
    assert abs(omega - ref["sft_potential"]) < 1e-8 * max(1, abs(ref["sft_potential"]))


@pytest.mark.parametrize("name", ["chain4", "bhz"])
def test_complex_integrand(name):
    if name == "chain4":
        p, hloc, hk, bh, bv = M.chain(nlat=4, u=4.0, nk=8)
    else:
        p, hloc, hk, bh, bv = M.bhz_small(nk=4)
    _say(f"complex integrand {name}")
    s = VcaSolver(p)
    s.vca_solve(hloc, hk, bh, bv)
    ref = _oracle_solve(p, hloc, hk, bh, bv)
    hkl = R.hk_to_lso(ref["model"].p, hk)
    zs = np.array([0.3j, 2.0j, 0.5 + 0.2j, -3.0 + 0.01j, 6.0 * np.exp(0.1j), 6.0 * np.exp(3.0j)])
    f = s.omega_integrand_complex(zs)
    fo = np.array([R.sum_kmesh(ref["model"], ref["gf"], hkl, z, complex_out=True) for z in zs])
    assert np.allclose(f, fo, rtol=1e-9, atol=1e-10), (f, fo)
    # the real part is the log|det| integrand
    assert np.allclose(f.real, s.omega_integrand(zs), rtol=1e-12, atol=1e-13)


def _ft(p, beta, bandwidth=4.0, nstates=2):
    p.beta, p.lanc_nstates_total, p.bandwidth = beta, nstates, bandwidth
    return p


def test_finite_t_solve_matches_oracle():
    """Finite T (LANC_NSTATES_TOTAL=2, annealed until the kept spectrum
    reaches CUTOFF: ground state plus the full triplet), ARPACK sectors
    above 10 states, U=4 chain: state list, zeta, Boltzmann weights, Nmax
    and the Matsubara + contour integral against the oracle."""
    p, hloc, hk, _, _ = M.chain(nlat=4, u=4.0, nk=8)
    p.lanc_method = "arpack"
    p.lanc_dim_threshold = 10
    _ft(p, 20.0)
    _say("finite-T vca_solve chain4 beta=20")
    s = VcaSolver(p)
    omega = s.vca_solve(hloc, hk)
    ref = _oracle_solve(p, hloc, hk)
    res = s.results()
    assert res["finite_t"] == 1
    assert res["nstates"] == len(ref["states"]), (res["nstates"], len(ref["states"]))
    assert res["spectrum_iters"] == ref["spectrum_iters"]
    assert abs(res["zeta"] - ref["zeta"]) < 1e-10 * ref["zeta"]
    assert res["nmax"] == ref["nmax"]
    assert abs(res["omegaprime"] - ref["omegaprime"]) < 1e-10 * abs(ref["omegaprime"])
    assert abs(omega - ref["sft_potential"]) < 1e-8 * max(1, abs(ref["sft_potential"]))


@pytest.mark.parametrize("beta,tprime", [(50.0, -1.0), (20.0, -0.7)])
def test_finite_t_exact_at_u0(beta, tprime):
    p, hloc, hk, _, _ = M.chain(nlat=4, u=0.0, nk=16, tprime=tprime)
    _ft(p, beta)
    _say(f"finite-T U=0 chain4 beta={beta} t'={tprime}")
    s = VcaSolver(p)
    omega = s.vca_solve(hloc, hk)
    hkl = R.hk_to_lso(M.to_oracle(p), hk)
    eps = np.concatenate([np.linalg.eigvalsh(h) for h in hkl])
    exact = -(3 - p.nspin) / beta * np.sum(np.log1p(np.exp(-beta * eps))) / len(hkl)
    assert abs(omega - exact) < 2e-8, (omega, exact)


def test_logdet_qr_equals_pivoted_lu(monkeypatch):
    """The FP64 log-det of 8-row blocks three ways -- a matrix per lane
    (k_logdet_lane64, the default), Householder QR with 4 lanes per matrix,
    pivoted LU with a row per lane -- gives the same complex integrand
    (log|det| and the principal phase), on 8 k-points (fewer than the lane
    kernel's 32 per workgroup) and 7 frequencies (not a multiple of its 8
    per wavefront), blocks of 4 padded to 8 rows."""
    p, hloc, hk, _, _ = M.bhz_small(nk=8)
    _say("lane vs QR vs LU log-det, BHZ 2 sites")
    s = VcaSolver(p)
    s.vca_solve(hloc, hk)
    zs = np.array([0.05j, 0.7j, 3.0j, 40.0j, 0.4 + 0.3j, -2.0 + 0.02j, 5.0 * np.exp(2.0j)])
    monkeypatch.setenv("VCA_OMEGA_LU", "0")
    f_lane = s.omega_integrand_complex(zs)
    monkeypatch.setenv("VCA_OMEGA_LANE", "0")
    f_qr = s.omega_integrand_complex(zs)
    monkeypatch.setenv("VCA_OMEGA_LU", "1")
    f_lu = s.omega_integrand_complex(zs)
    assert np.allclose(f_qr, f_lu, rtol=1e-12, atol=1e-13), (f_qr, f_lu)
    assert np.allclose(f_lane, f_qr, rtol=1e-12, atol=1e-13), (f_lane, f_qr)
    assert np.allclose(f_lane.real, s.omega_integrand(zs), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("case", ["bhz_blocks_of_8", "chain12_block_of_16"])
def test_mfma_logdet_equals_vector_alu_logdet(case, monkeypatch):
    """The FP32 integrand with the contraction on the matrix cores
    (k_logdet_mfma) and on the vector ALUs (k_logdet<float>): log|det| and
    the principal phase agree to FP32 rounding, on 8-row spin blocks and on
    a 12-site block padded to 16 rows; both within 1e-6 of the FP64
    integrand at the frequencies where it is not ill-conditioned."""
    if case == "bhz_blocks_of_8":
        p, hloc, hk, _, _ = configs.config_bhz_2x2(nk=16, nfreq=64, precision="fp64")
    else:
        p, hloc, hk, _, _ = M.chain(nlat=12, u=0.0, nk=64, tprime=-0.8)
        p.sectors = [(6, 6)]
    _say(f"MFMA vs VALU log-det, {case}")
    s = VcaSolver(p)
    s.vca_solve(hloc, hk)
    zs = np.array([0.3j, 1.1j, 3.0j, 40.0j, 0.4 + 0.3j, -2.0 + 0.2j, 5.0 * np.exp(2.0j)])
    f64 = s.omega_integrand_complex(zs, "fp64")
    monkeypatch.setenv("VCA_OMEGA_MFMA", "0")
    f_valu = s.omega_integrand_complex(zs, "mixed")
    monkeypatch.setenv("VCA_OMEGA_MFMA", "1")
    f_mfma = s.omega_integrand_complex(zs, "mixed")
    scale = np.abs(f64).max()
    _say(f"  MFMA vs VALU {np.abs(f_mfma - f_valu).max() / scale:.1e}, MFMA vs FP64 "
         f"{np.abs(f_mfma - f64).max() / scale:.1e}")
    assert np.abs(f_mfma - f_valu).max() < 2e-6 * scale
    assert np.abs(f_mfma - f64).max() < 2e-6 * scale
    assert np.allclose(f_mfma.real, s.omega_integrand(zs, "mixed"), rtol=1e-12, atol=1e-13)
