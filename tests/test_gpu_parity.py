"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars: hop tables (connectivity, signs, amplitudes) bit-exact; FP64 H x v
within 1e-12 relative (summation order differs); MIXED H x v within 2e-6
relative; E0 within 1e-10 relative; G'(z) and the Omega integrand within
1e-8 relative.
"""
import math

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


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=n) + 1j * rng.normal(size=n)


def _coo_sorted(rows, cols, amps):
    order = np.lexsort((cols, rows))
    return rows[order], cols[order], amps[order]


CASES = {
    "chain6": M.chain,
    "multiorb": M.multiorb,
    "bath": M.bath_model,
    "bath_hfshift": lambda: M.bath_model(hfshift=True),
    "bhz": M.bhz_small,
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_hop_tables_bit_exact(name):
    p, hloc, hk, bh, bv = CASES[name]()
    om = M.oracle_model(p, hloc, bh, bv)
    s = _solver(p, hloc, bh, bv)
    ns = p.ns
    for (nup, ndw) in [(ns // 2, ns // 2), (ns // 2 + 1, ns // 2 - 1), (1, ns - 1)]:
        _say(f"tables {name} sector ({nup},{ndw})")
        sec = s.sector(nup, ndw)
        for spin, npart in ((1, nup), (2, ndw)):
            g = _coo_sorted(*sec.spin_coo(spin))
            o = _coo_sorted(*om.spin_hops(spin, npart))
            assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])
            assert np.array_equal(g[2], o[2])
        sec.close()


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hxv_small(name, precision):
    p, hloc, hk, bh, bv = CASES[name]()
    om = M.oracle_model(p, hloc, bh, bv)
    s = _solver(p, hloc, bh, bv)
    ns = p.ns
    for (nup, ndw) in [(ns // 2, ns // 2), (ns // 2 + 1, ns // 2 - 1), (0, 1), (ns, ns - 1)]:
        _say(f"hxv {name} {precision} sector ({nup},{ndw})")
        sec = s.sector(nup, ndw, precision)
        v = _rand(sec.dim, 1000 + nup)
        hv = sec.hxv(v)
        ref = om.hxv(nup, ndw, v)
        err = np.linalg.norm(hv - ref) / max(np.linalg.norm(ref), 1e-300)
        assert err < (1e-12 if precision == "fp64" else 2e-6), err
        sec.close()


def test_hxv_square_3x4_full_sector():
    """12-site 3x4 Hubbard, sector (6,6): 853,776 states."""
    p, hloc, hk, _, _ = configs.config_square_3x4(nk=2)
    om = M.oracle_model(p, hloc)
    s = _solver(p, hloc)
    _say("hxv square3x4 (6,6)")
    sec = s.sector(6, 6)
    assert sec.dim == 853776
    v = _rand(sec.dim, 77)
    err = np.linalg.norm(sec.hxv(v) - om.hxv(6, 6, v, workers=8)) / np.linalg.norm(v)
    assert err < 1e-12, err


N_HXV_VARIANTS = 29  # launch shapes of the one-column kernels (kernels.hip kVariants)
N_HXV_DIRECT = 15    # the first 15: direct form; then the transposed form (15-18
                     # whole columns, 19-20 one real plane at a time: real
                     # amplitudes only), the overlapped short-column form (21-23)
                     # and its signed-column variant (24-26: one real hopping
                     # amplitude), the transposed form with complex-pair passes
                     # (27-28: real amplitudes)


@pytest.mark.parametrize("variant", range(N_HXV_VARIANTS))
def test_hxv_launch_variants(variant, monkeypatch):
    """Every launch shape the autotuner can pick gives the oracle's H x v
    (one-column kernels, plain and buffer-load forms, and the transposed
    form, which the size threshold is lowered for: chain of 10 sites, bath
    model, complex BHZ on a sector with DimUp != DimDw)."""
    monkeypatch.setenv("VCA_HXV_VARIANT", str(variant))
    if variant >= N_HXV_DIRECT:
        monkeypatch.setenv("VCA_TFORM_MIN_DIM", "0")
    bhz = VcaParams(nlat=5, norb=2, nspin=2, uloc=(2.0, 2.0, 0, 0, 0))
    bhz_h, bhz_hk = configs.bhz(5, 1, 2.0, 0.3, nk=(2, 1))
    for name, (p, hloc, hk, bh, bv), sector in (
            ("chain10", M.chain(nlat=10, u=4.0), (5, 5)),
            ("bath5+5", M.bath_model(nlat=5, nb=5), (5, 5)),
            ("bhz5x1", (bhz, bhz_h, bhz_hk, None, None), (5, 4))):
        om = M.oracle_model(p, hloc, bh, bv)
        s = _solver(p, hloc, bh, bv)
        for prec in ("fp64", "mixed"):
            _say(f"variant {variant} {name} {sector} {prec}")
            sec = s.sector(*sector, prec)
            v = _rand(sec.dim, 31)
            ref = om.hxv(*sector, v)
            err = np.linalg.norm(sec.hxv(v) - ref) / np.linalg.norm(ref)
            assert err < (1e-12 if prec == "fp64" else 2e-6), (name, prec, err)
            sec.close()


def test_lanczos_gs_against_dense():
    p, hloc, hk, _, _ = M.chain(nlat=8, u=4.0)
    om = M.oracle_model(p, hloc)
    s = _solver(p, hloc)
    for prec in ("fp64", "mixed"):
        _say(f"lanczos gs chain8 (4,4) {prec}")
        sec = s.sector(4, 4, prec)
        e0, vec, it = sec.lanczos_gs(tol=1e-12, niter=400)
        ed = np.linalg.eigvalsh(om.dense_h(4, 4))[0]
        assert abs(e0 - ed) < 1e-10 * abs(ed), (prec, e0, ed)
        r = om.hxv(4, 4, vec) - e0 * vec
        assert np.linalg.norm(r) < 1e-5
        assert abs(np.linalg.norm(vec) - 1) < 1e-10
        sec.close()


def test_lanczos_tridiag_first_steps():
    p, hloc, hk, bh, bv = M.bath_model()
    om = M.oracle_model(p, hloc, bh, bv)
    s = _solver(p, hloc, bh, bv)
    ns = p.ns
    nup, ndw = ns // 2, ns // 2
    _say("tridiag bath sector")
    sec = s.sector(nup, ndw)
    v0 = _rand(sec.dim, 5)
    v0 /= np.linalg.norm(v0)
    a, b = sec.lanczos_tridiag(v0, 12)
    ao, bo = R.lanczos_tridiag(lambda x: om.hxv(nup, ndw, x), v0, 12)
    assert np.allclose(a, ao, rtol=1e-10, atol=1e-10)
    assert np.allclose(b, bo, rtol=1e-10, atol=1e-10)


def _oracle_solve(p, hloc, hk, bh=None, bv=None, **kw):
    return R.vca_solve(M.to_oracle(p), hloc, hk, bh, bv, **kw)


@pytest.mark.parametrize("name", ["chain4", "chain4_u0_tp07", "bath", "bath_hfshift_off",
                                  "bhz", "multiorb", "chain4_symmetric_gf"])
def test_vca_solve_pipeline(name):
    if name == "chain4":
        p, hloc, hk, bh, bv = configs.config_chain4(nk=16)
    elif name == "chain4_u0_tp07":
        p, hloc, hk, bh, bv = M.chain(nlat=4, u=0.0, nk=16, tprime=-0.7)
    elif name == "bath":
        p, hloc, hk, bh, bv = M.bath_model(nlat=2, nb=2)
    elif name == "bath_hfshift_off":
        # HFMODE with HFSHIFT=F: the bath levels shifted by -U/2
        # (VCA_BATH_SETUP.f90:127-136)
        p, hloc, hk, bh, bv = M.bath_model(nlat=2, nb=2, hfshift=False)
    elif name == "multiorb":
        # Norb=2, Nspin=2: Ust, Jh, Hartree shifts, complex inter-orbital hops
        p, hloc, hk, bh, bv = M.multiorb()
    elif name == "chain4_symmetric_gf":
        # VCA_GF_SYMMETRIC: 2 channels per off-diagonal pair (chan2)
        p, hloc, hk, bh, bv = configs.config_chain4(nk=16)
        p.gf_symmetric = True
    else:
        p, hloc, hk, bh, bv = M.bhz_small(nk=4)
    _say(f"vca_solve {name}")
    s = VcaSolver(p)
    omega = s.vca_solve(hloc, hk, bh, bv)
    ref = _oracle_solve(p, hloc, hk, bh, bv, epsabs=1e-11, epsrel=1e-12)
    res = s.results()
    assert abs(res["omegaprime"] - ref["e0"]) < 1e-10 * max(1, abs(ref["e0"]))
    # G'(z) at a few points off the real axis
    for z in (0.3j, 1.7j, 0.5 + 0.2j, -2.0 + 1.0j):
        g = s.vca_gf_cluster(z)
        go = R.gf_cluster(ref["gf"], z)
        assert np.abs(g - go).max() < 1e-8 * max(1, np.abs(go).max()), z
    # Omega integrand at fixed frequencies (no quadrature ambiguity)
    hkl = R.hk_to_lso(ref["model"].p, hk)
    zs = np.array([0.05j, 0.4j, 1.0j, 3.0j, 20.0j])
    f = s.omega_integrand(zs)
    fo = np.array([R.sum_kmesh(ref["model"], ref["gf"], hkl, z) for z in zs])
    assert np.allclose(f, fo, rtol=1e-8, atol=1e-10), (f, fo)
    assert abs(omega - ref["sft_potential"]) < 1e-8 * max(1, abs(ref["sft_potential"]))


def test_chain_affine_batches_in_a_solve(monkeypatch):
    """The batched chain steps with each chain on one XCD (forced) give the
    same solve as with the columns spread over the XCDs: only the order of
    the alpha partial sums differs."""
    monkeypatch.setenv("VCA_CHAIN_SET_MAX_DIM", "0")  # one batch per target sector
    p, hloc, hk, _, _ = M.chain(nlat=10, u=4.0, nk=4)
    p.lanc_ngfiter = 80
    out = {}
    for xcd in ("0", "1"):
        monkeypatch.setenv("VCA_XCD_CHAINS", xcd)
        _say(f"10-site chain solve, xcd_chains={xcd}")
        s = VcaSolver(p)
        om = s.vca_solve(hloc, hk)
        out[xcd] = (om, s.vca_gf_cluster(0.3j), s.vca_gf_cluster(1.0 + 0.5j))
    assert abs(out["0"][0] - out["1"][0]) <= 1e-12 * abs(out["0"][0])
    for a, b in zip(out["0"][1:], out["1"][1:]):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(a).max()


def test_sft_exact_at_u0():
    """U=0: Omega(t') = (3-Nspin)/Nk sum_k sum_{e_kn<0} e_kn for every t'."""
    for tprime in (-1.0, -0.7):
        p, hloc, hk, _, _ = M.chain(nlat=4, u=0.0, nk=16, tprime=tprime)
        _say(f"u0 sft chain4 t'={tprime}")
        s = VcaSolver(p)
        omega = s.vca_solve(hloc, hk)
        hkl = R.hk_to_lso(M.to_oracle(p), hk)
        exact = (3 - p.nspin) * np.mean([np.sum(np.minimum(np.linalg.eigvalsh(h), 0)) for h in hkl])
        assert abs(omega - exact) < 1e-8 * abs(exact), (omega, exact)


def test_graph_replays_match_eager_launches(monkeypatch):
    """Chunks of Lanczos steps replayed from a captured graph (several
    replays, with eager work and syncs in between) give bit-identical
    alpha/beta to the same steps launched one by one."""
    p, hloc, hk, _, _ = M.chain(nlat=10, u=4.0)
    s = _solver(p, hloc)
    sec = s.sector(5, 5)
    v0 = _rand(sec.dim, 9)
    _say("graph vs eager, chain10 (5,5), 50 steps x 3 runs")
    monkeypatch.setenv("VCA_NO_GRAPHS", "1")
    a_e, b_e = sec.lanczos_tridiag(v0, 50)
    monkeypatch.setenv("VCA_NO_GRAPHS", "0")
    for _ in range(3):
        a_g, b_g = sec.lanczos_tridiag(v0, 50)
        assert np.array_equal(a_g, a_e) and np.array_equal(b_g, b_e)
        sec.hxv(v0)  # eager work between the replays


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("xcd", ["0", "1"])
def test_hxv_device_batch_equals_single(prec, xcd, monkeypatch):
    """vca_hxv_device_batch on nb stacked vectors gives, bit for bit, the nb
    single products (and the oracle's within the usual bar), with the
    columns of each chain spread over the XCDs or (xcd=1) each chain on one
    XCD, for a chain count that is not a multiple of 8."""
    from vca_amd.api import DeviceArray
    monkeypatch.setenv("VCA_XCD_CHAINS", xcd)
    p, hloc, hk, bh, bv = M.bath_model(nlat=5, nb=5)
    om = M.oracle_model(p, hloc, bh, bv)
    s = _solver(p, hloc, bh, bv)
    sec = s.sector(5, 5, prec)
    nb = 13
    _say(f"batched hxv bath5+5 (5,5) nb={nb} {prec} xcd_chains={xcd}")
    dt = np.complex128 if prec == "fp64" else np.complex64
    rng = np.random.default_rng(4)
    v = (rng.normal(size=(nb, sec.dim)) + 1j * rng.normal(size=(nb, sec.dim))).astype(dt)
    dv = DeviceArray(v.shape, dt).upload(v)
    dh = DeviceArray(v.shape, dt)
    sec.hxv_device_batch(nb, dv.ptr, dh.ptr)
    hv = dh.download()
    d1 = DeviceArray((sec.dim,), dt)
    h1 = DeviceArray((sec.dim,), dt)
    for c in range(nb):
        d1.upload(v[c])
        sec.hxv_device(d1.ptr, h1.ptr)
        assert np.array_equal(h1.download(), hv[c])
    ref = om.hxv(5, 5, v[1].astype(np.complex128))
    err = np.linalg.norm(hv[1] - ref) / np.linalg.norm(ref)
    assert err < (1e-12 if prec == "fp64" else 2e-6), err


@pytest.mark.parametrize("variant", ["24", "25", "26"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_signed_column_form_in_lanczos_and_batches(prec, variant, monkeypatch):
    """The signed-column overlapped H x v (one real hopping amplitude: the
    10-site chain; the column staged as +t x and -t x, no amplitude gather)
    inside the Lanczos steps, where alpha's partials reread the input column:
    the oracle's Lanczos coefficients (12 steps; 1e-10 FP64, 1e-5 MIXED), and
    stacked products equal to the single ones bit for bit."""
    from vca_amd.api import DeviceArray
    monkeypatch.setenv("VCA_HXV_VARIANT", variant)
    p, hloc, hk, bh, bv = M.chain(nlat=10, u=4.0)
    om = M.oracle_model(p, hloc, bh, bv)
    s = _solver(p, hloc, bh, bv)
    sec = s.sector(5, 5, prec)
    _say(f"signed-column form {variant}: tridiag chain10 (5,5) {prec}")
    v0 = _rand(sec.dim, 5)
    v0 /= np.linalg.norm(v0)
    a, b = sec.lanczos_tridiag(v0, 12)
    ao, bo = R.lanczos_tridiag(lambda x: om.hxv(5, 5, x), v0, 12)
    tol = 1e-10 if prec == "fp64" else 1e-5
    assert np.allclose(a, ao, rtol=tol, atol=tol) and np.allclose(b, bo, rtol=tol, atol=tol)
    _say(f"signed-column form {variant}: batched products chain10 (5,5) nb=13 {prec}")
    dt = np.complex128 if prec == "fp64" else np.complex64
    nb = 13
    rng = np.random.default_rng(8)
    v = (rng.normal(size=(nb, sec.dim)) + 1j * rng.normal(size=(nb, sec.dim))).astype(dt)
    dv = DeviceArray(v.shape, dt).upload(v)
    dh = DeviceArray(v.shape, dt)
    sec.hxv_device_batch(nb, dv.ptr, dh.ptr)
    hv = dh.download()
    d1 = DeviceArray((sec.dim,), dt)
    h1 = DeviceArray((sec.dim,), dt)
    for c in (0, 6, 12):
        d1.upload(v[c])
        sec.hxv_device(d1.ptr, h1.ptr)
        assert np.array_equal(h1.download(), hv[c])
        ref = om.hxv(5, 5, v[c].astype(np.complex128))
        err = np.linalg.norm(hv[c] - ref) / np.linalg.norm(ref)
        assert err < (1e-12 if prec == "fp64" else 2e-6), (c, err)
    sec.close()


@pytest.mark.parametrize("variant", ["16", "20", "27"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_transposed_form_in_lanczos_and_batches(prec, variant, monkeypatch):
    """The transposed H x v form (forced on small sectors; variant 16 with
    whole complex columns, 20 one real plane at a time, 27 whole complex
    columns through the column-pair tables) inside the Lanczos
    steps: its two gather passes' partials of alpha, the buffer
    parity and the skipping of finished chains.  Bars: the oracle's Lanczos
    coefficients (12 steps, bath model); stacked products equal to the
    single ones bit for bit; a whole solve (GS, batched chains) equal to the
    direct form's within rounding."""
    from vca_amd.api import DeviceArray
    monkeypatch.setenv("VCA_TFORM_MIN_DIM", "0")
    monkeypatch.setenv("VCA_HXV_VARIANT", variant)
    p, hloc, hk, bh, bv = M.bath_model(nlat=5, nb=5)
    om = M.oracle_model(p, hloc, bh, bv)
    s = _solver(p, hloc, bh, bv)
    sec = s.sector(5, 5, prec)
    _say(f"transposed form: tridiag bath5+5 (5,5) {prec}")
    v0 = _rand(sec.dim, 5)
    v0 /= np.linalg.norm(v0)
    a, b = sec.lanczos_tridiag(v0, 12)
    ao, bo = R.lanczos_tridiag(lambda x: om.hxv(5, 5, x), v0, 12)
    tol = 1e-10 if prec == "fp64" else 1e-5
    assert np.allclose(a, ao, rtol=tol, atol=tol) and np.allclose(b, bo, rtol=tol, atol=tol)
    _say(f"transposed form: batched products bath5+5 (5,5) nb=13 {prec}")
    dt = np.complex128 if prec == "fp64" else np.complex64
    nb = 13
    rng = np.random.default_rng(8)
    v = (rng.normal(size=(nb, sec.dim)) + 1j * rng.normal(size=(nb, sec.dim))).astype(dt)
    dv = DeviceArray(v.shape, dt).upload(v)
    dh = DeviceArray(v.shape, dt)
    sec.hxv_device_batch(nb, dv.ptr, dh.ptr)
    hv = dh.download()
    d1 = DeviceArray((sec.dim,), dt)
    h1 = DeviceArray((sec.dim,), dt)
    for c in (0, 6, 12):
        d1.upload(v[c])
        sec.hxv_device(d1.ptr, h1.ptr)
        assert np.array_equal(h1.download(), hv[c])
    sec.close()
    if prec != "fp64":
        return
    monkeypatch.setenv("VCA_CHAIN_SET_MAX_DIM", "0")  # chains on the sector's own H x v
    pc, hlc, hkc, _, _ = M.chain(nlat=10, u=4.0, nk=4)
    pc.lanc_ngfiter = 80
    out = {}
    for form, var in (("transposed", variant), ("direct", "13")):
        monkeypatch.setenv("VCA_HXV_VARIANT", var)
        _say(f"10-site chain solve, {form} H x v")
        sc = VcaSolver(pc)
        om_ = sc.vca_solve(hlc, hkc)
        out[form] = (om_, sc.vca_gf_cluster(0.3j), sc.vca_gf_cluster(1.0 + 0.5j))
    assert abs(out["transposed"][0] - out["direct"][0]) <= 1e-11 * abs(out["direct"][0])
    for x, y in zip(out["transposed"][1:], out["direct"][1:]):
        assert np.abs(x - y).max() <= 1e-9 * np.abs(y).max()


def test_complex_pair_passes_equal_whole_column_passes(monkeypatch):
    """The transposed form's complex-pair gather passes (variants 27, 28)
    and its whole-column passes (variant 16) sum the same hops in the same
    order (the complex64 table takes the bank-aware hop order of the
    one-plane table the pair table is made from): MIXED products equal bit
    for bit; FP64 (whose complex128 table keeps the original order) within
    1e-14 of each other; the 10-site chain and the bath model, 3 stacked
    vectors."""
    from vca_amd.api import DeviceArray
    monkeypatch.setenv("VCA_TFORM_MIN_DIM", "0")
    for name, (p, hloc, hk, bh, bv) in (("chain10", M.chain(nlat=10, u=4.0)),
                                        ("bath5+5", M.bath_model(nlat=5, nb=5))):
        s = _solver(p, hloc, bh, bv)
        for prec in ("fp64", "mixed"):
            dt = np.complex128 if prec == "fp64" else np.complex64
            out = {}
            for var in ("16", "27", "28"):
                monkeypatch.setenv("VCA_HXV_VARIANT", var)
                _say(f"complex-pair passes: {name} (5,5) {prec} variant {var}")
                sec = s.sector(5, 5, prec)
                rng = np.random.default_rng(21)
                v = (rng.normal(size=(3, sec.dim)) + 1j * rng.normal(size=(3, sec.dim))).astype(dt)
                dv = DeviceArray(v.shape, dt).upload(v)
                dh = DeviceArray(v.shape, dt)
                sec.hxv_device_batch(3, dv.ptr, dh.ptr)
                out[var] = dh.download()
                sec.close()
            for var in ("27", "28"):
                if prec == "mixed":
                    assert np.array_equal(out[var], out["16"]), (name, var)
                else:
                    err = np.abs(out[var] - out["16"]).max() / np.abs(out["16"]).max()
                    assert err < 1e-14, (name, var, err)


@pytest.mark.parametrize("form", ["direct", "transposed2"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_real_vector_ground_state(prec, form, monkeypatch):
    """Real Hamiltonians: the ground-state Lanczos on real vectors (forced on
    small sectors by lowering the transposed form's size threshold) against
    the same run on complex vectors: E0 within 1e-10 relative of each other,
    the eigenvector a normalised, real eigenvector of the oracle's H
    (residual 1e-5).  Sectors: 10-site chain
    (5,5), bath model (5,5), and the 11-site chain (3,3) whose 27,225 states
    leave the pad real of the last complex pair.  Direct and transposed
    real-vector H x v."""
    _set_env(monkeypatch, REAL_FORMS[form])
    monkeypatch.setenv("VCA_TFORM_MIN_DIM", "0")
    for name, (p, hloc, hk, bh, bv), sector in (
            ("chain10", M.chain(nlat=10, u=4.0), (5, 5)),
            ("bath5+5", M.bath_model(nlat=5, nb=5), (5, 5)),
            ("chain11", M.chain(nlat=11, u=4.0), (3, 3))):
        om = M.oracle_model(p, hloc, bh, bv)
        s = _solver(p, hloc, bh, bv)
        out = {}
        for real in ("1", "0"):
            monkeypatch.setenv("VCA_REAL_GS", real)
            _say(f"ground state {name} {sector} {prec} real={real}")
            sec = s.sector(*sector, prec)
            out[real] = sec.lanczos_gs(tol=1e-12, niter=400)
            sec.close()
        (e_r, v_r, _), (e_c, _, _) = out["1"], out["0"]
        assert abs(e_r - e_c) < 1e-10 * abs(e_c), (name, e_r, e_c)
        assert np.abs(v_r.imag).max() == 0.0
        assert abs(np.linalg.norm(v_r) - 1) < 1e-10
        r = om.hxv(*sector, v_r) - e_r * v_r
        assert np.linalg.norm(r) < 1e-5, (name, np.linalg.norm(r))


@pytest.mark.parametrize("rtol", [None, "1e-11"])
def test_real_vector_ground_state_in_a_solve(rtol, monkeypatch):
    """A whole solve whose ground state runs on real vectors, against the
    solve on complex vectors (the oracle's own solve of this 10-site cluster
    takes over 10 minutes on one core; the pipeline tests pin smaller ones
    to it).  The two ground-state runs start from different vectors, so
    under the Lanczos stopping rule |dE| < 1e-12 alone they agree to what
    that rule fixes, not bit for bit: E0 within 1e-10 relative, Omega within
    1e-8 (1.1e-10 seen), and G', first order in the ground-state vector that
    the rule fixes to ~1e-6 (the eigenvector error is ~sqrt of the energy
    error), within 2e-6 of its largest element (5.9e-7 seen).  With the
    Ritz-residual stop as well (VCA_GS_RESIDUAL=1e-11) both runs converge
    the vector, and they agree to E0 1e-13, Omega 1e-12 and G' 1e-9: the
    2e-6 is the stopping rule's, not the real vectors'.  The chains
    themselves add only rounding (test_real_vector_chains_in_a_solve:
    1e-10).  The GPU's complex path matches the oracle to 1e-8 in G' because
    both follow the same start vector."""
    monkeypatch.setenv("VCA_TFORM_MIN_DIM", "0")
    if rtol is None:
        monkeypatch.delenv("VCA_GS_RESIDUAL", raising=False)
    else:
        monkeypatch.setenv("VCA_GS_RESIDUAL", rtol)
    p, hloc, hk, _, _ = M.chain(nlat=10, u=4.0, nk=4)
    p.lanc_ngfiter = 80
    out = {}
    for real in ("1", "0"):
        monkeypatch.setenv("VCA_REAL_GS", real)
        _say(f"10-site chain solve, real-vector ground state={real}")
        sc = VcaSolver(p)
        om_ = sc.vca_solve(hloc, hk)
        out[real] = (om_, sc.results()["omegaprime"], sc.vca_gf_cluster(0.3j),
                     sc.vca_gf_cluster(1.0 + 0.5j))
    e_bar, om_bar, g_bar = (1e-10, 1e-8, 2e-6) if rtol is None else (1e-13, 1e-12, 1e-9)
    assert abs(out["1"][1] - out["0"][1]) <= e_bar * abs(out["0"][1])
    assert abs(out["1"][0] - out["0"][0]) <= om_bar * abs(out["0"][0])
    for x, y in zip(out["1"][2:], out["0"][2:]):
        err = np.abs(x - y).max() / np.abs(y).max()
        _say(f"  G' real against complex ground state: {err:.1e}")
        assert err <= g_bar


REAL_FORMS = {
    # the direct form (k_hxv_buf_real): small sectors keep the default shape
    "direct": {},
    # the transposed form on one real plane, two columns per workgroup
    # (k_hxv_plane2) or one (k_hxv_plane)
    "transposed2": {"VCA_TFORM_MIN_DIM": "0", "VCA_HXV_VARIANT": "20", "VCA_PLANE2": "1"},
    "transposed1": {"VCA_TFORM_MIN_DIM": "0", "VCA_HXV_VARIANT": "20", "VCA_PLANE2": "0"},
    # the overlapped direct form (k_hxv_ov_real, 512 lanes x 2 rows)
    "overlapped": {"VCA_REAL_VARIANT": "9"},
}

# launch shapes of the real direct form: (lanes, rows or slots per lane,
# kind): 0 single rows, 1 row pairs, 2 overlapped, 3 overlapped row pairs
REAL_SHAPES = [(256, 4, 0), (256, 2, 0), (512, 2, 0), (512, 1, 0), (1024, 1, 0), (1024, 2, 0),
               (256, 1, 1), (512, 1, 1), (1024, 1, 1), (512, 2, 2), (1024, 1, 2), (256, 4, 2),
               (512, 1, 3), (256, 2, 3), (1024, 1, 3), (128, 8, 2), (128, 4, 3)]


def _real_shape_fits(var, dimup):
    bs, nr, kind = REAL_SHAPES[var]
    if kind in (1, 3) and dimup % 2:
        return False
    return kind < 2 or dimup <= bs * nr * (2 if kind == 3 else 1)


@pytest.mark.parametrize("var", range(len(REAL_SHAPES)))
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_real_direct_shapes(prec, var, monkeypatch):
    """Every launch shape of the real-vector direct form (VCA_REAL_VARIANT)
    gives the oracle's H x v (1e-12 / 2e-6 relative) on the sectors it fits:
    an even DimUp of 252 and an odd one of 165; a shape that does not fit is
    refused with VCA_ERR_ARG."""
    from vca_amd.api import DeviceArray, VcaError
    monkeypatch.setenv("VCA_REAL_VARIANT", str(var))
    dt = np.float64 if prec == "fp64" else np.float32
    for name, (p, hloc, hk, bh, bv), sector in (
            ("bath5+5", M.bath_model(nlat=5, nb=5), (5, 5)),
            ("chain11", M.chain(nlat=11, u=4.0), (3, 3))):
        s = _solver(p, hloc, bh, bv)
        sec = s.sector(*sector, prec)
        dimup = sec.dimup
        v = np.random.default_rng(5).normal(size=sec.dim).astype(dt)
        dv = DeviceArray(v.shape, dt).upload(v)
        dh = DeviceArray(v.shape, dt)
        _say(f"real H x v shape {var} {name} {sector} {prec} dimup {dimup}")
        if not _real_shape_fits(var, dimup):
            with pytest.raises(VcaError) as ei:
                sec.hxv_device_real(dv.ptr, dh.ptr)
            assert ei.value.code == -1
            sec.close()
            continue
        sec.hxv_device_real(dv.ptr, dh.ptr)
        om = M.oracle_model(p, hloc, bh, bv)
        ref = om.hxv(*sector, v.astype(np.complex128)).real
        err = np.linalg.norm(dh.download() - ref) / np.linalg.norm(ref)
        assert err < (1e-12 if prec == "fp64" else 2e-6), (name, var, err)
        sec.close()


def _set_env(monkeypatch, env):
    for k, v in env.items():
        monkeypatch.setenv(k, v)


@pytest.mark.parametrize("form", sorted(REAL_FORMS))
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_hxv_device_real(prec, form, monkeypatch):
    """vca_hxv_device_real on real vectors gives the oracle's H x v (1e-12 /
    2e-6 relative) in each real-vector form, on a sector with an even and
    one with an odd number of columns; a sector with complex amplitudes is
    refused with VCA_ERR_ARG."""
    from vca_amd.api import DeviceArray, VcaError
    _set_env(monkeypatch, REAL_FORMS[form])
    dt = np.float64 if prec == "fp64" else np.float32
    for name, (p, hloc, hk, bh, bv), sector in (
            ("bath5+5", M.bath_model(nlat=5, nb=5), (5, 5)),
            ("chain11", M.chain(nlat=11, u=4.0), (3, 3))):
        om = M.oracle_model(p, hloc, bh, bv)
        s = _solver(p, hloc, bh, bv)
        sec = s.sector(*sector, prec)
        _say(f"real H x v {name} {sector} {prec}")
        v = np.random.default_rng(3).normal(size=sec.dim).astype(dt)
        dv = DeviceArray(v.shape, dt).upload(v)
        dh = DeviceArray(v.shape, dt)
        sec.hxv_device_real(dv.ptr, dh.ptr)
        ref = om.hxv(*sector, v.astype(np.complex128))
        assert np.abs(ref.imag).max() == 0.0
        err = np.linalg.norm(dh.download() - ref.real) / np.linalg.norm(ref.real)
        assert err < (1e-12 if prec == "fp64" else 2e-6), (name, err)
        sec.close()
    bhz = VcaParams(nlat=5, norb=2, nspin=2, uloc=(2.0, 2.0, 0, 0, 0))
    bhz_h, _ = configs.bhz(5, 1, 2.0, 0.3, nk=(2, 1))
    s = _solver(bhz, bhz_h)
    sec = s.sector(5, 4, prec)
    d = DeviceArray((sec.dim,), dt)
    with pytest.raises(VcaError) as ei:
        sec.hxv_device_real(d.ptr, d.ptr)
    assert ei.value.code == -1


@pytest.mark.parametrize("form", sorted(REAL_FORMS))
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_real_vector_chain_steps(prec, form, monkeypatch):
    """A continued-fraction chain from a real start vector on a real
    Hamiltonian runs on real vectors; its coefficients equal the oracle's
    Lanczos recurrence (1e-10 FP64, 1e-5 MIXED) and those of the same chain
    on complex vectors, on an even and an odd sector."""
    _set_env(monkeypatch, REAL_FORMS[form])
    for name, (p, hloc, hk, bh, bv), sector in (
            ("bath5+5", M.bath_model(nlat=5, nb=5), (5, 5)),
            ("chain11", M.chain(nlat=11, u=4.0), (3, 3))):
        om = M.oracle_model(p, hloc, bh, bv)
        s = _solver(p, hloc, bh, bv)
        sec = s.sector(*sector, prec)
        _say(f"real chain {name} {sector} {prec} {form}")
        v0 = np.random.default_rng(11).normal(size=sec.dim) + 0j
        v0 /= np.linalg.norm(v0)
        a, b = sec.lanczos_tridiag(v0, 20)
        ao, bo = R.lanczos_tridiag(lambda x: om.hxv(*sector, x), v0, 20)
        tol = 1e-10 if prec == "fp64" else 1e-5
        assert np.allclose(a, ao, rtol=tol, atol=tol), (name, np.abs(a - ao).max())
        assert np.allclose(b, bo, rtol=tol, atol=tol), (name, np.abs(b - bo).max())
        monkeypatch.setenv("VCA_REAL_CHAINS", "0")
        ac, bc = sec.lanczos_tridiag(v0, 20)
        monkeypatch.delenv("VCA_REAL_CHAINS")
        assert np.allclose(a, ac, rtol=tol, atol=tol) and np.allclose(b, bc, rtol=tol, atol=tol)
        sec.close()


@pytest.mark.parametrize("case", ["chain10_xcd0", "chain10_xcd1", "chain11_sector33",
                                  "chain10_overlapped"])
def test_real_vector_chains_in_a_solve(case, monkeypatch):
    """The excitation chains of a real Hamiltonian on real vectors (the
    ground state's global phase removed), batched with the columns of each
    chain spread over the XCDs or each chain on one XCD, against the same
    solve on complex chains.  chain11_sector33: the 11-site chain restricted
    to the sector (3,3), whose target sector (3,2) has an odd dimension
    (165 x 55 = 9,075: the pad real of every chain in the batch).
    chain10_overlapped: the real batches on the overlapped direct form
    (k_hxv_ov_real, 512 lanes x 2 rows), each chain on one XCD.  The
    ground state is the same complex vector in both, so E0 agrees to
    rounding.  With psi = e^{i phi}(r + i s), a real H and an operator O of
    real coefficients, <O psi|R(z)|O psi> = <Or|R|Or> + <Os|R|Os> exactly
    (R = (z - H)^-1 is complex-symmetric, so the cross terms cancel): the
    real chains, which drop s, differ by a term second order in the state's
    imaginary remainder, and otherwise by rounding, which 80-step chains do
    not amplify (no loss of orthogonality yet).  Bars 1e-10 of G''s largest
    element and 1e-12 relative in Omega (up to 6e-14 and 2e-14 seen)."""
    monkeypatch.setenv("VCA_CHAIN_SET_MAX_DIM", "0")  # one batch per target sector
    if case == "chain11_sector33":
        p, hloc, hk, _, _ = M.chain(nlat=11, u=4.0, nk=4)
        p.sectors = [(3, 3)]
    else:
        if case == "chain10_overlapped":
            monkeypatch.setenv("VCA_XCD_CHAINS", "1")
            monkeypatch.setenv("VCA_REAL_VARIANT", "9")
        else:
            monkeypatch.setenv("VCA_XCD_CHAINS", case[-1])
        p, hloc, hk, _, _ = M.chain(nlat=10, u=4.0, nk=4)
    p.lanc_ngfiter = 80
    out = {}
    for real in ("1", "0"):
        monkeypatch.setenv("VCA_REAL_CHAINS", real)
        _say(f"{case} solve, real-vector chains={real}")
        sc = VcaSolver(p)
        om_ = sc.vca_solve(hloc, hk)
        out[real] = (om_, sc.results()["omegaprime"], sc.vca_gf_cluster(0.3j),
                     sc.vca_gf_cluster(1.0 + 0.5j))
    assert abs(out["1"][1] - out["0"][1]) <= 1e-12 * abs(out["0"][1])
    print(f"{case}: dOmega {abs(out['1'][0] - out['0'][0]) / abs(out['0'][0]):.2e}, dG' " +
          ", ".join(f"{np.abs(x - y).max() / np.abs(y).max():.2e}"
                    for x, y in zip(out["1"][2:], out["0"][2:])), flush=True)
    assert abs(out["1"][0] - out["0"][0]) <= 1e-12 * abs(out["0"][0])
    for x, y in zip(out["1"][2:], out["0"][2:]):
        assert np.abs(x - y).max() <= 1e-10 * np.abs(y).max(), np.abs(x - y).max()
