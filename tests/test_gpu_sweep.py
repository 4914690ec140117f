"""The concurrent sector sweep of the diagonalisation (every (Nup, Ndw)
sector solved on worker threads, results taken in list order).

Bars: the sweep on 16 workers gives the same state list, Omega and G' as
the same sweep one sector at a time (1e-12: each sector's solve does not
depend on which thread runs it); an error in a worker reaches the caller
with its code and message.
"""
import numpy as np
import pytest

from vca_amd import configs
from vca_amd.api import VcaError, VcaParams, VcaSolver

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _cluster_2x4(**kw):
    """8-site 2x4 cluster, Nspin=2, U=4, a small staggered field: every
    sector (dims 1..4,900), ARPACK-equivalent eigenpairs."""
    onsite = np.zeros((8, 2))
    for iy in range(4):
        for ix in range(2):
            i = configs.site(ix, iy, 2)
            stag = 0.1 if (ix + iy) % 2 == 0 else -0.1
            onsite[i, 0], onsite[i, 1] = -stag, stag
    kw.setdefault("lanc_method", "arpack")
    p = VcaParams(nlat=8, nspin=2, uloc=(4.0, 0, 0, 0, 0), lanc_ngfiter=60, **kw)
    hloc, _ = configs.square_lattice(2, 4, 2, -1.0, onsite, nk=(1, 1))
    _, hk = configs.square_lattice(2, 4, 2, -1.0, np.zeros((8, 2)), nk=(4, 4))
    return p, hloc, hk


def _solve(monkeypatch, threads, **kw):
    monkeypatch.setenv("VCA_SECTOR_THREADS", str(threads))
    p, hloc, hk = _cluster_2x4(**kw)
    s = VcaSolver(p)
    om = s.vca_solve(hloc, hk)
    return s, om, s.results()


@pytest.mark.parametrize("mode", ["arpack", "arpack_finite_t", "lanczos"])
def test_concurrent_sweep_equals_sequential(monkeypatch, mode):
    """lanczos: one state per sector by the plain Lanczos, whose steps run
    from captured graphs on the calling thread and eagerly on the workers
    (graph capture on one thread would make allocations on the others fail)."""
    kw = {"arpack": {}, "arpack_finite_t": dict(beta=20.0, lanc_nstates_total=4, bandwidth=6.0),
          "lanczos": dict(lanc_method="lanczos")}[mode]
    _say(f"sweep 2x4 every sector {mode}: 1 worker")
    s1, om1, r1 = _solve(monkeypatch, 1, **kw)
    _say(f"sweep 2x4 every sector {mode}: 16 workers")
    s16, om16, r16 = _solve(monkeypatch, 16, **kw)
    assert r1["nstates"] == r16["nstates"] and r1["spectrum_iters"] == r16["spectrum_iters"]
    assert abs(r1["omegaprime"] - r16["omegaprime"]) <= 1e-12 * abs(r1["omegaprime"])
    assert abs(om1 - om16) <= 1e-12 * abs(om1)
    for z in (0.4j, 1.1 + 0.3j):
        g1, g16 = s1.vca_gf_cluster(z), s16.vca_gf_cluster(z)
        assert np.abs(g1 - g16).max() <= 1e-11 * np.abs(g1).max()


def test_worker_error_reaches_the_caller(monkeypatch):
    """One restart and 20 steps cannot converge 6 eigenpairs of the larger
    sectors: the worker's VCA_ERR_NOCONV and its message come back."""
    _say("sweep 2x4 with an unconvergeable eigensolve, 16 workers")
    with pytest.raises(VcaError) as ei:
        _solve(monkeypatch, 16, lanc_niter=1, lanc_ncv_factor=1, lanc_ncv_add=14)
    assert ei.value.code == -5
    assert "sector" in str(ei.value) and "converged" in str(ei.value)
