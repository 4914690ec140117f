"""CPU checks of vca_lanczos_of_sum, the host routine from which vca_solve
takes the (c^+_i + i c^+_j) and (c_i - i c_j) channels of a real Hamiltonian
(VCA_GF_NORMAL.f90:121-318, 561-958) out of the two diagonal chains instead
of running them.

The bar is the chain the reference runs: plain Lanczos from the complex
start vector a - i b under the same real H (the oracle's
lanczos_tridiag on the 6-site chain's (4,3) sector, 300 states, and on the
bath model's (4,3) sector), against the routine fed the oracle's chains of
a and b.  The identity behind it -- the measure of a - i b is the sum of
those of a and b, and n Lanczos steps read only the first 2n moments, which
each n-node Gauss quadrature keeps -- makes the coefficients equal up to
rounding: 1e-10 relative over 24 steps, and G(z) of the two tridiagonals
within 1e-12 of each other.  (The sum of the two n-step fractions, a
2n-node quadrature, is another approximation of the same function: near
the real axis it differs from any n-step chain by the truncation.)  Plain
Lanczos loses orthogonality within a few dozen steps on these sectors, so
the steps are kept short."""
import numpy as np
import pytest

from oracle import vca_ref as R
from vca_amd import api

from . import _models as M


def _cf(alpha, beta, nrm, z):
    """G(z) = nrm^2 <e0|(z - T)^-1|e0> of the tridiagonal (alpha, beta)."""
    w, v = R.tridiag_eigh(alpha, beta)
    return nrm * nrm * np.sum(v[0, :] ** 2 / (z - w))


def _sector_h(model):
    p, hloc, _, bh, bv = model
    om = M.oracle_model(p, hloc, bh, bv)
    h = om.dense_h(4, 3)
    assert np.abs(h.imag).max() == 0.0
    return h.real


@pytest.mark.parametrize("name", ["chain6", "bath3+3"])
def test_chain_of_a_minus_ib_from_the_chains_of_a_and_b(name):
    model = M.chain(nlat=6) if name == "chain6" else M.bath_model(nlat=3, nb=3)
    h = _sector_h(model)
    dim = h.shape[0]
    rng = np.random.default_rng(3)
    a = rng.normal(size=dim)
    b = rng.normal(size=dim) * 0.7
    n = 24
    mv = lambda x: h @ x  # noqa: E731
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    aa, ba = R.lanczos_tridiag(mv, (a / na).astype(np.complex128), n)
    ab, bb = R.lanczos_tridiag(mv, (b / nb).astype(np.complex128), n)
    c = (a - 1j * b)
    nc = np.linalg.norm(c)
    ac, bc = R.lanczos_tridiag(mv, c / nc, n)
    ad, bd, nd = api.lanczos_of_sum(aa, ba, na, ab, bb, nb)
    scale = np.abs(np.linalg.eigvalsh(h)).max()
    assert abs(nd - nc) <= 1e-14 * nc
    assert bd[0] == 0.0
    assert np.abs(ad - ac).max() <= 1e-10 * scale, np.abs(ad - ac).max() / scale
    assert np.abs(bd - bc).max() <= 1e-10 * scale, np.abs(bd - bc).max() / scale
    for z in (0.4j, 2.0j, 1.5 + 0.3j, -3.0 + 1.0j):
        gd = _cf(ad, bd, nd, z)
        gc = _cf(ac, bc, nc, z)
        assert abs(gd - gc) <= 1e-12 * abs(gc), (z, abs(gd - gc) / abs(gc))


def test_one_empty_chain_and_a_whole_space():
    """A chain of norm 0 (an empty channel) leaves the other one's
    coefficients; chains as long as the space (12 states, 12 steps) give the
    exact measure, and so does their sum."""
    rng = np.random.default_rng(5)
    x = rng.normal(size=(12, 12))
    h = x + x.T
    mv = lambda v: h @ v  # noqa: E731
    a, b = rng.normal(size=12), rng.normal(size=12)
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    aa, ba = R.lanczos_tridiag(mv, (a / na).astype(np.complex128), 12)
    ab, bb = R.lanczos_tridiag(mv, (b / nb).astype(np.complex128), 12)
    ad, bd, nd = api.lanczos_of_sum(aa, ba, na, ab, bb, 0.0)
    assert nd == pytest.approx(na, rel=1e-15)
    assert np.allclose(ad, aa, rtol=0, atol=1e-11) and np.allclose(bd, ba, rtol=0, atol=1e-11)
    ad, bd, nd = api.lanczos_of_sum(aa, ba, na, ab, bb, nb)
    c = a - 1j * b
    ev, u = np.linalg.eigh(h)
    for z in (0.3j, 1.0 + 0.5j):
        exact = np.sum(np.abs(u.T @ c) ** 2 / (z - ev))
        assert abs(_cf(ad, bd, nd, z) - exact) <= 1e-11 * abs(exact)
    ad, bd, nd = api.lanczos_of_sum(aa, ba, 0.0, ab, bb, 0.0)
    assert nd == 0.0 and not ad.any() and not bd.any()
