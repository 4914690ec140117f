"""Input generators: Hermiticity, shapes and the reduced-zone convention.

The cluster H_k on the reduced zone must reproduce the lattice band: for a
chain of L sites with hopping t, the eigenvalues of H(k) over the Nk cluster
momenta are 2t cos(q) at q = k + 2 pi m / L; for an Lx x Ly square cluster,
2t (cos qx + cos qy).  This pins the k grid, the boundary phases e^{+-ikL}
and the site indexing N = Nx (iy-1) + ix."""
import numpy as np

from oracle import vca_ref as R
from vca_amd import configs

from . import _models as M


def _lso(p, h):
    return R.nnn2lso(h, p.nlat, p.nspin, p.norb)


def test_chain_band():
    L, nk, t = 4, 8, -1.0
    hloc, hk = configs.chain(L, t, t, nk)
    ev = np.sort(np.concatenate([np.linalg.eigvalsh(hk[:, :, 0, 0, 0, 0, i]) for i in range(nk)]))
    q = 2 * np.pi * np.arange(nk * L) / (nk * L)
    assert np.allclose(ev, np.sort(2 * t * np.cos(q)), atol=1e-12)


def test_square_band():
    nx, ny, nk, t = 3, 2, 4, -1.0
    hloc, hk = configs.square_lattice(nx, ny, 1, t, np.zeros((nx * ny, 1)), nk=(nk, nk))
    ev = np.sort(np.concatenate([np.linalg.eigvalsh(hk[..., 0, 0, 0, 0, i])
                                 for i in range(hk.shape[6])]))
    qx = 2 * np.pi * np.arange(nk * nx) / (nk * nx)
    qy = 2 * np.pi * np.arange(nk * ny) / (nk * ny)
    exact = np.sort((2 * t * (np.cos(qx)[:, None] + np.cos(qy)[None, :])).ravel())
    assert np.allclose(ev, exact, atol=1e-12)


def test_hermitian_inputs():
    for f in (configs.config_chain4, configs.config_square_3x4, configs.config_batched_2x5,
              configs.config_bath_2x4, configs.config_bhz_2x2):
        p, hloc, hk, bh, bv = f(nk=2)
        h = _lso(p, hloc)
        assert np.abs(h - h.conj().T).max() < 1e-14
        for i in range(hk.shape[6]):
            m = _lso(p, hk[..., i])
            assert np.abs(m - m.conj().T).max() < 1e-14
        if bh is not None:
            b = bh[:, :, 0, 0, 0, 0]
            assert np.abs(b - b.conj().T).max() < 1e-14
            assert bv.shape == (p.nlat, p.nlat_bath, 1, 1, 1, 1)


def test_cluster_equals_lattice_at_zero_boundary():
    """Hloc equals the intra-cluster part of H_k (what remains when the
    inter-cluster terms are removed), so V_k = H_k - Hloc is pure boundary."""
    p, hloc, hk, _, _ = configs.config_square_3x4(nk=2)
    hl = _lso(p, hloc)
    v0 = _lso(p, hk[..., 0]) - hl
    # V_k connects only boundary sites (ix = 0 / Nx-1 or iy = 0 / Ny-1)
    nx, ny = 3, 4
    inner = [configs.site(ix, iy, nx) for ix in range(1, nx - 1) for iy in range(1, ny - 1)]
    assert np.abs(v0[inner, :]).max() == 0 and np.abs(v0[:, inner]).max() == 0


def test_sector_dimensions_of_the_workloads():
    import math
    assert math.comb(12, 6) ** 2 == 853776
    assert math.comb(16, 8) ** 2 == 165636900
    assert math.comb(10, 5) ** 2 == 63504
    p, *_ = configs.config_bhz_2x2(nk=2)
    assert p.ns == 8 and math.comb(8, 4) ** 2 == 4900


def test_oracle_params_mapping():
    p, *_ = configs.config_bath_2x4(nk=2)
    po = M.to_oracle(p)
    assert (po.nlat, po.nlat_bath, po.norb_bath, po.ns) == (8, 8, 1, 16)
    assert po.uloc[0] == 4.0 and po.sectors == [(8, 8)]
