"""Small models shared by the tests: each returns (VcaParams, hloc, hk,
bath_h, bath_v) and the matching oracle Params."""
from __future__ import annotations

import numpy as np

from oracle import vca_ref as R
from vca_amd import configs
from vca_amd.api import VcaParams


def to_oracle(p: VcaParams) -> R.Params:
    return R.Params(nlat=p.nlat, norb=p.norb, nspin=p.nspin, nlat_bath=p.nlat_bath,
                    norb_bath=p.norb_bath, uloc=tuple(p.uloc), ust=p.ust, jh=p.jh, xmu=p.xmu,
                    beta=p.beta, hfmode=p.hfmode, hfshift=p.hfshift,
                    gf_symmetric=p.gf_symmetric, lanc_method=p.lanc_method,
                    lanc_nstates_sector=p.lanc_nstates_sector, lanc_niter=p.lanc_niter,
                    lanc_ngfiter=p.lanc_ngfiter, lanc_tolerance=p.lanc_tolerance,
                    lanc_dim_threshold=p.lanc_dim_threshold, gs_threshold=p.gs_threshold,
                    sectors=p.sectors, lanc_nstates_total=p.lanc_nstates_total,
                    lanc_nstates_step=p.lanc_nstates_step,
                    lanc_niter_spectrum=p.lanc_niter_spectrum,
                    lanc_ncv_factor=p.lanc_ncv_factor, lanc_ncv_add=p.lanc_ncv_add,
                    cutoff=p.cutoff, bandwidth=p.bandwidth, lmats=p.lmats, lreal=p.lreal,
                    wini=p.wini, wfin=p.wfin, eps=p.eps)


def chain(nlat=6, u=4.0, nk=8, tprime=-1.0):
    p = VcaParams(nlat=nlat, uloc=(u, 0, 0, 0, 0), lanc_method="lanczos")
    hloc, hk = configs.chain(nlat, tprime, -1.0, nk)
    return p, hloc, hk, None, None


def multiorb(seed=7):
    """2 sites x 2 orbitals, Nspin=2 with spin-dependent levels, Ust, Jh,
    hfmode, complex inter-orbital hopping."""
    rng = np.random.default_rng(seed)
    p = VcaParams(nlat=2, norb=2, nspin=2, uloc=(3.0, 2.5, 0, 0, 0), ust=1.2, jh=0.3, xmu=0.1,
                  lanc_method="lanczos")
    h = np.zeros((2, 2, 2, 2, 2, 2), dtype=np.complex128, order="F")
    for s in range(2):
        m = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
        m = 0.5 * (m + m.conj().T)
        m[np.diag_indices(4)] = m.diagonal().real + (0.2 if s == 0 else -0.2)
        for il in range(2):
            for jl in range(2):
                for io in range(2):
                    for jo in range(2):
                        h[il, jl, s, s, io, jo] = m[io + 2 * il, jo + 2 * jl]
    hk = np.repeat(h[..., None], 4, axis=6)
    hk = np.asfortranarray(hk)
    return p, h, hk, None, None


def bath_model(nlat=3, nb=3, seed=11, hfshift=False):
    rng = np.random.default_rng(seed)
    p = VcaParams(nlat=nlat, nlat_bath=nb, norb_bath=1, uloc=(3.0, 0, 0, 0, 0),
                  hfshift=hfshift, lanc_method="lanczos")
    hloc, hk = configs.chain(nlat, -1.0, -1.0, 4)
    hb = rng.normal(0, 0.2, (nb, nb))
    hb = hb + hb.T
    hb[np.diag_indices(nb)] = rng.uniform(-1.5, 1.5, nb)
    vb = rng.normal(0, 0.4, (nlat, nb))
    bath_h = np.zeros((nb, nb, 1, 1, 1, 1), dtype=np.complex128, order="F")
    bath_v = np.zeros((nlat, nb, 1, 1, 1, 1), dtype=np.complex128, order="F")
    bath_h[:, :, 0, 0, 0, 0] = hb
    bath_v[:, :, 0, 0, 0, 0] = vb
    return p, hloc, hk, bath_h, bath_v


def bhz_small(nk=4):
    p = VcaParams(nlat=2, norb=2, nspin=2, uloc=(2.0, 2.0, 0, 0, 0), lanc_method="lanczos")
    hloc, hk = configs.bhz(2, 1, 2.0, 0.3, nk=(nk, 1))
    return p, hloc, hk, None, None


def oracle_model(p, hloc, bath_h=None, bath_v=None):
    return R.Model(to_oracle(p), hloc, bath_h, bath_v)
