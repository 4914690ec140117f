"""Input generators for the benchmark workloads (cluster hopping Hloc,
lattice hopping H_k on the reduced Brillouin zone, baths).

Conventions restated from the reference drivers:
  * site index N = Nx*(iy-1) + ix (drivers/vca_hm_2dsquare_bath.f90:294-300);
  * Hloc(i, j) is the amplitude of c^+_i c_j;
  * H_k carries the cluster hopping plus the inter-cluster hops with phase
    e^{+-i k L} (drivers/vca_hm_chain1d.f90:204-227,
    drivers/vca_hm_2dsquare_bath.f90:222-269), k on the grid
    2 pi n / (N_k L) per direction (the DMFTtools grid divided by the
    cluster size, drivers/vca_hm_chain1d.f90:187-188);
  * BHZ blocks t_m = M sigma_z, t_x = -t sigma_z + (s/2) i lambda sigma_x,
    t_y = -t sigma_z + [[0,-lambda/2],[lambda/2,0]]
    (drivers/vca_bhz_2d.f90:443-472), s = (-1)^(ispin+1).
"""
from __future__ import annotations

import numpy as np

from .api import VcaParams

SZ = np.array([[1, 0], [0, -1]], dtype=np.complex128)
SX = np.array([[0, 1], [1, 0]], dtype=np.complex128)


def _zeros(nlat, nspin, norb, nk=None):
    shape = (nlat, nlat, nspin, nspin, norb, norb) + ((nk,) if nk else ())
    return np.zeros(shape, dtype=np.complex128, order="F")


def site(ix, iy, nx):
    """0-based site index of (ix, iy) (0-based)."""
    return ix + nx * iy


def kgrid(nkx, nky, nx, ny):
    kx = 2 * np.pi * np.arange(nkx) / (nkx * nx)
    ky = 2 * np.pi * np.arange(nky) / (nky * ny) if nky else np.zeros(1)
    return [(a, b) for b in ky for a in kx]


def square_lattice(nx, ny, nspin, t, onsite, tk_pairs_k=None, nk=(1, 1)):
    """Hloc and H_k of an nx x ny open cluster of a square lattice with
    nearest-neighbour hopping t (matrix element), on-site block onsite[i, s].
    Returns (hloc, hk)."""
    nlat = nx * ny
    hloc = _zeros(nlat, nspin, 1)
    for s in range(nspin):
        for iy in range(ny):
            for ix in range(nx):
                i = site(ix, iy, nx)
                hloc[i, i, s, s, 0, 0] = onsite[i, s]
                if ix + 1 < nx:
                    j = site(ix + 1, iy, nx)
                    hloc[j, i, s, s, 0, 0] = t
                    hloc[i, j, s, s, 0, 0] = t
                if iy + 1 < ny:
                    j = site(ix, iy + 1, nx)
                    hloc[j, i, s, s, 0, 0] = t
                    hloc[i, j, s, s, 0, 0] = t
    ks = kgrid(nk[0], nk[1], nx, ny)
    hk = _zeros(nlat, nspin, 1, len(ks))
    for ik, (kx, ky) in enumerate(ks):
        for s in range(nspin):
            for iy in range(ny):
                for ix in range(nx):
                    i = site(ix, iy, nx)
                    if ix + 1 < nx:
                        j = site(ix + 1, iy, nx)
                        hk[j, i, s, s, 0, 0, ik] = t
                        hk[i, j, s, s, 0, 0, ik] = t
                    if iy + 1 < ny:
                        j = site(ix, iy + 1, nx)
                        hk[j, i, s, s, 0, 0, ik] = t
                        hk[i, j, s, s, 0, 0, ik] = t
            # inter-cluster hops across the x and y boundaries
            for iy in range(ny):
                a, b = site(0, iy, nx), site(nx - 1, iy, nx)
                hk[b, a, s, s, 0, 0, ik] += t * np.exp(1j * kx * nx)
                hk[a, b, s, s, 0, 0, ik] += t * np.exp(-1j * kx * nx)
            if nk[1]:
                for ix in range(nx):
                    a, b = site(ix, 0, nx), site(ix, ny - 1, nx)
                    hk[b, a, s, s, 0, 0, ik] += t * np.exp(1j * ky * ny)
                    hk[a, b, s, s, 0, 0, ik] += t * np.exp(-1j * ky * ny)
    return hloc, hk


def chain(nlat, t_cluster, t_lattice, nk, nspin=1):
    """1D chain cluster (drivers/vca_hm_chain1d.f90:204-227, zero-initialised)."""
    hloc = _zeros(nlat, nspin, 1)
    hk = _zeros(nlat, nspin, 1, nk)
    ks = 2 * np.pi * np.arange(nk) / (nk * nlat)
    for s in range(nspin):
        for i in range(nlat - 1):
            hloc[i, i + 1, s, s, 0, 0] = t_cluster
            hloc[i + 1, i, s, s, 0, 0] = t_cluster
            hk[i, i + 1, s, s, 0, 0, :] = t_lattice
            hk[i + 1, i, s, s, 0, 0, :] = t_lattice
        hk[0, nlat - 1, s, s, 0, 0, :] += t_lattice * np.exp(1j * ks * nlat)
        hk[nlat - 1, 0, s, s, 0, 0, :] += t_lattice * np.exp(-1j * ks * nlat)
    return hloc, hk


def bhz(nx, ny, m, lam, t=1.0, nk=(1, 1)):
    """2D BHZ cluster, Norb=2, Nspin=2 (drivers/vca_bhz_2d.f90:312-472)."""
    nlat, nspin, norb = nx * ny, 2, 2

    def tx(s):
        sz = (-1.0) ** s  # s = 0 for spin up: (-1)^(ispin+1) with ispin=1
        return -t * SZ + 0.5 * sz * 1j * lam * SX

    ty = -t * SZ + np.array([[0, -lam / 2], [lam / 2, 0]], dtype=np.complex128)
    tm = m * SZ

    # work array [ilat, jlat, ispin, iorb, jorb] (spin-diagonal model)
    def cluster():
        h = np.zeros((nlat, nlat, nspin, norb, norb), dtype=np.complex128)
        for s in range(nspin):
            for iy in range(ny):
                for ix in range(nx):
                    i = site(ix, iy, nx)
                    h[i, i, s] = tm
                    if ix + 1 < nx:
                        j = site(ix + 1, iy, nx)
                        h[j, i, s] = tx(s)
                        h[i, j, s] = tx(s).conj().T
                    if iy + 1 < ny:
                        j = site(ix, iy + 1, nx)
                        h[j, i, s] = ty
                        h[i, j, s] = ty.T
        return h

    def to_nnn(h, out):
        for s in range(nspin):
            out[:, :, s, s, :, :] = h[:, :, s]

    hloc = _zeros(nlat, nspin, norb)
    to_nnn(cluster(), hloc)
    ks = kgrid(nk[0], nk[1], nx, ny)
    hk = _zeros(nlat, nspin, norb, len(ks))
    base = cluster()
    for ik, (kx, ky) in enumerate(ks):
        h = base.copy()
        for s in range(nspin):
            for iy in range(ny):
                a, b = site(0, iy, nx), site(nx - 1, iy, nx)
                h[b, a, s] += tx(s).conj().T * np.exp(1j * kx * nx)
                h[a, b, s] += tx(s) * np.exp(-1j * kx * nx)
            for ix in range(nx):
                a, b = site(ix, 0, nx), site(ix, ny - 1, nx)
                h[b, a, s] += ty.T * np.exp(1j * ky * ny)
                h[a, b, s] += ty * np.exp(-1j * ky * ny)
        blk = np.zeros((nlat, nlat, nspin, nspin, norb, norb), dtype=np.complex128)
        to_nnn(h, blk)
        hk[..., ik] = blk
    return hloc, hk


# ------------------------------------------------------------- workloads
def config_chain4(nk=32):
    """4-site Hubbard chain, U=4, half filling (HFMODE=T, XMU=0), t=t'=-1."""
    p = VcaParams(nlat=4, uloc=(4.0, 0, 0, 0, 0), lanc_method="lanczos",
                  lanc_ngfiter=200, beta=100.0, omega_nfreq=4096)
    hloc, hk = chain(4, -1.0, -1.0, nk)
    return p, hloc, hk, None, None


def config_square_3x4(nk=64, ngfiter=300):
    """12-site 3x4 square cluster, U=8, half filling, GS sector (6,6)."""
    p = VcaParams(nlat=12, uloc=(8.0, 0, 0, 0, 0), lanc_method="lanczos",
                  lanc_ngfiter=ngfiter, lanc_tolerance=1e-12, sectors=[(6, 6)])
    hloc, hk = square_lattice(3, 4, 1, -1.0, np.zeros((12, 1)), nk=(nk, nk))
    return p, hloc, hk, None, None


def config_batched_2x5(nk=16, seed_eps=3001, seed_h=3002):
    """10-site 2x5 cluster, U=6, Nspin=2, random on-site field and Weiss AF
    field in Hloc only; every sector, LANC_METHOD=arpack (the reference
    default: up to LANC_NSTATES_SECTOR=6 eigenpairs per sector)."""
    eps = np.random.default_rng(seed_eps).uniform(-0.2, 0.2, 10)
    h = np.random.default_rng(seed_h).uniform(0.0, 0.3)
    onsite = np.zeros((10, 2))
    for iy in range(5):
        for ix in range(2):
            i = site(ix, iy, 2)
            stag = 1.0 if (ix + iy) % 2 == 0 else -1.0
            onsite[i, 0] = eps[i] - stag * h
            onsite[i, 1] = eps[i] + stag * h
    p = VcaParams(nlat=10, nspin=2, uloc=(6.0, 0, 0, 0, 0), lanc_method="arpack")
    hloc, _ = square_lattice(2, 5, 2, -1.0, onsite, nk=(1, 1))
    _, hk = square_lattice(2, 5, 2, -1.0, np.zeros((10, 2)), nk=(nk, nk))
    return p, hloc, hk, None, None


def config_bath_2x4(nk=16, seed_h=4001, seed_v=4002, precision="mixed"):
    """8-site 2x4 cluster + 8 bath levels, U=4, sector (8,8)."""
    rh = np.random.default_rng(seed_h)
    hb = rh.normal(0.0, 0.1, (8, 8))
    hb = np.triu(hb, 1)
    hb = hb + hb.T
    hb[np.diag_indices(8)] = rh.uniform(-2.0, 2.0, 8)
    vb = np.random.default_rng(seed_v).normal(0.0, 0.3, (8, 8))
    bath_h = np.zeros((8, 8, 1, 1, 1, 1), dtype=np.complex128, order="F")
    bath_v = np.zeros((8, 8, 1, 1, 1, 1), dtype=np.complex128, order="F")
    bath_h[:, :, 0, 0, 0, 0] = hb
    bath_v[:, :, 0, 0, 0, 0] = vb
    p = VcaParams(nlat=8, nlat_bath=8, norb_bath=1, uloc=(4.0, 0, 0, 0, 0),
                  lanc_method="lanczos", sectors=[(8, 8)], precision=precision)
    hloc, hk = square_lattice(2, 4, 1, -1.0, np.zeros((8, 1)), nk=(nk, nk))
    return p, hloc, hk, bath_h, bath_v


def config_bhz_2x2(nk=128, nfreq=8192, precision="mixed"):
    """BHZ 2x2, Norb=2, Nspin=2, M=2, lambda=0.3, U=2 per orbital, half
    filling, Omega on the Matsubara grid at beta=200."""
    p = VcaParams(nlat=4, norb=2, nspin=2, uloc=(2.0, 2.0, 0, 0, 0), lanc_method="lanczos",
                  beta=200.0, omega_mode="matsubara", omega_nfreq=nfreq,
                  sectors=[(4, 4)], precision=precision)
    hloc, hk = bhz(2, 2, 2.0, 0.3, nk=(nk, nk))
    return p, hloc, hk, None, None
