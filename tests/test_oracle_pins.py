"""Pins of the CPU oracle by routes independent of its own code: brute
force over bit strings, a Jordan-Wigner construction of H in the full Fock
space, closed-form spectra, exact Lehmann sums and closed-form integrals."""
import math

import numpy as np
import pytest

from oracle import vca_ref as R
from vca_amd import configs

from . import _models as M
from ._closed_forms import matsubara_tail as _matsubara_tail
from ._closed_forms import matsubara_u0_exact as _matsubara_u0_exact


# ------------------------------------------------------------ integers --
def colex_rank(s):
    r, k = 0, 1
    while s:
        p = (s & -s).bit_length() - 1
        r += math.comb(p, k)
        k += 1
        s &= s - 1
    return r


@pytest.mark.parametrize("ns", [1, 4, 7, 10, 12])
def test_rank_equals_binary_search(ns):
    lib = R.lib()
    for n in range(ns + 1):
        m = R.Model(R.Params(nlat=ns), np.zeros((ns, ns, 1, 1, 1, 1), complex)).map(n)
        assert len(m) == math.comb(ns, n)
        assert np.all(np.diff(m) > 0)
        for i, s in enumerate(m):
            assert lib.orc_binary_search(m.ctypes.data_as(__import__("ctypes").c_void_p),
                                         len(m), int(s)) == i + 1
            assert colex_rank(int(s)) == i


# ------------------------------------------------------- Jordan-Wigner --
def jw_hamiltonian(p: R.Params, hloc, bath_h=None, bath_v=None):
    """H on the full Fock space of 2*Ns modes, modes 0..Ns-1 = up orbitals,
    Ns..2Ns-1 = down orbitals, basis index b = up | dw << Ns; fermionic signs
    from the Jordan-Wigner strings; terms from the physical definition."""
    ns = p.ns
    nm = 2 * ns
    dim = 1 << nm
    import scipy.sparse as sp

    def c(mode):
        rows, cols, vals = [], [], []
        for b in range(dim):
            if (b >> mode) & 1:
                sgn = -1.0 if bin(b & ((1 << mode) - 1)).count("1") & 1 else 1.0
                rows.append(b ^ (1 << mode))
                cols.append(b)
                vals.append(sgn)
        return sp.csr_matrix((vals, (rows, cols)), shape=(dim, dim), dtype=complex)

    cs = [c(m) for m in range(nm)]
    cd = [x.getH() for x in cs]
    n = [cd[m] @ cs[m] for m in range(nm)]
    H = sp.csr_matrix((dim, dim), dtype=complex)
    model = R.Model(p, hloc, bath_h, bath_v)
    nl, no, nsp = p.nlat, p.norb, p.nspin
    imp = lambda il, io: io + il * no
    bpos = lambda ib, ob: nl * no + ib * p.norb_bath + ob
    for spin in range(2):
        si = 0 if spin == 0 else nsp - 1
        off = spin * ns
        for il in range(nl):
            for jl in range(nl):
                for io in range(no):
                    for jo in range(no):
                        a, b = imp(il, io), imp(jl, jo)
                        v = hloc[il, jl, si, si, io, jo]
                        if a == b:
                            H = H + (v - p.xmu) * n[off + a]
                        elif v != 0:
                            H = H + v * (cd[off + a] @ cs[off + b])
        if p.nlat_bath * p.norb_bath:
            bh, bv = model.bath_h, model.bath_v
            for ib in range(p.nlat_bath):
                for jb in range(p.nlat_bath):
                    for ob in range(p.norb_bath):
                        for pb in range(p.norb_bath):
                            a, b = bpos(ib, ob), bpos(jb, pb)
                            v = bh[ib, jb, si, si, ob, pb]
                            if a == b:
                                H = H + (v - p.xmu) * n[off + a]
                            elif v != 0:
                                H = H + v * (cd[off + a] @ cs[off + b])
            for il in range(nl):
                for io in range(no):
                    for jb in range(p.nlat_bath):
                        for pb in range(p.norb_bath):
                            a, b = imp(il, io), bpos(jb, pb)
                            v = bv[il, jb, si, si, io, pb]
                            H = H + v * (cd[off + a] @ cs[off + b])
                            H = H + np.conj(v) * (cd[off + b] @ cs[off + a])
    for il in range(nl):
        for io in range(no):
            a = imp(il, io)
            H = H + p.uloc[io] * (n[a] @ n[ns + a])
            for jo in range(no):
                if jo != io:
                    H = H + p.ust * (n[a] @ n[ns + imp(il, jo)])
                if jo > io:
                    b = imp(il, jo)
                    H = H + (p.ust - p.jh) * (n[a] @ n[b] + n[ns + a] @ n[ns + b])
            if p.hfmode:
                H = H - 0.5 * p.uloc[io] * (n[a] + n[ns + a])
                H = H + 0.25 * p.uloc[io] * sp.identity(dim, format="csr")
                for jo in range(io + 1, no):
                    b = imp(il, jo)
                    s4 = n[a] + n[ns + a] + n[b] + n[ns + b]
                    u2 = 2 * p.ust - p.jh
                    H = H - 0.5 * u2 * s4 + 0.25 * u2 * sp.identity(dim, format="csr")
    return H.tocsr()


def project(H, ns, model, nup, ndw):
    mu, md = model.map(nup), model.map(ndw)
    idx = np.array([int(u) | (int(d) << ns) for d in md for u in mu])
    return H[idx][:, idx].toarray()


JW_CASES = {
    "chain5": lambda: M.chain(nlat=5, u=3.0),
    "multiorb": M.multiorb,
    "bath": lambda: M.bath_model(nlat=2, nb=2),
    "bath_noshift": lambda: M.bath_model(nlat=2, nb=2, hfshift=True),
    "bhz": M.bhz_small,
}


@pytest.mark.parametrize("name", sorted(JW_CASES))
def test_dense_h_equals_jordan_wigner(name):
    p, hloc, hk, bh, bv = JW_CASES[name]()
    po = M.to_oracle(p)
    model = R.Model(po, hloc, bh, bv)
    H = jw_hamiltonian(po, hloc, bh, bv)
    ns = po.ns
    for nup in range(ns + 1):
        for ndw in range(ns + 1):
            hs = model.dense_h(nup, ndw)
            hj = project(H, ns, model, nup, ndw)
            assert np.abs(hs - hj).max() < 1e-12, (nup, ndw)
            assert np.abs(hs - hs.conj().T).max() < 1e-13


def test_hxv_direct_sweeps_equal_dense():
    p, hloc, hk, bh, bv = M.bath_model(nlat=3, nb=3)
    model = M.oracle_model(p, hloc, bh, bv)
    rng = np.random.default_rng(0)
    for (nup, ndw) in [(3, 3), (2, 4), (6, 1)]:
        h = model.dense_h(nup, ndw)
        v = rng.normal(size=h.shape[0]) + 1j * rng.normal(size=h.shape[0])
        assert np.abs(h @ v - model.hxv(nup, ndw, v)).max() < 1e-12
        assert np.abs(h @ v - model.hxv(nup, ndw, v, workers=3)).max() < 1e-12


def test_hxv_slices_sum_to_one_application():
    """The slices of the DW decomposition (the bounded CPU-baseline sample)
    add up to the dense H v, whether run together on threads or one at a
    time."""
    p, hloc, hk, bh, bv = M.bath_model(nlat=3, nb=3)
    model = M.oracle_model(p, hloc, bh, bv)
    h = model.dense_h(3, 3)
    v = np.random.default_rng(1).normal(size=h.shape[0]) * (1 + 0.5j)
    for nsl, threads in ((7, 1), (7, 3)):
        hv = model.hxv_slices(3, 3, v, np.zeros_like(v), nsl, 0, nsl, threads)
        assert np.abs(h @ v - hv).max() < 1e-12
    hv = np.zeros_like(v)
    for k in range(5):
        model.hxv_slices(3, 3, v, hv, 5, k, 1)
    assert np.abs(h @ v - hv).max() < 1e-12


# ------------------------------------------------------------- spectra --
def test_dimer_closed_form():
    for u in (0.0, 2.0, 8.0):
        for hf in (False, True):
            p = R.Params(nlat=2, uloc=(u, 0, 0, 0, 0), hfmode=hf)
            h = R.zeros_nnn(p)
            h[0, 1, 0, 0, 0, 0] = h[1, 0, 0, 0, 0, 0] = -1.0
            e = np.linalg.eigvalsh(R.Model(p, h).dense_h(1, 1))[0]
            exact = (u - math.sqrt(u * u + 16)) / 2 if not hf else -math.sqrt(u * u + 16) / 2
            assert abs(e - exact) < 1e-13


def test_u0_slater_determinant():
    p, hloc, hk, _, _ = M.chain(nlat=6, u=0.0)
    po = M.to_oracle(p)
    model = R.Model(po, hloc)
    eps = np.linalg.eigvalsh(R.nnn2lso(hloc, 6, 1, 1))
    for nup, ndw in [(3, 3), (2, 4), (1, 0), (6, 5)]:
        e = np.linalg.eigvalsh(model.dense_h(nup, ndw))[0]
        assert abs(e - (eps[:nup].sum() + eps[:ndw].sum())) < 1e-12


def test_atomic_limit():
    p = R.Params(nlat=3, uloc=(5.0, 0, 0, 0, 0), hfmode=False, xmu=1.0)
    h = R.zeros_nnn(p)
    for i, e in enumerate((0.3, -0.2, 0.1)):
        h[i, i, 0, 0, 0, 0] = e
    model = R.Model(p, h)
    mu, md = model.map(2), model.map(1)
    d = model.diag(2, 1)
    k = 0
    for sd in md:
        for su in mu:
            e = 0.0
            for i, eps in enumerate((0.3, -0.2, 0.1)):
                nu, nd = (su >> i) & 1, (sd >> i) & 1
                e += (eps - 1.0) * (nu + nd) + 5.0 * nu * nd
            assert abs(d[k] - e) < 1e-14
            k += 1


def test_particle_hole_symmetry():
    p, hloc, hk, _, _ = M.chain(nlat=6, u=4.0)
    model = M.oracle_model(p, hloc)
    for nup, ndw in [(2, 3), (1, 1), (3, 3), (0, 2)]:
        e1 = np.linalg.eigvalsh(model.dense_h(nup, ndw))[0]
        e2 = np.linalg.eigvalsh(model.dense_h(6 - nup, 6 - ndw))[0]
        assert abs(e1 - e2) < 1e-12


# ------------------------------------------------------ Green functions --
def test_u0_gf_is_resolvent():
    p, hloc, hk, _, _ = M.chain(nlat=4, u=0.0, nk=4)
    po = M.to_oracle(p)
    r = R.vca_solve(po, hloc, hk)
    hl = R.nnn2lso(hloc, 4, 1, 1)
    for z in (0.4j, 1.3 + 0.2j, -0.7 + 1.0j):
        g = R.nnn2lso(R.gf_cluster(r["gf"], z), 4, 1, 1)
        exact = np.linalg.inv((z + po.xmu) * np.eye(4) - hl)
        assert np.abs(g - exact).max() < 1e-10


def test_hubbard_atom_gf():
    u = 3.0
    p = R.Params(nlat=1, uloc=(u, 0, 0, 0, 0))
    r = R.vca_solve(p, np.zeros((1, 1, 1, 1, 1, 1), complex), np.zeros((1, 1, 1, 1, 1, 1, 1), complex))
    assert len(r["states"]) == 2  # (1,0) and (0,1), degenerate
    for z in (0.5j, 2.0 + 0.1j):
        g = R.gf_cluster(r["gf"], z)[0, 0, 0, 0, 0, 0]
        assert abs(g - (0.5 / (z - u / 2) + 0.5 / (z + u / 2))) < 1e-13


def exact_lehmann(p, hloc, z):
    """G_ij(z) from full diagonalisation in Fock space (T=0, unique GS)."""
    po = M.to_oracle(p)
    H = jw_hamiltonian(po, hloc).toarray()
    ns = po.ns
    e, v = np.linalg.eigh(H)
    g0 = v[:, 0]
    nm = 2 * ns
    dim = 1 << nm

    def cmat(mode):
        c = np.zeros((dim, dim))
        for b in range(dim):
            if (b >> mode) & 1:
                c[b ^ (1 << mode), b] = -1.0 if bin(b & ((1 << mode) - 1)).count("1") & 1 else 1.0
        return c

    n = p.nlat * p.norb
    G = np.zeros((n, n), complex)
    cs = [cmat(m) for m in range(n)]
    res = np.linalg.inv(z * np.eye(dim) - (H - e[0] * np.eye(dim)))
    res2 = np.linalg.inv(z * np.eye(dim) + (H - e[0] * np.eye(dim)))
    for i in range(n):
        for j in range(n):
            G[i, j] = (g0.conj() @ cs[i] @ res @ cs[j].T @ g0
                       + g0.conj() @ cs[j].T @ res2 @ cs[i] @ g0)
    return G


def test_offdiagonal_recombination_exact():
    for sym in (False, True):
        p, hloc, hk, _, _ = M.chain(nlat=4, u=4.0, nk=2)
        p.gf_symmetric = sym
        r = R.vca_solve(M.to_oracle(p), hloc, hk)
        assert len(r["states"]) == 1
        for z in (0.3j, 1.1 + 0.5j):
            g = R.nnn2lso(R.gf_cluster(r["gf"], z), 4, 1, 1)
            ge = exact_lehmann(p, hloc, z)
            assert np.abs(g - ge).max() < 1e-10, sym


# ------------------------------------------ grid G, Sigma = G0^-1 - G^-1 --
def test_frequency_grids():
    p = R.Params(beta=37.0, lmats=11, lreal=9, wini=-3.0, wfin=2.0)
    wm, wr = R.frequency_grids(p)
    assert np.abs(wm - np.array([(2 * n + 1) * math.pi / 37.0 for n in range(11)])).max() < 1e-15
    assert wr[0] == -3.0 and abs(wr[-1] - 2.0) < 1e-15
    assert np.abs(np.diff(wr) - 5.0 / 8.0).max() < 1e-15


def _u0(model):
    if model == "chain":
        p, hloc, hk, bh, bv = M.chain(nlat=4, u=0.0, nk=2)
    elif model == "bath":
        p, hloc, hk, bh, bv = M.bath_model(nlat=2, nb=2)
        p.uloc = (0.0, 0, 0, 0, 0)
    else:
        p, hloc, hk, bh, bv = M.bhz_small(nk=2)
        p.uloc = (0.0, 0.0, 0, 0, 0)
    p.lmats, p.lreal, p.beta, p.eps = 24, 24, 30.0, 0.05
    return p, hloc, hk, bh, bv


@pytest.mark.parametrize("model", ["chain", "bath", "bhz"])
def test_sigma_vanishes_at_u0(model):
    """At U=0 the cluster G is the resolvent of Hloc + Delta, so
    Sigma = G0^-1 - G^-1 = 0 and G = G0 on both grids: pins the grid pole
    sums, the recombination, Delta and the lso reshapes together."""
    p, hloc, hk, bh, bv = _u0(model)
    r = R.vca_solve(M.to_oracle(p), hloc, hk, bh, bv, epsabs=1e-8, epsrel=1e-8)
    for axis in (0, 1):
        g, s, g0 = R.sigma_grid(r["model"], r["gf"], axis)
        assert g.shape == (p.nlat, p.nlat, p.nspin, p.nspin, p.norb, p.norb, 24)
        assert np.abs(s).max() < 1e-9, (model, axis, np.abs(s).max())
        assert np.abs(g - g0).max() < 1e-10


def test_sigma_hubbard_atom():
    """Half-filled Hubbard atom (HFMODE, xmu=0): G = (1/(z-U/2) + 1/(z+U/2))/2,
    so Sigma(z) = U^2/(4z) exactly, on the Matsubara and real grids."""
    u = 3.0
    p = R.Params(nlat=1, uloc=(u, 0, 0, 0, 0), beta=25.0, lmats=40, lreal=31, eps=0.02)
    r = R.vca_solve(p, np.zeros((1, 1, 1, 1, 1, 1), complex),
                    np.zeros((1, 1, 1, 1, 1, 1, 1), complex))
    wm, wr = R.frequency_grids(p)
    for axis, z in ((0, 1j * wm), (1, wr + 1j * p.eps)):
        _, s, g0 = R.sigma_grid(r["model"], r["gf"], axis)
        assert np.abs(s[0, 0, 0, 0, 0, 0] - u * u / (4 * z)).max() < 1e-11
        assert np.abs(g0[0, 0, 0, 0, 0, 0] - 1 / z).max() < 1e-13


def test_sigma_equals_exact_lehmann():
    """U=4 chain: Sigma from the oracle's grid equals (z + mu) - Hloc -
    G_exact^-1 with G_exact from full Fock-space diagonalisation."""
    p, hloc, hk, _, _ = M.chain(nlat=4, u=4.0, nk=2)
    p.lmats, p.lreal, p.beta = 6, 5, 10.0
    po = M.to_oracle(p)
    r = R.vca_solve(po, hloc, hk)
    wm, wr = R.frequency_grids(po)
    hl = R.nnn2lso(hloc, 4, 1, 1)
    for axis, zs in ((0, 1j * wm), (1, wr + 1j * po.eps)):
        _, s, _ = R.sigma_grid(r["model"], r["gf"], axis)
        for i in (0, len(zs) - 1):
            z = zs[i]
            se = (z + po.xmu) * np.eye(4) - hl - np.linalg.inv(exact_lehmann(p, hloc, z))
            assert np.abs(R.nnn2lso(s[..., i], 4, 1, 1) - se).max() < 1e-9


# --------------------------------------------------- Omega, quadrature --
def test_quadrature_closed_form():
    for a in (0.3, 1.0, 4.0):
        val, err = R.quad_semi_infinite(lambda x: np.log1p(a * a / (x * x)), 1e-13, 1e-13)
        assert abs(val - math.pi * a) < 1e-9 * math.pi * a


@pytest.mark.parametrize("tprime", [-1.0, -0.6])
def test_sft_exact_u0(tprime):
    p, hloc, hk, _, _ = M.chain(nlat=4, u=0.0, nk=16, tprime=tprime)
    po = M.to_oracle(p)
    r = R.vca_solve(po, hloc, hk, epsabs=1e-12, epsrel=1e-13)
    hkl = R.hk_to_lso(po, hk)
    exact = 2 * np.mean([np.sum(np.minimum(np.linalg.eigvalsh(h), 0)) for h in hkl])
    assert abs(r["sft_potential"] - exact) < 1e-9 * abs(exact)


def test_matsubara_tail_series():
    """The tail series against direct summation far out plus the midpoint
    rule's integral beyond it."""
    beta, n0, c = 10.0, 40, 7.3
    n = np.arange(n0, 4_000_000)
    w = (2 * n + 1) * math.pi / beta
    direct = np.log1p(c / w ** 2).sum()
    w_end = 2 * 4_000_000 * math.pi / beta
    direct += beta / (2 * math.pi) * c / w_end  # int_{w_end}^inf c/w^2 dw / (2pi/beta)
    assert abs(direct - _matsubara_tail(c, beta, n0)) < 1e-12


@pytest.mark.parametrize("case", ["chain4", "chain4_tprime", "bhz"])
def test_matsubara_sum_exact_u0(case):
    """The fixed-grid Omega (matsubara_integration: (3-Nspin)(2/beta)
    sum_{n<N} f(i w_n), w_n = (2n+1)pi/beta) at U=0 against the cosh
    product minus the exactly summed tail beyond N: pins the odd
    frequencies, the 2/beta weight, the (3-Nspin) factor and the truncation
    at N together (a grid (2n-1)pi/beta from n=0, a weight 1/beta or a factor
    2 for Nspin=2 each miss by orders of magnitude more than the bar)."""
    if case == "bhz":
        p, hloc, hk, _, _ = M.bhz_small(nk=4)
        p.uloc = (0.0, 0.0, 0, 0, 0)
        beta, nfreq = 20.0, 300
    else:
        tp = -1.0 if case == "chain4" else -0.6
        p, hloc, hk, _, _ = M.chain(nlat=4, u=0.0, nk=8, tprime=tp)
        beta, nfreq = 100.0, 4096
    p.beta = beta
    po = M.to_oracle(p)
    r = R.vca_solve(po, hloc, hk, omega_mode="matsubara", nfreq=nfreq)
    hkl = R.hk_to_lso(po, hk)
    hll = R.nnn2lso(hloc, po.nlat, po.nspin, po.norb)
    exact = _matsubara_u0_exact(hll, hkl, po.nspin, beta, nfreq)
    assert abs(r["omega_integral"] - exact) < 1e-10 * max(1.0, abs(exact)), (
        r["omega_integral"], exact)
    # the truncation matters at this N: the untruncated closed form differs
    inf = _matsubara_u0_exact(hll, hkl, po.nspin, beta, 10 ** 9)
    assert abs(inf - exact) > 1e3 * 1e-10 * max(1.0, abs(exact))


def test_sft_exact_u0_bhz():
    """Spin-orbit-free but orbitally complex (BHZ) U=0 cluster."""
    p, hloc, hk, _, _ = M.bhz_small(nk=4)
    p.uloc = (0.0, 0.0, 0, 0, 0)
    po = M.to_oracle(p)
    r = R.vca_solve(po, hloc, hk, epsabs=1e-12, epsrel=1e-13)
    hkl = R.hk_to_lso(po, hk)
    exact = np.mean([np.sum(np.minimum(np.linalg.eigvalsh(h), 0)) for h in hkl])
    assert abs(r["sft_potential"] - exact) < 1e-8 * abs(exact)


# ------------------------------------------------------------ finite T --
@pytest.mark.parametrize("beta,tprime,tol", [(50.0, -1.0, 1e-11), (20.0, -0.7, 2e-8)])
def test_sft_exact_u0_finite_t(beta, tprime, tol):
    """U=0 at finite T: the reference's Matsubara sum to Nmax plus the
    contour integral (VCA_OMEGA.f90:153-204), with Omega' from the kept
    states (VCA_MAIN.f90:186-190) and Boltzmann GF weights, equals
    -(3-Nspin)/(beta Nk) sum_k sum_n ln(1 + exp(-beta e_kn)) for every t'.
    At t' != t the kept spectrum is cut at exp(-beta dE) <= CUTOFF, which
    bounds the deviation (tol)."""
    p, hloc, hk, _, _ = M.chain(nlat=4, u=0.0, nk=16, tprime=tprime)
    po = M.to_oracle(p)
    po.beta, po.lanc_nstates_total, po.bandwidth = beta, 2, 4.0
    r = R.vca_solve(po, hloc, hk, epsabs=1e-11, epsrel=1e-12)
    hkl = R.hk_to_lso(po, hk)
    eps = np.concatenate([np.linalg.eigvalsh(h) for h in hkl])
    exact = -(3 - po.nspin) / beta * np.sum(np.log1p(np.exp(-beta * eps))) / len(hkl)
    assert abs(r["sft_potential"] - exact) < tol, (r["sft_potential"], exact)
    assert r["nmax"] > 0


def test_finite_t_state_list_semantics():
    """es_add_state with a size limit keeps the lowest states in ascending
    order, new states before equal energies (VCA_EIGENSPACE.f90:133-212)."""
    st = lambda e, tag: R.State(e, (tag, 0), np.zeros(1))
    lst = []
    for e, tag in ((3.0, 0), (1.0, 1), (2.0, 2), (1.0, 3), (0.5, 4), (5.0, 5)):
        R.es_add_sized(lst, st(e, tag), 3)
    assert [(s.energy, s.sector[0]) for s in lst] == [(0.5, 4), (1.0, 3), (1.0, 1)]


def test_finite_t_low_temperature_limit():
    """A gapped cluster at beta = 200: the finite-T Omega (state list,
    annealing, Matsubara sum + contour) equals the T=0 Omega (one ground
    state, adaptive quadrature) up to exp(-beta gap) terms."""
    p, hloc, hk, _, _ = M.chain(nlat=4, u=4.0, nk=8, tprime=-1.0)
    po = M.to_oracle(p)
    r0 = R.vca_solve(po, hloc, hk, epsabs=1e-11, epsrel=1e-12)
    po.beta, po.lanc_nstates_total, po.bandwidth = 200.0, 2, 4.0
    rt = R.vca_solve(po, hloc, hk, epsabs=1e-11, epsrel=1e-12)
    assert abs(rt["sft_potential"] - r0["sft_potential"]) < 1e-8


def test_sum_kmesh_complex_is_principal_log_det():
    """sum_kmesh_complex takes log(det) per k (principal branch), not the
    sum of the logs of the LU pivots: compare with numpy's determinant."""
    p, hloc, hk, _, _ = M.bhz_small(nk=4)
    po = M.to_oracle(p)
    po.lanc_method = "lanczos"
    r = R.vca_solve(po, hloc, hk, epsabs=1e-8, epsrel=1e-8)
    model, gd = r["model"], r["gf"]
    hkl = R.hk_to_lso(po, hk)
    for z in (0.3 + 0.2j, -1.5 + 0.05j, 2.0j, 4.0 * np.exp(0.3j)):
        g = R.nnn2lso(R.gf_cluster(gd, z), po.nlat, po.nspin, po.norb)
        hl = R.nnn2lso(model.hloc, po.nlat, po.nspin, po.norb)
        brute = np.mean([np.log(np.linalg.det(np.eye(po.nlso) + (hl - h) @ g)) for h in hkl])
        f = R.sum_kmesh(model, gd, hkl, z, complex_out=True)
        assert abs(f - brute) < 1e-10, (z, f, brute)


def test_arpack_sector_equals_dense():
    """ARPACK path of the oracle (eigsh, which='SA', the reference's ncv
    rule) on a sector above LANC_DIM_THRESHOLD equals dense eigh."""
    p, hloc, hk, _, _ = M.chain(nlat=8, u=4.0)
    po = M.to_oracle(p)
    po.lanc_method = "arpack"
    model = R.Model(po, hloc)
    ev, vec = R.solve_sector(model, 4, 3, 6)
    ed = np.linalg.eigvalsh(model.dense_h(4, 3))[:6]
    assert np.allclose(ev, ed, atol=1e-10)
