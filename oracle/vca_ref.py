"""CPU oracle for the VCA solver-core path (TEST INFRASTRUCTURE ONLY).

Only tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker / the CPU
baseline.  The product path (``vca_amd`` + ``libvca_gpu.so``) never does.

It restates the reference's algorithm in numpy on top of the C restatement
in ``vca_oracle.c`` (Fock maps, signs, matrix elements, H x v, excitation
vectors).  Each function cites the reference file:line it follows.

Pinning (see tests/test_oracle_pins.py): the matrix elements are checked
against an independent Jordan-Wigner construction; spectra against the
Hubbard dimer, U=0 and atomic-limit closed forms and particle-hole symmetry;
the Green's function against the U=0 resolvent and the Hubbard atom; the
off-diagonal recombination against exact Lehmann sums; Omega against the
U=0 SFT exactness identity; the Matsubara-grid sum against the U=0
cosh-product closed form with its tail beyond the last frequency summed
exactly; the quadrature against closed-form integrals.
Parity unpinned (third-party numerics the reference calls but does not
ship): the exact SciFortran Lanczos conventions (start vector, stopping
rule), ARPACK's ordering of near-degenerate pairs, QUADPACK's default
tolerances and the DMFTtools k-grid offsets.
"""
from __future__ import annotations

import ctypes
import os
import math
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    """Load (building if needed) the C oracle ``build/liboracle.so``."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "build", "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", _HERE])
        L = ctypes.CDLL(path)
        L.orc_build_map.restype = ctypes.c_int
        L.orc_binary_search.restype = ctypes.c_int
        L.orc_hxv.restype = ctypes.c_int64
        L.orc_hxv_parallel.restype = ctypes.c_int64
        L.orc_hxv_slices.restype = ctypes.c_int64
        L.orc_spin_hops.restype = ctypes.c_int64
        L.orc_diag.restype = ctypes.c_int
        L.orc_excite.restype = ctypes.c_int
        _LIB = L
    return _LIB


# --------------------------------------------------------------- model ---
@dataclass
class Params:
    """Subset of the reference input variables (VCA_INPUT_VARS.f90:112-172)
    that the solver-core path reads; defaults are the reference defaults."""
    nlat: int = 1
    norb: int = 1
    nspin: int = 1
    nlat_bath: int = 0
    norb_bath: int = 0
    uloc: tuple = (2.0, 0.0, 0.0, 0.0, 0.0)
    ust: float = 0.0
    jh: float = 0.0
    xmu: float = 0.0
    beta: float = 1000.0
    hfmode: bool = True
    hfshift: bool = True
    gf_symmetric: bool = False
    lanc_method: str = "arpack"
    lanc_nstates_sector: int = 6
    lanc_nstates_total: int = 1
    lanc_niter: int = 512
    lanc_ngfiter: int = 200
    lanc_tolerance: float = 1e-12
    lanc_dim_threshold: int = 256
    gs_threshold: float = 1e-9
    sectors: list | None = None  # list of (nup, ndw); None = every sector
    lanc_nstates_step: int = 2
    lanc_niter_spectrum: int = 10
    lanc_ncv_factor: int = 3
    lanc_ncv_add: int = 5
    cutoff: float = 1e-9
    bandwidth: float = 0.0
    lmats: int = 5000
    lreal: int = 5000
    wini: float = -5.0
    wfin: float = 5.0
    eps: float = 0.01

    @property
    def finite_t(self):
        """VCA_SETUP.f90:157-162: finite T unless one state is kept."""
        return self.lanc_nstates_total > 1

    @property
    def ns(self):
        return self.nlat * self.norb + self.nlat_bath * self.norb_bath

    @property
    def nlso(self):
        return self.nlat * self.nspin * self.norb


def zeros_nnn(p: Params, extra=()):
    return np.zeros((p.nlat, p.nlat, p.nspin, p.nspin, p.norb, p.norb) + tuple(extra),
                    dtype=np.complex128, order="F")


def nnn2lso(h, nlat, nspin, norb):
    """VCA_AUX_FUNX.f90:255-279: lso index = iorb + (ilat-1)Norb + (ispin-1)Norb*Nlat."""
    n = nlat * nspin * norb
    out = np.zeros((n, n), dtype=np.complex128)
    for il in range(nlat):
        for jl in range(nlat):
            for s in range(nspin):
                for t in range(nspin):
                    for io in range(norb):
                        for jo in range(norb):
                            out[io + il * norb + s * norb * nlat,
                                jo + jl * norb + t * norb * nlat] = h[il, jl, s, t, io, jo]
    return out


def lso2nnn(h, nlat, nspin, norb):
    out = np.zeros((nlat, nlat, nspin, nspin, norb, norb), dtype=np.complex128, order="F")
    for il in range(nlat):
        for jl in range(nlat):
            for s in range(nspin):
                for t in range(nspin):
                    for io in range(norb):
                        for jo in range(norb):
                            out[il, jl, s, t, io, jo] = h[io + il * norb + s * norb * nlat,
                                                          jo + jl * norb + t * norb * nlat]
    return out


def rect_n2j(v, nlat, nlat2, nspin, norb, norb2):
    """VCA_AUX_FUNX.f90:444-468 (rectangular reshape of bath_v)."""
    out = np.zeros((nlat * nspin * norb, nlat2 * nspin * norb2), dtype=np.complex128)
    for il in range(nlat):
        for jl in range(nlat2):
            for s in range(nspin):
                for t in range(nspin):
                    for io in range(norb):
                        for jo in range(norb2):
                            out[io + il * norb + s * norb * nlat,
                                jo + jl * norb2 + t * norb2 * nlat2] = v[il, jl, s, t, io, jo]
    return out


def set_bath(p: Params, bath_h, bath_v):
    """VCA_BATH_SETUP.f90:112-138: copy and, if hfmode and not hfshift, shift
    the bath levels by -(U/2 + Ust(Norb-1)/2 + (Ust-Jh)(Norb-1)/2)."""
    h = np.array(bath_h, dtype=np.complex128, order="F", copy=True)
    v = np.array(bath_v, dtype=np.complex128, order="F", copy=True)
    if p.hfmode and not p.hfshift:
        for il in range(p.nlat_bath):
            for io in range(p.norb_bath):
                for s in range(p.nspin):
                    h[il, il, s, s, io, io] -= (0.5 * p.uloc[io] + 0.5 * p.ust * (p.norb - 1)
                                                + 0.5 * (p.ust - p.jh) * (p.norb - 1))
    return h, v


class _CModel(ctypes.Structure):
    _fields_ = [("nlat", ctypes.c_int32), ("norb", ctypes.c_int32),
                ("nspin", ctypes.c_int32), ("nlat_bath", ctypes.c_int32),
                ("norb_bath", ctypes.c_int32), ("hfmode", ctypes.c_int32),
                ("uloc", ctypes.c_double * 5), ("ust", ctypes.c_double),
                ("jh", ctypes.c_double), ("xmu", ctypes.c_double),
                ("hloc", ctypes.c_void_p), ("bath_h", ctypes.c_void_p),
                ("bath_v", ctypes.c_void_p)]


class Model:
    """Cluster model: parameters plus Hloc (and the shifted bath) in the
    reference's rank-6 column-major layouts."""

    def __init__(self, p: Params, hloc, bath_h=None, bath_v=None):
        self.p = p
        self.hloc = np.asfortranarray(np.asarray(hloc, dtype=np.complex128))
        assert self.hloc.shape == (p.nlat, p.nlat, p.nspin, p.nspin, p.norb, p.norb)
        nb = p.nlat_bath * p.norb_bath
        if nb > 0:
            self.bath_h, self.bath_v = set_bath(p, bath_h, bath_v)
        else:
            self.bath_h = np.zeros((1,) * 6, dtype=np.complex128, order="F")
            self.bath_v = np.zeros((1,) * 6, dtype=np.complex128, order="F")
        self._c = _CModel(p.nlat, p.norb, p.nspin, p.nlat_bath, p.norb_bath,
                          1 if p.hfmode else 0,
                          (ctypes.c_double * 5)(*[float(x) for x in (list(p.uloc) + [0] * 5)[:5]]),
                          p.ust, p.jh, p.xmu,
                          self.hloc.ctypes.data, self.bath_h.ctypes.data,
                          self.bath_v.ctypes.data)

    @property
    def ns(self):
        return self.p.ns

    # ---- Fock space -----------------------------------------------------
    def map(self, n):
        ns = self.ns
        d = lib().orc_build_map(ns, n, None)
        out = np.zeros(d, dtype=np.int32)
        lib().orc_build_map(ns, n, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def dims(self, nup, ndw):
        return math.comb(self.ns, nup), math.comb(self.ns, ndw)

    # ---- Hamiltonian ----------------------------------------------------
    def hxv(self, nup, ndw, v, workers=0):
        v = np.ascontiguousarray(v, dtype=np.complex128)
        hv = np.empty_like(v)
        if workers:
            bad = lib().orc_hxv_parallel(ctypes.byref(self._c), nup, ndw,
                                         v.ctypes.data_as(ctypes.c_void_p),
                                         hv.ctypes.data_as(ctypes.c_void_p), workers)
        else:
            bad = lib().orc_hxv(ctypes.byref(self._c), nup, ndw,
                                v.ctypes.data_as(ctypes.c_void_p),
                                hv.ctypes.data_as(ctypes.c_void_p))
        if bad != 0:
            raise RuntimeError(f"oracle hxv: {bad} hop targets outside the sector")
        return hv

    def hxv_slices(self, nup, ndw, v, hv, nslices, first, count, threads=1):
        """Slices [first, first+count) of an nslices-way DW decomposition of
        one H x v (VCA_HAMILTONIAN.f90:77-90) accumulated into hv: a bounded
        sample of the work of one application (count/nslices of it)."""
        assert v.dtype == np.complex128 and hv.dtype == np.complex128
        assert v.flags.c_contiguous and hv.flags.c_contiguous
        bad = lib().orc_hxv_slices(ctypes.byref(self._c), nup, ndw,
                                   v.ctypes.data_as(ctypes.c_void_p),
                                   hv.ctypes.data_as(ctypes.c_void_p), nslices, first, count,
                                   threads)
        if bad != 0:
            raise RuntimeError(f"oracle hxv_slices: status {bad}")
        return hv

    def diag(self, nup, ndw):
        du, dd = self.dims(nup, ndw)
        out = np.zeros(du * dd, dtype=np.complex128)
        lib().orc_diag(ctypes.byref(self._c), nup, ndw, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def spin_hops(self, spin, npart):
        """COO (rows, cols, amps) of the one-spin hop matrix (stored order)."""
        n = lib().orc_spin_hops(ctypes.byref(self._c), spin, npart, None, None, None)
        rows = np.zeros(n, dtype=np.int32)
        cols = np.zeros(n, dtype=np.int32)
        amps = np.zeros(n, dtype=np.complex128)
        lib().orc_spin_hops(ctypes.byref(self._c), spin, npart,
                            rows.ctypes.data_as(ctypes.c_void_p),
                            cols.ctypes.data_as(ctypes.c_void_p),
                            amps.ctypes.data_as(ctypes.c_void_p))
        return rows, cols, amps

    def spin_matrix(self, spin, npart):
        d = math.comb(self.ns, npart)
        rows, cols, amps = self.spin_hops(spin, npart)
        h = np.zeros((d, d), dtype=np.complex128)
        np.add.at(h, (rows, cols), amps)
        return h

    def dense_h(self, nup, ndw):
        """Dense sector H = diag + H_dw (x) 1 + 1 (x) H_up, the Kronecker form
        of VCA_HAMILTONIAN_SPARSE_HxV.f90:103-123."""
        du, dd = self.dims(nup, ndw)
        hu = self.spin_matrix(1, nup)
        hd = self.spin_matrix(2, ndw)
        h = np.diag(self.diag(nup, ndw))
        h += np.kron(np.eye(dd), hu) + np.kron(hd, np.eye(du))
        return h

    def excite(self, nup, ndw, psi, spin, pos, dagger, coef, out):
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        rc = lib().orc_excite(ctypes.byref(self._c), nup, ndw,
                              psi.ctypes.data_as(ctypes.c_void_p), spin, pos,
                              1 if dagger else 0, ctypes.c_double(np.real(coef)),
                              ctypes.c_double(np.imag(coef)), out.ctypes.data_as(ctypes.c_void_p))
        if rc != 0:
            raise RuntimeError("oracle excite failed")
        return out


# ------------------------------------------------------------ Lanczos ----
def start_vector(dim, seed=5001):
    """Deterministic start vector: splitmix64(seed, i) mapped to [-1, 1),
    normalised.  Identical to the one the GPU library generates."""
    i = np.arange(dim, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + (i + np.uint64(1)) *
             np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    x = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0) * 2.0 - 1.0
    v = x.astype(np.complex128)
    return v / np.linalg.norm(v)


def lanczos_tridiag(matvec, v0, nlanc):
    """Plain Lanczos from the normalised v0, no reorthogonalisation: the
    recurrence behind sp_lanc_tridiag (called at VCA_GF_NORMAL.f90:218-223).
    Returns alpha[0:nlanc], beta[0:nlanc] with beta[0] = 0 and beta[k] the
    coupling of steps k-1 and k (the reference uses beta(2:Nlanc),
    VCA_GF_NORMAL.f90:1009).  After an exact breakdown the rest is zero."""
    alpha = np.zeros(nlanc)
    beta = np.zeros(nlanc)
    v_prev = np.zeros_like(v0)
    v = v0.copy()
    b = 0.0
    for k in range(nlanc):
        w = matvec(v)
        a = np.vdot(v, w).real
        w = w - a * v - b * v_prev
        alpha[k] = a
        if k + 1 < nlanc:
            b = np.linalg.norm(w)
            beta[k + 1] = b
            if b <= 1e-300:
                break
            v_prev, v = v, w / b
    return alpha, beta


def tridiag_eigh(alpha, beta):
    n = len(alpha)
    t = np.diag(alpha) + np.diag(beta[1:n], 1) + np.diag(beta[1:n], -1)
    return np.linalg.eigh(t)


def lanczos_gs(matvec, dim, tol, niter, seed=5001):
    """Plain Lanczos ground state (sp_lanc_eigh semantics, called at
    VCA_DIAG.f90:172-184): iterate until successive lowest Ritz values differ
    by less than tol, then a second pass rebuilds the eigenvector."""
    v0 = start_vector(dim, seed)
    alpha, beta = [], [0.0]
    v_prev = np.zeros_like(v0)
    v = v0.copy()
    b = 0.0
    e_old = np.inf
    for k in range(min(niter, dim)):
        w = matvec(v)
        a = np.vdot(v, w).real
        w = w - a * v - b * v_prev
        alpha.append(a)
        e, _ = tridiag_eigh(np.array(alpha), np.array(beta[:len(alpha)]))
        if abs(e[0] - e_old) < tol:
            break
        e_old = e[0]
        b = np.linalg.norm(w)
        if b <= 1e-300:
            break
        beta.append(b)
        v_prev, v = v, w / b
    m = len(alpha)
    e, z = tridiag_eigh(np.array(alpha), np.array(beta[:m]))
    y = z[:, 0]
    x = np.zeros_like(v0)
    v_prev = np.zeros_like(v0)
    v = v0.copy()
    for k in range(m):
        x += y[k] * v
        if k + 1 == m:
            break
        w = matvec(v) - alpha[k] * v - beta[k] * v_prev
        v_prev, v = v, w / beta[k + 1]
    x /= np.linalg.norm(x)
    return e[0], x


# ----------------------------------------------------- sector bookkeeping
def sector_list(p: Params):
    """Sector scan order: isector = 1 + Ndw + Nup*(Ns+1) (VCA_SETUP.f90:410-420)."""
    if p.sectors is not None:
        return [tuple(s) for s in p.sectors]
    ns = p.ns
    return [(nup, ndw) for nup in range(ns + 1) for ndw in range(ns + 1)]


@dataclass
class State:
    energy: float
    sector: tuple
    vec: np.ndarray


def solve_sector(model: Model, nup, ndw, neigen, matvec_factory=None):
    """Eigenpairs of one sector (VCA_DIAG.f90:98-212): dense eigh when
    dim <= lanc_dim_threshold or Neigen == dim, otherwise ARPACK
    (scipy.sparse.linalg.eigsh, which='SA', ncv = ncv_factor*max(Neigen,
    nstates_sector) + ncv_add, :100-102) or the plain Lanczos
    ('lanczos')."""
    p = model.p
    du, dd = model.dims(nup, ndw)
    dim = du * dd
    lanc_solve = not (neigen == dim or dim <= p.lanc_dim_threshold)
    if not lanc_solve:
        h = model.dense_h(nup, ndw)
        e, z = np.linalg.eigh(h)
        return e[:neigen], z[:, :neigen]
    mv = (matvec_factory(nup, ndw) if matvec_factory
          else (lambda x, a=nup, b=ndw: model.hxv(a, b, x)))
    if p.lanc_method == "lanczos":
        e0, x = lanczos_gs(mv, dim, p.lanc_tolerance, p.lanc_niter)
        return np.array([e0]), x[:, None]
    from scipy.sparse.linalg import LinearOperator, eigsh
    op = LinearOperator((dim, dim), matvec=mv, dtype=np.complex128)
    ncv = min(dim, p.lanc_ncv_factor * max(neigen, p.lanc_nstates_sector) + p.lanc_ncv_add)
    e, z = eigsh(op, k=neigen, which="SA", ncv=ncv, tol=p.lanc_tolerance,
                 v0=start_vector(dim), maxiter=min(dim, p.lanc_niter))
    order = np.argsort(e)
    return e[order], z[:, order]


def es_insert(states: list, st: State):
    """es_insert_state (VCA_EIGENSPACE.f90:160-212): ascending energies, a
    new state goes before the first state of equal or higher energy."""
    pos = 0
    while pos < len(states) and not (st.energy <= states[pos].energy):
        pos += 1
    states.insert(pos, st)


def es_add_sized(states: list, st: State, size: int):
    """es_add_state with size (VCA_EIGENSPACE.f90:133-155)."""
    if len(states) < size:
        es_insert(states, st)
    elif st.energy < states[-1].energy:
        states.pop()
        es_insert(states, st)


@dataclass
class DiagResult:
    states: list
    zeta: float
    iters: int


def diagonalize_full(model: Model, matvec_factory=None) -> DiagResult:
    """vca_diag_c (VCA_DIAG.f90:49-267) with vca_post_diag (:333-474): the
    sector scan; at T = 0 the ground-state selection (:224-235); at finite T
    the size-limited state list and the spectrum annealing loop, which
    adapts the states per sector and the list size until
    exp(-beta (Emax - Emin)) < cutoff, trimming the list at the end."""
    p = model.p
    ft = p.finite_t
    nst_sector, nst_total = p.lanc_nstates_sector, p.lanc_nstates_total
    if ft:  # even counts, VCA_SETUP.f90:165-176
        nst_sector += nst_sector % 2
        nst_total += nst_total % 2
    neigen_sector = {}
    for nup in range(p.ns + 1):
        for ndw in range(p.ns + 1):
            du, dd = model.dims(nup, ndw)
            neigen_sector[(nup, ndw)] = min(du * dd, nst_sector)
    converged, it, zeta = False, 0, 1.0
    states: list[State] = []
    while not converged and it < max(1, p.lanc_niter_spectrum):
        it += 1
        states = []
        oldzero = 1000.0
        for (nup, ndw) in sector_list(p):
            du, dd = model.dims(nup, ndw)
            neigen = 1 if p.lanc_method == "lanczos" else min(du * dd, neigen_sector[(nup, ndw)])
            evals, evecs = solve_sector(model, nup, ndw, neigen, matvec_factory)
            for i in range(len(evals)):
                st = State(float(evals[i]), (nup, ndw), evecs[:, i].copy())
                if ft:
                    es_add_sized(states, st, nst_total)
                elif st.energy < oldzero - 10.0 * p.gs_threshold:
                    oldzero = st.energy
                    states = [st]
                elif abs(st.energy - oldzero) <= p.gs_threshold:
                    oldzero = min(oldzero, st.energy)
                    es_insert(states, st)
        egs, emax = states[0].energy, states[-1].energy
        converged = (math.exp(-p.beta * (emax - egs)) < p.cutoff) if ft else True
        if ft:
            zeta = sum(math.exp(-p.beta * (st.energy - egs)) for st in states)
            count = {}
            for st in states:
                count[st.sector] = count.get(st.sector, 0) + 1
            for key in neigen_sector:
                c = count.get(key, 0)
                ne = neigen_sector[key] + (1 if c > 0 else -1)
                if ne > c:
                    ne = c + 1
                neigen_sector[key] = max(ne, 1)
            if math.exp(-p.beta * (emax - egs)) > p.cutoff:
                nst_total += p.lanc_nstates_step
            else:
                while len(states) > 1 and math.exp(-p.beta * (states[-1].energy - egs)) <= p.cutoff:
                    states.pop()
                nst_total = max(len(states), p.lanc_nstates_step) + p.lanc_nstates_step
        else:
            zeta = float(len(states))
    return DiagResult(states, zeta, it)


def diagonalize(model: Model, matvec_factory=None):
    """The kept states of diagonalize_full."""
    return diagonalize_full(model, matvec_factory).states


# --------------------------------------------------------- GF chains -----
@dataclass
class GFData:
    """Lehmann store: impGmatrix(ilat,jlat,s,s,iorb,jorb)%state(:)%channel(:)
    with weights and poles (VCA_VARS_GLOBAL.f90:61-71), T=0."""
    p: Params
    emin: float
    zeta: float
    # key (ilat, jlat, ispin, iorb, jorb) -> list over states of list over
    # channels of (weights complex[Nexc], poles real[Nexc])
    elems: dict = field(default_factory=dict)
    max_exc: float = -np.inf


def _chain_poles(vnorm2, ei, alpha, beta, isign, emin, zeta, temp_beta=None):
    """add_to_lanczos_gf_normal, VCA_GF_NORMAL.f90:970-1022; at finite T
    (temp_beta given) the weights carry exp(-beta (Ei - Egs)) when
    beta (Ei - Egs) < 200, else 0 (:987-994)."""
    e, z = tridiag_eigh(alpha, beta)
    peso_bz = vnorm2 / zeta
    if temp_beta is not None:
        x = temp_beta * (ei - emin)
        peso_bz = peso_bz * (math.exp(-x) if x < 200 else 0.0)
    w = peso_bz * z[0, :] * z[0, :]
    de = e - ei
    return w.astype(np.complex128), isign * de, de.max()


def excitation(model: Model, st: State, ops):
    """O|psi> for O = sum_k coef_k * (c or c^+)_{pos_k, spin}, all ops of one
    type (VCA_GF_NORMAL.f90:184-198, 624-660).  ops: list of (spin, pos,
    dagger, coef).  Returns (target sector, vector) or (None, None)."""
    nup, ndw = st.sector
    spin, _, dagger, _ = ops[0]
    d = 1 if dagger else -1
    jnup, jndw = (nup + d, ndw) if spin == 1 else (nup, ndw + d)
    ns = model.ns
    if not (0 <= jnup <= ns and 0 <= jndw <= ns):
        return None, None
    du, dd = model.dims(jnup, jndw)
    out = np.zeros(du * dd, dtype=np.complex128)
    for (s, pos, dg, coef) in ops:
        model.excite(nup, ndw, st.vec, s, pos, dg, coef, out)
    return (jnup, jndw), out


def build_gf(model: Model, states: list, matvec_factory=None, zeta=None):
    """build_gf_normal (VCA_GF_NORMAL.f90:37-104) at T=0: diagonal chains
    c^+_i, c_i and, for i != j, the 4 mixed channels (c^+_i + c^+_j),
    (c_i + c_j), (c^+_i + i c^+_j) with weight -i norm2, (c_i - i c_j) with
    weight -i norm2; 2 channels when gf_symmetric."""
    p = model.p
    emin = min(s.energy for s in states)
    if zeta is None:
        zeta = float(len(states))
    gd = GFData(p, emin, zeta)
    tb = p.beta if p.finite_t else None

    def run_chain(st, ops, vscale, isign):
        jsec, v = excitation(model, st, ops)
        if jsec is None:
            return (np.zeros(0, np.complex128), np.zeros(0))
        norm2 = np.vdot(v, v).real
        if norm2 == 0.0:
            # O|psi> = 0: the reference would divide by zero
            # (VCA_GF_NORMAL.f90:201-203); an empty channel is its limit
            return (np.zeros(0, np.complex128), np.zeros(0))
        v = v / math.sqrt(norm2)
        du, dd = model.dims(*jsec)
        nlanc = min(du * dd, p.lanc_ngfiter)
        mv = (matvec_factory(*jsec) if matvec_factory
              else (lambda x, a=jsec[0], b=jsec[1]: model.hxv(a, b, x)))
        alpha, beta = lanczos_tridiag(mv, v, nlanc)
        w, poles, mx = _chain_poles(vscale * norm2, st.energy, alpha, beta, isign, emin, zeta, tb)
        gd.max_exc = max(gd.max_exc, mx)
        return (w, poles)

    for ispin in range(1, p.nspin + 1):
        for iorb in range(1, p.norb + 1):
            for isite in range(1, p.nlat + 1):
                is_ = iorb + (isite - 1) * p.norb
                per_state = []
                for st in states:
                    per_state.append([
                        run_chain(st, [(ispin, is_, True, 1.0)], 1.0, +1),
                        run_chain(st, [(ispin, is_, False, 1.0)], 1.0, -1)])
                gd.elems[(isite, isite, ispin, iorb, iorb)] = per_state
                for jsite in range(1, p.nlat + 1):
                    for jorb in range(1, p.norb + 1):
                        if isite == jsite and iorb == jorb:
                            continue
                        js = jorb + (jsite - 1) * p.norb
                        per_state = []
                        for st in states:
                            ch = [run_chain(st, [(ispin, is_, True, 1.0), (ispin, js, True, 1.0)], 1.0, +1),
                                  run_chain(st, [(ispin, is_, False, 1.0), (ispin, js, False, 1.0)], 1.0, -1)]
                            if not p.gf_symmetric:
                                ch.append(run_chain(st, [(ispin, is_, True, 1.0), (ispin, js, True, 1j)], -1j, +1))
                                ch.append(run_chain(st, [(ispin, is_, False, 1.0), (ispin, js, False, -1j)], -1j, -1))
                            per_state.append(ch)
                        gd.elems[(isite, jsite, ispin, iorb, jorb)] = per_state
    return gd


def gf_cluster(gd: GFData, z):
    """vca_gf_cluster_scalar, VCA_IO/gf_cluster.f90:1-64: pole sums, then the
    off-diagonal recombination G_ij = (G^mix - (1 - i chan4)(G_ii + G_jj))/2."""
    p = gd.p
    chan4 = 0.0 if p.gf_symmetric else 1.0
    g = zeros_nnn(p)
    for (il, jl, s, io, jo), per_state in gd.elems.items():
        acc = 0j
        for chans in per_state:
            for (w, poles) in chans:
                if len(w):
                    acc += np.sum(w / (z - poles))
        g[il - 1, jl - 1, s - 1, s - 1, io - 1, jo - 1] = acc
    for s in range(p.nspin):
        for io in range(p.norb):
            for jo in range(p.norb):
                for il in range(p.nlat):
                    for jl in range(p.nlat):
                        if il == jl and io == jo:
                            continue
                        g[il, jl, s, s, io, jo] = 0.5 * (g[il, jl, s, s, io, jo]
                                                         - (1 - chan4 * 1j) * g[il, il, s, s, io, io]
                                                         - (1 - chan4 * 1j) * g[jl, jl, s, s, jo, jo])
    return g


def delta_bath(model: Model, z):
    """delta_bath_freq_main, VCA_BATH_FUNCTIONS.f90:127-150:
    Delta = V (z + mu - h)^-1 V^dagger in lso order."""
    p = model.p
    if p.nlat_bath * p.norb_bath == 0:
        return np.zeros((p.nlso, p.nlso), dtype=np.complex128)
    hb = nnn2lso(model.bath_h, p.nlat_bath, p.nspin, p.norb_bath)
    vv = rect_n2j(model.bath_v, p.nlat, p.nlat_bath, p.nspin, p.norb, p.norb_bath)
    m = (z + p.xmu) * np.eye(hb.shape[0]) - hb
    return vv @ np.linalg.inv(m) @ vv.conj().T


def frequency_grids(p: Params):
    """allocate_grids, VCA_GF_SHARED.f90:43-56: wm = pi/beta (2n - 1),
    n = 1..Lmats; wr = linspace(wini, wfin, Lreal), SciFortran's form
    start + (i - 1) (stop - start)/(num - 1) (SciFortran is not shipped with
    the reference; its linspace with both end points is restated)."""
    wm = (math.pi / p.beta) * (2 * np.arange(1, p.lmats + 1) - 1).astype(np.float64)
    step = (p.wfin - p.wini) / float(p.lreal - 1)
    wr = p.wini + np.arange(p.lreal, dtype=np.float64) * step
    return wm, wr


def gf_cluster_grid(gd: GFData, zs):
    """impGmats / impGreal of build_gf_normal: the pole sums of
    add_to_lanczos_gf_normal (VCA_GF_NORMAL.f90:1024-1031) at every grid point,
    then the off-diagonal recombination (:89-101).  Returns
    [Nlat,Nlat,Nspin,Nspin,Norb,Norb,L]."""
    p = gd.p
    zs = np.asarray(zs, dtype=np.complex128)
    chan4 = 0.0 if p.gf_symmetric else 1.0
    g = zeros_nnn(p, (len(zs),))
    for (il, jl, s, io, jo), per_state in gd.elems.items():
        acc = np.zeros(len(zs), np.complex128)
        for chans in per_state:
            for (w, poles) in chans:
                if len(w):
                    acc += (w[None, :] / (zs[:, None] - poles[None, :])).sum(axis=1)
        g[il - 1, jl - 1, s - 1, s - 1, io - 1, jo - 1] = acc
    f = 1 - chan4 * 1j
    for s in range(p.nspin):
        for il in range(p.nlat):
            for jl in range(p.nlat):
                for io in range(p.norb):
                    for jo in range(p.norb):
                        if il == jl and io == jo:
                            continue
                        g[il, jl, s, s, io, jo] = 0.5 * (g[il, jl, s, s, io, jo]
                                                         - f * g[il, il, s, s, io, io]
                                                         - f * g[jl, jl, s, s, jo, jo])
    return g


def _nnn2lso_many(a, p: Params):
    """[Nlat,Nlat,Nspin,Nspin,Norb,Norb,L] -> [L, nlso, nlso] (lso index
    iorb + ilat Norb + ispin Norb Nlat, VCA_AUX_FUNX.f90:255-279)."""
    n = p.nlso
    # axes (il, jl, s, t, io, jo, z) -> (z, s, il, io, t, jl, jo)
    return np.ascontiguousarray(a.transpose(6, 2, 0, 4, 3, 1, 5)).reshape(a.shape[6], n, n)


def _lso2nnn_many(m, p: Params):
    L = m.shape[0]
    t = m.reshape(L, p.nspin, p.nlat, p.norb, p.nspin, p.nlat, p.norb)
    return np.asfortranarray(t.transpose(2, 5, 1, 4, 3, 6, 0))


def sigma_grid(model: Model, gd: GFData, axis):
    """build_gf_normal's grid G and build_sigma_normal
    (VCA_GF_NORMAL.f90:1045-1111) on the Matsubara (axis 0, z = i wm) or real
    (axis 1, z = wr + i eps) grid: G0^-1 = (z + xmu) 1 - Hloc - Delta(z),
    Sigma = G0^-1 - G^-1, G0 = (G0^-1)^-1.  Returns (G, Sigma, G0), each
    [Nlat,Nlat,Nspin,Nspin,Norb,Norb,L]."""
    p = model.p
    wm, wr = frequency_grids(p)
    zs = 1j * wm if axis == 0 else wr + 1j * p.eps
    g = gf_cluster_grid(gd, zs)
    n = p.nlso
    hl = nnn2lso(model.hloc, p.nlat, p.nspin, p.norb)
    g0inv = np.empty((len(zs), n, n), np.complex128)
    for i, z in enumerate(zs):
        g0inv[i] = (z + p.xmu) * np.eye(n) - hl - delta_bath(model, z)
    ginv = np.linalg.inv(_nnn2lso_many(g, p))
    return g, _lso2nnn_many(g0inv - ginv, p), _lso2nnn_many(np.linalg.inv(g0inv), p)


# ------------------------------------------------------------ pole file
def _gf_file_order(p: Params):
    """Loop order of save_gfprime / read_gfprime (VCA_AUX_FUNX.f90:633-647):
    ilat, jlat, ispin, iorb, jorb with jorb innermost (1-based keys)."""
    for il in range(1, p.nlat + 1):
        for jl in range(1, p.nlat + 1):
            for s in range(1, p.nspin + 1):
                for io in range(1, p.norb + 1):
                    for jo in range(1, p.norb + 1):
                        yield (il, jl, s, io, jo)


def save_gfprime(gd: GFData, path, formatted=False):
    """save_gfprime with the GFmatrix defined I/O (VCA_AUX_FUNX.f90:484-604,
    608-652).  Unformatted: one Fortran sequential record per element (the
    child transfers of a defined WRITE do not start records), 4-byte length
    markers, int32 counts, complex128 poles and weights.  Formatted:
    list-directed, one line per element, (re,im) complex items and the
    literal \\n item after each state."""
    import struct
    with open(path, "w" if formatted else "wb") as f:
        for key in _gf_file_order(gd.p):
            states = gd.elems.get(key, [])
            if formatted:
                items = [f"{len(states):12d}"]
                for chans in states:
                    items.append(f"{len(chans):12d}")
                    for (w, poles) in chans:
                        items.append(f"{len(poles):12d}")
                        items += [f"({complex(x).real:.17E},{complex(x).imag:.17E})" for x in poles]
                        items += [f"({complex(x).real:.17E},{complex(x).imag:.17E})" for x in w]
                    items.append("\\n")
                f.write(" ".join(items) + "\n")
            else:
                rec = [struct.pack("<i", len(states))]
                for chans in states:
                    rec.append(struct.pack("<i", len(chans)))
                    for (w, poles) in chans:
                        rec.append(struct.pack("<i", len(poles)))
                        rec.append(np.asarray(poles, dtype="<c16").tobytes())
                        rec.append(np.asarray(w, dtype="<c16").tobytes())
                body = b"".join(rec)
                f.write(struct.pack("<i", len(body)) + body + struct.pack("<i", len(body)))


def _fortran_list_tokens(text):
    """List-directed items of a text: blanks, commas and line ends separate;
    (re,im) is one item; r*item repeats; the \\n character items are
    dropped."""
    import re
    tok = re.compile(r"\s*(?:(\d+)\*)?(\([^)]*\)|[^\s,()]+)\s*,?")
    out = []
    for m in tok.finditer(text):
        item = m.group(2)
        if item == "\\n":
            continue
        out += [item] * (int(m.group(1)) if m.group(1) else 1)
    return out


def _fortran_complex(item):
    re_, im_ = item.strip("()").split(",")
    return complex(float(re_.replace("D", "E").replace("d", "e")),
                   float(im_.replace("D", "E").replace("d", "e")))


def read_gfprime(p: Params, path, formatted=False) -> GFData:
    """read_gfprime (VCA_AUX_FUNX.f90:654-697) into a GFData (emin and zeta
    are not in the file).  Poles keep their complex values."""
    import struct
    gd = GFData(p, 0.0, 1.0)
    if formatted:
        with open(path) as f:
            items = iter(_fortran_list_tokens(f.read()))
        for key in _gf_file_order(p):
            states = []
            for _ in range(int(next(items))):
                chans = []
                for _ in range(int(next(items))):
                    n = int(next(items))
                    poles = np.array([_fortran_complex(next(items)) for _ in range(n)])
                    w = np.array([_fortran_complex(next(items)) for _ in range(n)])
                    chans.append((w.astype(np.complex128), poles.astype(np.complex128)))
                states.append(chans)
            gd.elems[key] = states
        if next(items, None) is not None:
            raise ValueError("data after the last element")
        return gd
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    for key in _gf_file_order(p):
        (n1,) = struct.unpack_from("<i", data, pos)
        rec = data[pos + 4:pos + 4 + n1]
        (n2,) = struct.unpack_from("<i", data, pos + 4 + n1)
        if n1 != n2:
            raise ValueError("record markers differ")
        pos += n1 + 8
        q = 0

        def i32():
            nonlocal q
            (v,) = struct.unpack_from("<i", rec, q)
            q += 4
            return v

        def c16(n):
            nonlocal q
            a = np.frombuffer(rec, dtype="<c16", count=n, offset=q).astype(np.complex128)
            q += 16 * n
            return a

        states = []
        for _ in range(i32()):
            chans = []
            for _ in range(i32()):
                n = i32()
                poles = c16(n)
                w = c16(n)
                chans.append((w, poles))
            states.append(chans)
        if q != len(rec):
            raise ValueError("record longer than its counts")
        gd.elems[key] = states
    if pos != len(data):
        raise ValueError("records beyond the last element")
    return gd


def sum_kmesh(model: Model, gd: GFData, hk_lso, omega, complex_out=False):
    """sum_kmesh / sum_kmesh_complex, VCA_OMEGA.f90:37-116:
    (1/Nk) sum_k log|det(1 + (Delta + Hloc - H_k) G')| in lso order."""
    p = model.p
    g = nnn2lso(gf_cluster(gd, omega), p.nlat, p.nspin, p.norb)
    hl = nnn2lso(model.hloc, p.nlat, p.nspin, p.norb)
    dl = delta_bath(model, omega)
    n = p.nlso
    a = (dl + hl)[None, :, :] - hk_lso
    m = np.eye(n)[None] + a @ g[None]
    sign, logabs = np.linalg.slogdet(m)
    if complex_out:
        return np.mean(logabs + np.log(sign))
    return float(np.mean(logabs))


# ------------------------------------------------------------ quadrature
_GK15_X = np.array([0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                    0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                    0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                    0.207784955007898467600689403773245, 0.0])
_GK15_WK = np.array([0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                     0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                     0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                     0.204432940075298892414161999234649, 0.209482141084727828012999174891714])
_GK15_WG = np.array([0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                     0.381830050505118944950369775488975, 0.417959183673469387755102040816327])


def gk15_nodes(a, b):
    """Gauss-Kronrod 7/15 nodes and weights on [a, b] (Kronrod 1965)."""
    c, h = 0.5 * (a + b), 0.5 * (b - a)
    x = np.concatenate([c - h * _GK15_X[:7], [c], c + h * _GK15_X[6::-1]])
    wk = h * np.concatenate([_GK15_WK[:7], [_GK15_WK[7]], _GK15_WK[6::-1]])
    wg = np.zeros(15)
    # Gauss nodes are the odd-indexed Kronrod abscissae 1,3,5 and the centre
    for gi, ki in enumerate([1, 3, 5]):
        wg[ki] = h * _GK15_WG[gi]
        wg[14 - ki] = h * _GK15_WG[gi]
    wg[7] = h * _GK15_WG[3]
    return x, wk, wg


def quad_semi_infinite(f_batch, epsabs=1e-10, epsrel=1e-10, limit=400):
    """Adaptive G-K 7/15 on [0, inf) after x = (1-t)/t (the QAGI map used by
    QUADPACK's qagi, which the reference calls through SciFortran quad at
    VCA_OMEGA.f90:132).  f_batch evaluates the integrand on an array of
    points.  Bisects the interval of largest error estimate until the total
    error estimate is below max(epsabs, epsrel*|I|)."""
    def panel(a, b):
        t, wk, wg = gk15_nodes(a, b)
        x = (1.0 - t) / t
        fx = np.asarray(f_batch(x)) / (t * t)
        ik, ig = float(wk @ fx), float(wg @ fx)
        return ik, abs(ik - ig)

    panels = [(0.0, 1.0) + panel(0.0, 1.0)]
    for _ in range(limit):
        tot = sum(pn[2] for pn in panels)
        err = sum(pn[3] for pn in panels)
        if err <= max(epsabs, epsrel * abs(tot)):
            break
        k = int(np.argmax([pn[3] for pn in panels]))
        a, b, _, _ = panels.pop(k)
        m = 0.5 * (a + b)
        panels.append((a, m) + panel(a, m))
        panels.append((m, b) + panel(m, b))
    return sum(pn[2] for pn in panels), sum(pn[3] for pn in panels)


def frequency_integration(model, gd, hk_lso, epsabs=1e-10, epsrel=1e-12):
    """VCA_OMEGA.f90:123-146: (3-Nspin)/pi * int_0^inf sum_kmesh(i zeta)."""
    f = lambda xs: np.array([sum_kmesh(model, gd, hk_lso, 1j * x) for x in xs])
    val, _ = quad_semi_infinite(f, epsabs, epsrel)
    return (3 - model.p.nspin) * val / math.pi


def matsubara_integration(model, gd, hk_lso, beta, nfreq):
    """Fixed-grid form: (3-Nspin) * (2/beta) * sum_{n<nfreq} f(i w_n),
    w_n = (2n+1) pi/beta, the midpoint rule of step 2pi/beta for
    (3-Nspin)/pi int_0^inf f (the Matsubara-sum part of
    VCA_OMEGA.f90:153-177)."""
    wn = (2 * np.arange(nfreq) + 1) * math.pi / beta
    s = sum(sum_kmesh(model, gd, hk_lso, 1j * w) for w in wn)
    return (3 - model.p.nspin) * 2.0 / beta * s


def quad_finite(f_batch, a, b, epsabs=1e-10, epsrel=1e-10, limit=400):
    """Adaptive G-K 7/15 on [a, b] (QUADPACK qag-style bisection of the
    interval with the largest error estimate; the reference calls quad
    with key=6 at VCA_OMEGA.f90:179, whose converged value this matches at
    the tolerance)."""
    def panel(lo, hi):
        t, wk, wg = gk15_nodes(lo, hi)
        fx = np.asarray(f_batch(t))
        ik, ig = float(wk @ fx), float(wg @ fx)
        return ik, abs(ik - ig)

    panels = [(a, b) + panel(a, b)]
    for _ in range(limit):
        tot = sum(pn[2] for pn in panels)
        err = sum(pn[3] for pn in panels)
        if err <= max(epsabs, epsrel * abs(tot)):
            break
        k = int(np.argmax([pn[3] for pn in panels]))
        lo, hi, _, _ = panels.pop(k)
        m = 0.5 * (lo + hi)
        panels.append((lo, m) + panel(lo, m))
        panels.append((m, hi) + panel(m, hi))
    return sum(pn[2] for pn in panels), sum(pn[3] for pn in panels)


def finite_t_nmax(max_exc, bandwidth, beta):
    """VCA_OMEGA.f90:157-164: Nmax and the contour radius R."""
    nmax = int(2 * (abs(max_exc) + bandwidth) * beta / math.pi)
    nmax = nmax // 2 if nmax % 2 == 0 else (nmax + 1) // 2
    return nmax, 2 * (nmax + 1) * math.pi / beta


def frequency_integration_finite_t(model, gd, hk_lso, epsabs=1e-10, epsrel=1e-12):
    """VCA_OMEGA.f90:153-204: (3 - Nspin) [ (2/beta) sum_{n=0}^{Nmax}
    Re F(i w_n) + int_0^pi Re((1/pi) w f(w) F(w)) dzeta ], w = R e^{i zeta},
    f the Fermi function (0 where Re(beta (w - mu)) >= 100), F the complex
    sum_kmesh."""
    p = model.p
    beta = p.beta
    nmax, radius = finite_t_nmax(gd.max_exc, p.bandwidth, beta)
    s = sum(sum_kmesh(model, gd, hk_lso, 1j * (2 * n + 1) * math.pi / beta, complex_out=True).real
            for n in range(nmax + 1))
    s *= 2.0 / beta

    def g(zs):
        out = []
        for zeta in zs:
            w = radius * np.exp(1j * zeta)
            fermi = 0.0 if ((w - p.xmu) * beta).real >= 100 else 1.0 / (np.exp(beta * (w - p.xmu)) + 1.0)
            out.append(((1.0 / math.pi) * w * fermi *
                        sum_kmesh(model, gd, hk_lso, w, complex_out=True)).real)
        return np.array(out)

    contour, _ = quad_finite(g, 0.0, math.pi, epsabs, epsrel)
    return (3 - p.nspin) * (s + contour), nmax


def hk_to_lso(p: Params, hk):
    nk = hk.shape[6]
    return np.stack([nnn2lso(hk[..., k], p.nlat, p.nspin, p.norb) for k in range(nk)])


def vca_solve(p: Params, hloc, hk, bath_h=None, bath_v=None, omega_mode="quad",
              nfreq=4096, matvec_factory=None, epsabs=1e-10, epsrel=1e-12):
    """vca_solve_serial at T=0 (VCA_MAIN.f90:124-203): diagonalise, build the
    GF, integrate, Omega = E0 - integral."""
    model = Model(p, hloc, bath_h, bath_v)
    dr = diagonalize_full(model, matvec_factory)
    states = dr.states
    gd = build_gf(model, states, matvec_factory, zeta=dr.zeta)
    hkl = hk_to_lso(p, np.asarray(hk))
    e0 = min(s.energy for s in states)
    nmax = None
    if p.finite_t:
        integ, nmax = frequency_integration_finite_t(model, gd, hkl, epsabs, epsrel)
        # VCA_MAIN.f90:186-190
        omegaprime = e0 - math.log(sum(math.exp(-p.beta * (s.energy - e0)) for s in states)) / p.beta
    elif omega_mode == "quad":
        integ = frequency_integration(model, gd, hkl, epsabs, epsrel)
        omegaprime = e0
    else:
        integ = matsubara_integration(model, gd, hkl, p.beta, nfreq)
        omegaprime = e0
    return dict(sft_potential=omegaprime - integ, e0=e0, omegaprime=omegaprime,
                omega_integral=integ, states=states, gf=gd, model=model, zeta=dr.zeta,
                nmax=nmax, spectrum_iters=dr.iters)
