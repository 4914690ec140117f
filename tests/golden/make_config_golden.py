"""Writes the oracle fixtures of two workloads too large for the oracle to
run inside a test:

config_2x5.json   the 10-site 2x5 cluster, U=6, Nspin=2, seeded on-site and
                  Weiss fields, every (Nup, Ndw) sector with up to 6
                  eigenpairs (ARPACK semantics), 64 k-points: the lowest
                  eigenvalues of every sector above the dense threshold, the
                  kept ground states and zeta, G'(z) at four points, the
                  Omega integrand at five frequencies and Omega (adaptive
                  quadrature).
config_bhz2x2.json  the 2x2 BHZ cluster (Norb=2, Nspin=2, M=2, lambda=0.3,
                  U=2), sector (4,4): E0, G'(z) at four points and the Omega
                  integrand over the 128x128 k-grid at eight Matsubara
                  frequencies (beta=200) from the lowest to the last of 8192.

Both come from the oracle (oracle/vca_ref.py with the C loops of
oracle/vca_oracle.c, built by `make -C oracle`).  Complex numbers are stored
as [re, im].  Run:  python tests/golden/make_config_golden.py [2x5|bhz2x2]
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import vca_ref as R  # noqa: E402
from tests._models import to_oracle  # noqa: E402
from vca_amd import configs  # noqa: E402

GF_Z = [0.3j, 1.7j, 0.5 + 0.2j, -2.0 + 1.0j]


def _c(a):
    a = np.asarray(a, dtype=np.complex128)
    return np.stack([a.real, a.imag], axis=-1).tolist()


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def _matvec_factory(model, workers):
    return lambda nup, ndw: (lambda x: model.hxv(nup, ndw, x, workers=workers))


def golden_2x5():
    p, hloc, hk, _, _ = configs.config_batched_2x5(nk=8)
    po = to_oracle(p)
    model = R.Model(po, hloc)
    workers = os.cpu_count() or 1
    mvf = _matvec_factory(model, workers)
    sector_evals = {}
    solve = R.solve_sector

    def recording(model_, nup, ndw, neigen, matvec_factory=None):
        e, z = solve(model_, nup, ndw, neigen, matvec_factory)
        sector_evals[f"{nup},{ndw}"] = [float(x) for x in e]
        return e, z

    R.solve_sector = recording
    t0 = time.time()
    dr = R.diagonalize_full(model, mvf)
    R.solve_sector = solve
    _log(f"2x5: diagonalised in {time.time() - t0:.0f}s, {len(dr.states)} states")
    t0 = time.time()
    gd = R.build_gf(model, dr.states, mvf, zeta=dr.zeta)
    _log(f"2x5: chains in {time.time() - t0:.0f}s")
    hkl = R.hk_to_lso(po, hk)
    zs = [0.05j, 0.4j, 1.0j, 3.0j, 20.0j]
    integrand = [R.sum_kmesh(model, gd, hkl, z) for z in zs]
    integ = R.frequency_integration(model, gd, hkl, 1e-11, 1e-12)
    e0 = min(s.energy for s in dr.states)
    return {
        "workload": "2x5 cluster U=6 Nspin=2, configs.config_batched_2x5(nk=8), every sector",
        "e0": e0, "zeta": dr.zeta,
        "states": [[s.energy, s.sector[0], s.sector[1]] for s in dr.states],
        "sector_evals": sector_evals,
        "gf_z": _c(GF_Z), "gf": [_c(R.gf_cluster(gd, z)) for z in GF_Z],
        "integrand_z": _c(zs), "integrand": integrand,
        "omega_integral": integ, "sft_potential": e0 - integ,
    }


def golden_bhz2x2():
    p, hloc, hk, _, _ = configs.config_bhz_2x2(nk=128, nfreq=8192, precision="fp64")
    po = to_oracle(p)
    model = R.Model(po, hloc)
    t0 = time.time()
    dr = R.diagonalize_full(model)
    gd = R.build_gf(model, dr.states, zeta=dr.zeta)
    _log(f"bhz2x2: solved in {time.time() - t0:.0f}s")
    hkl = R.hk_to_lso(po, hk)
    nidx = [0, 1, 3, 10, 40, 200, 1000, 8191]
    wn = [(2 * n + 1) * np.pi / po.beta for n in nidx]
    integrand = [R.sum_kmesh(model, gd, hkl, 1j * w) for w in wn]
    e0 = min(s.energy for s in dr.states)
    return {
        "workload": "BHZ 2x2 Norb=2 Nspin=2 M=2 lambda=0.3 U=2, sector (4,4), "
                    "configs.config_bhz_2x2(nk=128, nfreq=8192), beta=200",
        "e0": e0, "nstates": len(dr.states),
        "gf_z": _c(GF_Z), "gf": [_c(R.gf_cluster(gd, z)) for z in GF_Z],
        "matsubara_n": nidx, "integrand": integrand,
    }


def main():
    which = sys.argv[1:] or ["2x5", "bhz2x2"]
    for w in which:
        out = {"2x5": golden_2x5, "bhz2x2": golden_bhz2x2}[w]()
        path = os.path.join(HERE, f"config_{w}.json")
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        _log(f"wrote {path}")


if __name__ == "__main__":
    main()
