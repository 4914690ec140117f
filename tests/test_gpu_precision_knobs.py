"""GPU: the chain and Omega precision of a solve set apart from the solver's
(VCA_CHAIN_PRECISION, VCA_OMEGA_PRECISION).

- A solve's own precision named explicitly changes nothing: bit for bit.
- A MIXED solve with FP64 chains: its chains run on FP64 vectors from the
  MIXED ground state, so against the FP64 solve G' differs only through
  the ground-state vector (within 2e-6 of its largest element on the
  10-site chain, 80-step chains: 2.5e-7 seen), closer than with FP32
  chains (1.8e-6 seen).
- A MIXED solve with FP64 log-dets: its Matsubara-sum integral is the FP64
  integrand of its own poles, summed (1e-12), where FP32 log-dets differ
  from it.
- Any other value is refused with VCA_ERR_ARG."""
import math

import numpy as np
import pytest

from vca_amd import configs
from vca_amd._lib import VcaError
from vca_amd.api import VcaSolver

from . import _models as M

pytestmark = pytest.mark.gpu

ZS = (0.3j, 1.7j, 0.5 + 0.2j)


def _say(*a):
    print(*a, flush=True)


def _chain10(prec):
    p, hloc, hk, _, _ = M.chain(nlat=10, u=4.0, nk=4)
    p.lanc_ngfiter = 80
    p.precision = prec
    return p, hloc, hk


def _gf(p, hloc, hk):
    s = VcaSolver(p)
    om = s.vca_solve(hloc, hk)
    return om, [s.vca_gf_cluster(z) for z in ZS]


def test_chain_precision_knob(monkeypatch):
    out = {}
    for name, prec, knob in (("fp64", "fp64", None), ("fp64_named", "fp64", "fp64"),
                             ("mixed", "mixed", None), ("mixed_named", "mixed", "mixed"),
                             ("mixed_fp64_chains", "mixed", "fp64")):
        if knob is None:
            monkeypatch.delenv("VCA_CHAIN_PRECISION", raising=False)
        else:
            monkeypatch.setenv("VCA_CHAIN_PRECISION", knob)
        _say(f"chain10 solve: {name}")
        out[name] = _gf(*_chain10(prec))
    for a, b in (("fp64", "fp64_named"), ("mixed", "mixed_named")):
        assert out[a][0] == out[b][0]
        for x, y in zip(out[a][1], out[b][1]):
            assert np.array_equal(x, y)

    def gerr(name):
        return max(np.abs(x - y).max() / np.abs(y).max()
                   for x, y in zip(out[name][1], out["fp64"][1]))

    e64, e32 = gerr("mixed_fp64_chains"), gerr("mixed")
    _say(f"  G' against FP64: FP64 chains {e64:.1e}, FP32 chains {e32:.1e}")
    assert e64 <= 2e-6 and e64 < e32, (e64, e32)
    monkeypatch.setenv("VCA_CHAIN_PRECISION", "fp16")
    with pytest.raises(VcaError):
        _gf(*_chain10("fp64"))


def test_omega_precision_knob(monkeypatch):
    p, hloc, hk, _, _ = configs.config_chain4(nk=16)
    p.precision = "mixed"
    p.omega_mode = "matsubara"
    zs = 1j * (2 * np.arange(p.omega_nfreq) + 1) * math.pi / p.beta
    res = {}
    for knob in (None, "fp64"):
        if knob is None:
            monkeypatch.delenv("VCA_OMEGA_PRECISION", raising=False)
        else:
            monkeypatch.setenv("VCA_OMEGA_PRECISION", knob)
        _say(f"chain4 MIXED solve, VCA_OMEGA_PRECISION={knob}")
        s = VcaSolver(p)
        s.vca_solve(hloc, hk)
        i64 = (3 - p.nspin) * 2.0 / p.beta * s.omega_integrand(zs, "fp64").sum()
        res[knob] = (s.results()["omega_integral"], i64)
    integral, i64 = res["fp64"]
    assert abs(integral - i64) <= 1e-12 * abs(i64), (integral, i64)
    integral32, i64_b = res[None]
    assert i64_b == i64  # the same poles either way
    assert integral32 != i64
    monkeypatch.setenv("VCA_OMEGA_PRECISION", "bf16")
    with pytest.raises(VcaError):
        VcaSolver(p).vca_solve(hloc, hk)
