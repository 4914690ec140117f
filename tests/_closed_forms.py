"""Closed forms shared by the oracle pins and the GPU tests."""
import math

import numpy as np


def lncosh(x):
    x = np.abs(x)
    return x + np.log1p(np.exp(-2.0 * x)) - math.log(2.0)


def matsubara_tail(c, beta, nfreq, terms=12):
    """sum_{n >= nfreq} ln(1 + c / w_n^2), w_n = (2n+1) pi / beta, by the
    series of ln(1+x) with sum_{n>=N} w_n^{-2k} = (beta/2pi)^{2k} zeta(2k, N+1/2)
    (Hurwitz zeta): exact for c < w_N^2."""
    from scipy.special import zeta
    c = np.asarray(c, dtype=float)
    assert np.all(c < ((2 * nfreq + 1) * math.pi / beta) ** 2)
    out = np.zeros_like(c)
    for k in range(1, terms + 1):
        zk = (beta / (2 * math.pi)) ** (2 * k) * zeta(2 * k, nfreq + 0.5)
        out += (-1) ** (k + 1) * c ** k / k * zk
    return out


def matsubara_u0_exact(hloc_lso, hk_lso, nspin, beta, nfreq):
    """(3-Nspin) (2/beta) sum_{n<N} f(i w_n) at U=0, where
    f(w) = (1/Nk) sum_k sum_j (1/2) ln[(w^2 + a_kj^2) / (w^2 + b_j^2)], a the
    eigenvalues of H_k, b those of Hloc.  The infinite sum is closed by
    cosh(x/2) = prod_{n>=0} (1 + x^2 / ((2n+1)^2 pi^2)):
    (2/beta) sum_{n>=0} (1/2) ln[(w_n^2+a^2)/(w_n^2+b^2)]
      = (1/beta) ln[cosh(beta a/2) / cosh(beta b/2)],
    and the part beyond N frequencies is subtracted in closed form."""
    b = np.linalg.eigvalsh(hloc_lso)
    tb = (lncosh(beta * b / 2) - matsubara_tail(b * b, beta, nfreq)).sum()
    tot = 0.0
    for h in hk_lso:
        a = np.linalg.eigvalsh(h)
        tot += (lncosh(beta * a / 2) - matsubara_tail(a * a, beta, nfreq)).sum() - tb
    return (3 - nspin) * tot / len(hk_lso) / beta
