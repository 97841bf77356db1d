"""Algorithmic work model of the three-centre sweep (for roofline figures).

The unit of work is one screened three-centre evaluation (primitive pair x
potential image x term).  Its FP64 operation count is the count of the
formulation the kernels evaluate, per angular class (la, lb), counting an
FMA as two operations and each special function (exp, erf, sqrt, reciprocal)
as one; the special functions are reported separately as evaluations.

three-Gaussian overlap (electron3CIntg formulation):
    36                       centre G, GA, GB, |P-C|^2, prefactor
  + 9 ((la+1)(lb+1) - 1)     three 1-D Obara-Saika tables
  + 4 NA NB                  Cartesian products and accumulation
nuclear (nuclear3CIntg formulation, L = la + lb):
    60 + 12 (L+1)            centres, exponent, Boys F_0..F_L
  + 8 sum_{a,b} (L - |a| - |b| + 1)   two-index recursion entries
  + 2 NA NB                  accumulation
with NA, NB the Cartesian component counts (1, 4, 10, 20 up to l).
"""
from __future__ import annotations

import numpy as np

NCART = [1, 4, 10, 20]


def _cart_tot(l):
    out = []
    for t in range(l + 1):
        out += [t] * ((t + 1) * (t + 2) // 2)
    return out


def e3c_flops(la, lb):
    return 36 + 9 * ((la + 1) * (lb + 1) - 1) + 4 * NCART[la] * NCART[lb]


def nuc_flops(la, lb):
    L = la + lb
    rec = sum(L - ta - tb + 1 for ta in _cart_tot(la) for tb in _cart_tot(lb))
    return 60 + 12 * (L + 1) + 8 * rec + 2 * NCART[la] * NCART[lb]


def sweep_flops(counts):
    """counts: [4, 4, 2] evaluations per class -> total FP64 flops."""
    counts = np.asarray(counts, dtype=np.float64)
    tot = 0.0
    for la in range(4):
        for lb in range(4):
            tot += counts[la, lb, 0] * e3c_flops(la, lb) + counts[la, lb, 1] * nuc_flops(la, lb)
    return tot
