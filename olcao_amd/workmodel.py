"""Algorithmic work model of the kernels the Mode-B iteration runs, for the
roofline figures.

FP64 operations are counted from the device code itself (the
McMurchie-Davidson formulation of csrc/prims.hpp that the three-centre
kernels evaluate), an FMA as two operations, a multiplication, addition or
subtraction as one; divisions, square roots, erf and the libm exponential
are counted separately as special-function evaluations (`*_specials`), the
table-driven exponential of the three-Gaussian items by its own arithmetic
(one rint, seven FMAs, two multiplications).  L = la + lb is the angular
momentum of the primitive pair; nherm(L) = (L+1)(L+2)(L+3)/6 the number of
Hermite sums R_tuv it accumulates.

Three-Gaussian item (acc_e3c_herm_tab<L>), one potential image x term:
    8                           X = P - C and |X|^2
  + 19                          mu X^2, exponential, prefactor
  + [L >= 1] (4 + 12 (L - 1))   Hermite polynomials h_k(X_x), three axes
  + (L+1)(L+2)                  products pref h_t(X) h_u(Y)
  + 2 nherm(L)                  W_tuv += . h_v(Z) (FMA)
Nuclear item (acc_nuc_herm<L>), one potential image:
    12                          exponents mu, kappa, X, |X|^2
  + boys(L)                     Boys F_0..F_L, the reference's closed forms
  + 3 + 2 L + 3 (L+1) + 3 (L+1)(L+2)   prefactor, powers, R^(n)_000
  + 3 sum_{m=1..L} (nherm(m) - 1)      the R^(n)_tuv recursion
  + nherm(L)                    accumulation
  specials: 1 reciprocal, exp, sqrt, erf, and the Boys divisions
Finish (k_herm_finish<la,lb>), once per primitive pair: the Hermite
expansion tables E (three axes) and every Cartesian component's sum
K sum_tuv E E E W, then the Cartesian -> solid-harmonic transform.

The tridiagonal reduction (k_trd) of a complex Hermitian n x n matrix does
(16/3) n^3 real flops (4 (4/3) n^3, zhetrd), a real symmetric one
(4/3) n^3 (dsytrd).
"""
from __future__ import annotations

import numpy as np

NCART = [1, 4, 10, 20]
NSH = [1, 4, 9, 16]


def nherm(L):
    return (L + 1) * (L + 2) * (L + 3) // 6


def e3c_flops(la, lb):
    L = la + lb
    h = 4 + 12 * (L - 1) if L >= 1 else 0
    return 8 + 19 + h + (L + 1) * (L + 2) + 2 * nherm(L)


def _boys_flops(L):
    """boys<L> closed-form branch (T > 0.1, the lattice sums' common case):
    F_0 two products; per m: the power T^m (m - 1 products), lead (three
    products), the inner sum over p (a product and an addition each), the
    difference (one FMA) and the scaling (one product for m > 1)."""
    return 2 + sum((m - 1) + 3 + 2 * m + 2 + (1 if m > 1 else 0) for m in range(1, L + 1))


def _boys_specials(L):
    return 4 + sum(1 + m for m in range(1, L + 1))


def nuc_flops(la, lb):
    L = la + lb
    return (12 + _boys_flops(L) + 3 + 2 * L + 3 * (L + 1) + 3 * (L + 1) * (L + 2)
            + 3 * sum(nherm(m) - 1 for m in range(1, L + 1)) + nherm(L))


def nuc_specials(la, lb):
    return 2 + _boys_specials(la + lb)


def _carts(l):
    return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]


def finish_flops(la, lb):
    """Per primitive pair: E tables (3 axes, ~5 flops per entry), the
    Cartesian sums over the components of every shell up to la and lb, the
    solid-harmonic transform (2 flops per Cartesian x harmonic pair)."""
    L = la + lb
    etab = 3 * sum(5 * (i + j + 1) for i in range(la + 1) for j in range(lb + 1))
    cart = 0
    for a in _carts(la):
        for b in _carts(lb):
            tx, ty, tz = a[0] + b[0] + 1, a[1] + b[1] + 1, a[2] + b[2] + 1
            cart += tx * ty * (1 + 2 * tz) + 1
    trans = 2 * NCART[la] * NCART[lb] * (NSH[la] + NSH[lb])
    return etab + cart + trans + 4 * L


def sweep_flops(counts, finish_pairs=None):
    """counts: [4, 4, 2] three-Gaussian and nuclear evaluations per class
    (olcao_get_work_counts); finish_pairs: [4, 4] primitive pairs per class
    (optional) -> total FP64 flops of the three-centre sweep."""
    counts = np.asarray(counts, dtype=np.float64).reshape(4, 4, 2)
    tot = 0.0
    for la in range(4):
        for lb in range(4):
            tot += counts[la, lb, 0] * e3c_flops(la, lb) + counts[la, lb, 1] * nuc_flops(la, lb)
            if finish_pairs is not None:
                tot += float(np.asarray(finish_pairs).reshape(4, 4)[la, lb]) * finish_flops(la, lb)
    return tot


def sweep_specials(counts):
    counts = np.asarray(counts, dtype=np.float64).reshape(4, 4, 2)
    return float(sum(counts[la, lb, 1] * nuc_specials(la, lb)
                     for la in range(4) for lb in range(4)))


def hetrd_flops(n, nk, cplx=True):
    """Real flops of the tridiagonal reductions of nk matrices."""
    return (16.0 / 3.0 if cplx else 4.0 / 3.0) * float(n) ** 3 * nk


def table():
    """Per-class counts, as committed in the profiles."""
    rows = []
    for la in range(4):
        for lb in range(4):
            rows.append((la, lb, e3c_flops(la, lb), nuc_flops(la, lb), nuc_specials(la, lb),
                         finish_flops(la, lb)))
    return rows
