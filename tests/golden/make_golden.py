"""Write golden vectors for the primitive integrals and the Boys function.

Evaluates the reference's own primitive-integral subroutines
(src/uolcao/gaussIntegrals.f90: overlap2CIntg :26, kinetic2CIntg :4422,
electron3CIntg :11624, boys :16080, nuclear3CIntg :16223, massvel2CIntg
:39070), compiled from the reference sources by `make -C oracle ref` into
oracle/_ref/libolcao_ref_prims.so, on seeded inputs, and stores inputs and
outputs in prims_golden.npz.  Run in the build container (the reference is
not present on GPU machines):

    make -C oracle ref && python tests/golden/make_golden.py

Sampling: for every (l1, l2) class in {s,p,d,f}^2 and every operator, the
exponents are log-uniform in [0.05, 1e7] and separations are drawn so the
Gaussian-product exponent xi d^2 (and for three-centre integrals the second
exponent mu |P-C|^2) lies inside the 1e-16 screening range, i.e. the values
the integral sweeps actually use.  Boys arguments cover both branches,
including the float(0.1) branch point.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref", "libolcao_ref_prims.so")
OUT = os.path.join(HERE, "prims_golden.npz")

OPS = {"overlap": 0, "kinetic": 1, "electron3c": 2, "nuclear3c": 3, "massvel": 4}
CASES_PER_CLASS = 10


def main():
    lib = C.CDLL(os.path.abspath(REF))
    dp = C.POINTER(C.c_double)
    lib.ref_prim.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, dp, dp, dp,
                             C.c_int, dp]
    lib.ref_boys.argtypes = [C.c_double, dp]
    rng = np.random.default_rng(20261203)
    out = {}
    for name, op in OPS.items():
        rows = []
        for la in range(4):
            for lb in range(4):
                sw = (1 << la) + 16 * (1 << lb)
                for case in range(CASES_PER_CLASS):
                    a1, a2 = np.exp(rng.uniform(np.log(0.05), np.log(1e7), 2))
                    if la > 0 or lb > 0:   # diffuse-enough exponents for l > 0
                        a1, a2 = np.exp(rng.uniform(np.log(0.05), np.log(2e5), 2))
                    a3 = np.exp(rng.uniform(np.log(0.1), np.log(1e5)))
                    if op == 3:
                        a3 = rng.choice([20.0, np.exp(rng.uniform(np.log(0.5), np.log(50)))])
                    xi = a1 * a2 / (a1 + a2)
                    d = rng.normal(size=3)
                    d /= np.linalg.norm(d)
                    if case == 0:
                        dist = 0.0
                    else:
                        dist = np.sqrt(rng.uniform(0.0, 30.0) / xi)
                    A = rng.uniform(-5, 5, 3)
                    B = A - d * dist
                    P = (a1 * A + a2 * B) / (a1 + a2)
                    z = a1 + a2
                    mu = z * a3 / (z + a3)
                    e = rng.normal(size=3)
                    e /= np.linalg.norm(e)
                    Cc = P + e * np.sqrt(rng.uniform(0.0, 25.0) / mu)
                    if case == 1:
                        Cc = A.copy()
                    sh = np.zeros((16, 16), order="F")
                    Ab, Bb, Cb = (np.ascontiguousarray(x, np.float64) for x in (A, B, Cc))
                    lib.ref_prim(op, a1, a2, a3, Ab.ctypes.data_as(dp), Bb.ctypes.data_as(dp),
                                 Cb.ctypes.data_as(dp), sw, sh.ctypes.data_as(dp))
                    rows.append((la, lb, sw, a1, a2, a3, *A, *B, *Cc, sh.copy(order="C")))
        out[name + "_cls"] = np.array([r[:3] for r in rows], np.int32)
        out[name + "_in"] = np.array([r[3:15] for r in rows], np.float64)
        out[name + "_sh"] = np.array([r[15] for r in rows], np.float64)
    T = np.concatenate([
        [0.0, 1e-12, 1e-6, 1e-3, 0.05, 0.0999999, 0.1, 0.10000000149011612,
         0.1000001, 0.1001, 0.11, 0.2, 0.5, 1.0, 2.0, 5.0, 10.0, 30.0, 80.0, 300.0],
        np.exp(rng.uniform(np.log(1e-8), np.log(200.0), 60))])
    F = np.zeros((len(T), 7))
    for i, t in enumerate(T):
        f = np.zeros(7)
        lib.ref_boys(float(t), f.ctypes.data_as(dp))
        F[i] = f
    out["boys_T"] = T
    out["boys_F"] = F
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
