"""Write golden vectors for the k-overlap primitive.

Evaluates the reference's own KOverlap2CIntg (src/uolcao/gaussIntegrals.f90
:129408, the block of exp(-i deltaK . r) between two Gaussians, complex),
compiled from the reference sources by `make -C oracle ref` into
oracle/_ref/libolcao_ref_prims.so, on seeded inputs, and stores inputs and
the real and imaginary parts of the blocks in prims_k_golden.npz.  Run in
the build container (the reference is not present on GPU machines):

    make -C oracle ref && python tests/golden/make_golden_k.py

Sampling: for every (l1, l2) class in {s,p,d,f}^2, exponents log-uniform in
[0.05, 2e5] (the diffuse end for l > 0), separations inside the 1e-16
screening range, and deltaK vectors of length 0 (the plain overlap) and
0.05..4 inverse bohr in random directions (a reciprocal vector over the
k-mesh divisions along its axis is 0.1..2 for the cells here).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref", "libolcao_ref_prims.so")
OUT = os.path.join(HERE, "prims_k_golden.npz")
CASES_PER_CLASS = 10


def main():
    lib = C.CDLL(os.path.abspath(REF))
    dp = C.POINTER(C.c_double)
    lib.ref_primk.argtypes = [C.c_double, C.c_double, dp, dp, dp, C.c_int, dp, dp]
    rng = np.random.default_rng(20261204)
    cls, inp, shr, shi = [], [], [], []
    for la in range(4):
        for lb in range(4):
            sw = (1 << la) + 16 * (1 << lb)
            for case in range(CASES_PER_CLASS):
                hi = 1e7 if la == 0 and lb == 0 else 2e5
                a1, a2 = np.exp(rng.uniform(np.log(0.05), np.log(hi), 2))
                xi = a1 * a2 / (a1 + a2)
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                dist = 0.0 if case == 0 else np.sqrt(rng.uniform(0.0, 30.0) / xi)
                A = rng.uniform(-5, 5, 3)
                B = A - d * dist
                e = rng.normal(size=3)
                e /= np.linalg.norm(e)
                dK = np.zeros(3) if case == 1 else e * np.exp(rng.uniform(np.log(0.05), np.log(4.0)))
                r = np.zeros((16, 16), order="F")
                i = np.zeros((16, 16), order="F")
                Ab, Bb, Kb = (np.ascontiguousarray(x, np.float64) for x in (A, B, dK))
                lib.ref_primk(a1, a2, Ab.ctypes.data_as(dp), Bb.ctypes.data_as(dp),
                              Kb.ctypes.data_as(dp), sw, r.ctypes.data_as(dp), i.ctypes.data_as(dp))
                cls.append((la, lb, sw))
                inp.append((a1, a2, *A, *B, *dK))
                shr.append(r.copy(order="C"))
                shi.append(i.copy(order="C"))
    out = {"k_cls": np.array(cls, np.int32), "k_in": np.array(inp, np.float64),
           "k_re": np.array(shr, np.float64), "k_im": np.array(shi, np.float64)}
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
