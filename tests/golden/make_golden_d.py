"""Write golden vectors for the force-derivative primitives.

Evaluates the reference's own dkinetic2CIntg (src/uolcao/gaussIntegrals.f90
:65512), dnuclear3CIntgCB (:88076) and delectron3CIntgCB (:112289), the
derivatives with respect to the ket centre B that computeForceIntg
(forces.F90:60) sums, compiled from the reference sources by
`make -C oracle ref` into oracle/_ref/libolcao_ref_prims.so, on seeded
inputs, and stores inputs and the (16,16,3) blocks in prims_d_golden.npz.
Run in the build container (the reference is not present on GPU machines):

    make -C oracle ref && python tests/golden/make_golden_d.py

Sampling: for every (l1, l2) class in {s,p,d,f}^2 (the nuclear f-f class
left out: the reference's Boys array stops at F_6 and that block needs
F_7), exponents log-uniform in [0.05, 2e5], the third exponent in
[0.1, 2e3], separations inside the 1e-16 screening range and the third
centre within a few widths of the pair centre.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref", "libolcao_ref_prims.so")
OUT = os.path.join(HERE, "prims_d_golden.npz")
CASES_PER_CLASS = 8


def main():
    lib = C.CDLL(os.path.abspath(REF))
    dp = C.POINTER(C.c_double)
    lib.ref_primd.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, dp, dp, dp, C.c_int, dp]
    rng = np.random.default_rng(20261205)
    cls, inp, shs = [], [], []
    for op in (10, 11, 12):
        for la in range(4):
            for lb in range(4):
                if op == 11 and la == 3 and lb == 3:
                    continue
                sw = (1 << la) + 16 * (1 << lb)
                for case in range(CASES_PER_CLASS):
                    a1, a2 = np.exp(rng.uniform(np.log(0.05), np.log(2e5), 2))
                    a3 = np.exp(rng.uniform(np.log(0.1), np.log(2e3)))
                    xi = a1 * a2 / (a1 + a2)
                    d = rng.normal(size=3)
                    d /= np.linalg.norm(d)
                    dist = 0.0 if case == 0 else np.sqrt(rng.uniform(0.0, 30.0) / xi)
                    A = rng.uniform(-5, 5, 3)
                    B = A - d * dist
                    P = (a1 * A + a2 * B) / (a1 + a2)
                    z3 = a1 + a2 + a3
                    mu = (a1 + a2) * a3 / z3
                    e = rng.normal(size=3)
                    e /= np.linalg.norm(e)
                    Cc = P + e * np.sqrt(rng.uniform(0.0, 20.0) / mu)
                    sh = np.zeros((16, 16, 3), order="F")
                    Ab, Bb, Cb = (np.ascontiguousarray(x, np.float64) for x in (A, B, Cc))
                    lib.ref_primd(op, a1, a2, a3, Ab.ctypes.data_as(dp), Bb.ctypes.data_as(dp),
                                  Cb.ctypes.data_as(dp), sw, sh.ctypes.data_as(dp))
                    cls.append((op, la, lb, sw))
                    inp.append((a1, a2, a3, *A, *B, *Cc))
                    shs.append(np.transpose(sh, (2, 0, 1)).copy())
    out = {"d_cls": np.array(cls, np.int32), "d_in": np.array(inp, np.float64),
           "d_sh": np.array(shs, np.float64)}
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
