"""Write golden vectors for the three-component primitives: the dipole
integrals <a| (r - C) |b> (dipole3CIntg, src/uolcao/gaussIntegrals.f90:49404)
and the momentum integrals <a| d/dr |b> (momentum2CIntg, :23041), evaluated
by the reference's own subroutines compiled from its sources by
`make -C oracle ref` into oracle/_ref/libolcao_ref_prims.so, on seeded
inputs; stored in prims_vec_golden.npz.  Run in the build container:

    make -C oracle ref && python tests/golden/make_golden_vec.py

Sampling as make_golden.py: every (l1, l2) class in {s,p,d,f}^2, exponents
log-uniform, separations inside the 1e-16 screening range, dipole centres
within a few bohr of the pair.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref", "libolcao_ref_prims.so")
OUT = os.path.join(HERE, "prims_vec_golden.npz")

OPS = {"dipole": 5, "momentum": 6}
CASES_PER_CLASS = 8


def main():
    lib = C.CDLL(os.path.abspath(REF))
    dp = C.POINTER(C.c_double)
    lib.ref_prim3.argtypes = [C.c_int, C.c_double, C.c_double, dp, dp, dp, C.c_int, dp]
    rng = np.random.default_rng(20261204)
    out = {}
    for name, op in OPS.items():
        rows = []
        for la in range(4):
            for lb in range(4):
                sw = (1 << la) + 16 * (1 << lb)
                for case in range(CASES_PER_CLASS):
                    a1, a2 = np.exp(rng.uniform(np.log(0.05), np.log(2e5), 2))
                    xi = a1 * a2 / (a1 + a2)
                    d = rng.normal(size=3)
                    d /= np.linalg.norm(d)
                    dist = 0.0 if case == 0 else np.sqrt(rng.uniform(0.0, 30.0) / xi)
                    A = rng.uniform(-5, 5, 3)
                    B = A - d * dist
                    Cc = rng.uniform(-3, 3, 3)
                    sh = np.zeros((3, 16, 16))            # Fortran (16,16,3) memory order
                    Ab, Bb, Cb = (np.ascontiguousarray(x, np.float64) for x in (A, B, Cc))
                    lib.ref_prim3(op, a1, a2, Ab.ctypes.data_as(dp), Bb.ctypes.data_as(dp),
                                  Cb.ctypes.data_as(dp), sw, sh.ctypes.data_as(dp))
                    # sh[c] holds component c column-major: transpose to [c, row, col]
                    rows.append((la, lb, sw, a1, a2, *A, *B, *Cc, sh.transpose(0, 2, 1).copy()))
        out[name + "_cls"] = np.array([r[:3] for r in rows], np.int32)
        out[name + "_in"] = np.array([r[3:14] for r in rows], np.float64)
        out[name + "_sh"] = np.array([r[14] for r in rows], np.float64)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
