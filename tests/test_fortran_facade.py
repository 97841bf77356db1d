"""The Fortran host path: a Fortran program (tests/fortran/facade_driver.f90)
that uses the facade module O_GPUFacade at the reference driver's call sites
(OLCAO.F90:248-279 setupSCF, :482 mainSCF, :645-662 intgPSCF, :758 bandPSCF,
and solveZHEGV as secularEqn.F90:227 calls it), run as its own process and
compared with the same library driven through ctypes and with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from olcao_amd import configs
from tests.conftest import ROOT, block_rel_err

DRIVER = os.path.join(ROOT, "tests", "fortran", "facade_driver")
HARTREE_EV = 27.211386245988


def write_system(s, num_states, path):
    """The system as parseInput leaves it, one array per record, each in
    the Fortran array's column-major order (1-based type numbers)."""
    f = s.flat()
    pot_dim = s.pot_dim
    coef = np.asarray(s.pot_coeffs, np.float64).reshape(s.spin, pot_dim)

    def row(a):
        return " ".join(repr(float(x)) if np.issubdtype(np.asarray(a).dtype, np.floating)
                        else str(int(x)) for x in np.asarray(a).reshape(-1))
    lines = [
        f"{f['nAtoms']} {f['nTypes']} {f['maxAlphas']} {f['maxRadial']} {f['nPotTypes']} "
        f"{f['maxPotAlphas']} {f['nPotSites']}",
        f"{f['nk']} {f['gamma']} {f['rel']} {f['spin']} {pot_dim} {num_states} "
        f"{f['basisCutoff']!r}",
        row(f["atomPos"]), row(f["atomType"] + 1), row(f["typeAlphas"]),
        row(f["typeNumAlphas"]), row(f["typeNumOrbAlphas"]), row(f["typeElement"]),
        row(f["typeNumValeRadial"]), row(f["typeNumCoreRadial"]), row(f["typeValeL"]),
        row(f["typeCoreL"]), row(f["typeValeCoef"]), row(f["typeCoreCoef"]),
        row(f["potAlphas"]), row(f["potNumAlphas"]), row(f["potZ"]), row(f["potNucAlpha"]),
        row(f["potElement"]), row(f["potPos"]), row(f["potSiteType"] + 1),
        row(f["cellVectors"]), row(f["kpoints"]), row(coef),
    ]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def read_output(path, num_states):
    with open(path) as fh:
        n, nk, spin = (int(x) for x in fh.readline().split())
        v = np.array([float(x.replace("D", "E")) for x in fh.read().split()])
    m = num_states * nk * spin
    evA = v[:m].reshape(spin, nk, num_states)
    evB = v[m:2 * m].reshape(spin, nk, num_states)
    evZ = v[2 * m:2 * m + num_states]
    rest = v[2 * m + num_states:]
    S = (rest[0:2 * n * n:2] + 1j * rest[1:2 * n * n:2]).reshape(n, n).T
    H = (rest[2 * n * n:4 * n * n:2] + 1j * rest[2 * n * n + 1:4 * n * n:2]).reshape(n, n).T
    return evA, evB, evZ, S, H


def test_driver_input_round_trip(tmp_path):
    """The input file holds each array in the order the Fortran reads it."""
    s = configs.small_test(kmesh=(2, 1, 1), two_types=True, core=True)
    p = tmp_path / "in.txt"
    write_system(s, 3, p)
    rows = p.read_text().splitlines()
    assert len(rows) == 24
    f = s.flat()
    np.testing.assert_array_equal(np.array(rows[2].split(), float), f["atomPos"].reshape(-1))
    np.testing.assert_array_equal(np.array(rows[12].split(), float),
                                  f["typeValeCoef"].reshape(-1))


@pytest.mark.gpu
def test_fortran_driver_through_facade(tmp_path):
    """A complex system with core states and two types: the Fortran
    program's Mode A and Mode B eigenvalues, the first k-point's S and H and
    the drop-in solveZHEGV eigenvalues equal the ctypes path's (1e-12) and
    the oracle's (integrals 3e-10 of the matrix scale after the core
    orthogonalisation's 1e-10 zeroing, eigenvalues 1e-6 eV)."""
    from olcao_amd import gpu
    if not os.path.exists(DRIVER):
        pytest.fail(f"Fortran driver not built: {DRIVER}")
    s = configs.small_test(kmesh=(2, 1, 1), two_types=True, core=True)
    g = gpu.OlcaoGPU(s)
    ns = g.vale_dim
    write_system(s, ns, tmp_path / "in.txt")
    print("fortran facade driver: running", flush=True)
    r = subprocess.run([DRIVER, str(tmp_path / "in.txt"), str(tmp_path / "out.txt")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    evA, evB, evZ, S, H = read_output(tmp_path / "out.txt", ns)
    # the same calls through ctypes
    Sg = g.gaussOverlapOL()[0]
    Hg = g.gaussOverlapHamPSCF(s.pot_coeffs)[0]
    wb = np.asarray(g.secularEqnPSCF(ns)).reshape(1, s.nk, ns)
    g.setupSCF()
    wa = np.asarray(g.secularEqnSCF(s.pot_coeffs, ns)).reshape(1, s.nk, ns)
    g.close()
    assert block_rel_err(S, Sg[0]) < 1e-12 and block_rel_err(H, Hg[0]) < 1e-12
    assert np.abs(evA - wa).max() * HARTREE_EV < 1e-9
    assert np.abs(evB - wb).max() * HARTREE_EV < 1e-9
    # against the oracle
    o = oracle.Oracle(s)
    So = o.integral(oracle.OL)
    Ho = o.integral(oracle.HAM, coef=s.pot_coeffs)
    S_or = oracle.ortho_dataset(*So, So[1], overlap=True)
    H_or = oracle.ortho_dataset(*Ho, So[1])
    # orthogonalised matrices: one side may zero a component just below
    # 1e-10 that the other keeps (intgOrtho.F90:142-190)
    assert block_rel_err(S, S_or[0]) < 3e-10 and block_rel_err(H, H_or[0]) < 3e-10
    wo, _ = oracle.solve(H_or, S_or)
    assert np.abs(evB[0] - wo[:, :ns]).max() * HARTREE_EV < 1e-6
    assert np.abs(evZ - wo[0, :ns]).max() * HARTREE_EV < 1e-6
