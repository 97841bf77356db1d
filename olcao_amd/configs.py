"""Seeded synthetic systems for the integral -> eigenvalue path.

The five workloads:
  diamond_c8   diamond-cubic C, 8-atom cell, 1s 2s 2p from 20 even-tempered
               exponents 0.12..1e6, 16 potential Gaussians 0.1..1e5,
               4x4x4 Monkhorst-Pack k-mesh (64 k)                    n = 40
  si64         2x2x2 diamond Si supercell, 0.05 A Gaussian jitter, s+p+d per
               Si (1s 2p 3d) from 24 exponents 0.05..1e7, 2x2x2 k   n = 576
  quartz       alpha-quartz SiO2 (P3_1 21, 9 atoms), Si s+p+d, O 1s 2s 2p,
               8x8x8 k-mesh (512 k)                                  n = 57
  asio2_1000   1000-atom amorphous SiO2 cube at 2.2 g/cm^3, Si 1s 2s 2p 3s 3p,
               O 2s 2p + 1s core, Gamma only (real)            n = 5665
  uneven_batch 64 independent cells of 8..216 C/Si/O atoms (diamond or
               zincblende, 0.1 A jitter, 2.2..3.5 g/cm^3), minimal or
               extended basis, 1..64 k-points each              n = 40..1296

plus small variants used by the tests.  Exponents are even-tempered,
alpha_i = amin (amax/amin)^(i/(N-1)), the rule of makePotAlphas
(src/uolcao/potTypes.f90:170-194).  Radial coefficients are drawn U(-1,1)
from the seed, then Loewdin-orthonormalised within each (type, l) channel
using the exact same-centre primitive overlap, so every orbital has unit
norm.  Geometry sources: skl/diamond and skl/silica (P3_1 21, Si 3a at
(0.465, 0, 1/3), O 6c at (0.417, 0.278, 0.222)).
"""
from __future__ import annotations

import math

import numpy as np

from .system import BOHR_PER_ANGSTROM, AtomType, PotType, SystemInput

ELEM_C, ELEM_O, ELEM_SI = 6, 8, 14
ALPHA_NUC = 20.0     # NUCLEAR_CHARGE__ALPHA of atomicPDB/{c,si,o}/pot1


def even_tempered(amin: float, amax: float, n: int) -> np.ndarray:
    i = np.arange(n, dtype=np.float64)
    return amin * (amax / amin) ** (i / (n - 1))


def radial_gram(alphas: np.ndarray, l: int) -> np.ndarray:
    """<g_a|g_b> of angNorm-normalised primitives r^l Y_lm exp(-a r^2)."""
    s = alphas[:, None] + alphas[None, :]
    return math.gamma(l + 1.5) / (2.0 * s ** (l + 1.5))


def radial_functions(rng, alphas, l, count):
    """`count` radial functions of angular momentum l over `alphas`,
    Loewdin-orthonormal with respect to the primitive overlap."""
    c = rng.uniform(-1.0, 1.0, size=(len(alphas), count))
    s = c.T @ radial_gram(alphas, l) @ c
    w, v = np.linalg.eigh(s)
    c = c @ (v @ np.diag(w ** -0.5) @ v.T)
    return [np.ascontiguousarray(c[:, i]) for i in range(count)]


def make_type(rng, element, alphas, num_orb_alphas, vale_l, core_l=()):
    """Atom type with radial functions for the listed l values (same-l
    functions are orthonormalised together)."""
    by_l = {}
    for l in list(vale_l) + list(core_l):
        by_l[l] = by_l.get(l, 0) + 1
    fns = {l: radial_functions(rng, alphas[:num_orb_alphas[l]], l, n)
           for l, n in sorted(by_l.items())}
    used = {l: 0 for l in by_l}

    def take(l):
        f = fns[l][used[l]]
        used[l] += 1
        return f
    core = [take(l) for l in core_l]
    vale = [take(l) for l in vale_l]
    return AtomType(element=element, alphas=np.asarray(alphas, float),
                    num_orb_alphas=tuple(num_orb_alphas), vale_l=list(vale_l),
                    vale_coef=vale, core_l=list(core_l), core_coef=core)


def mp_mesh(n, shift=(0.5, 0.5, 0.5)):
    """Monkhorst-Pack mesh in fractional reciprocal coordinates,
    f = -1/2 + (i - 1 + shift)/n, i outermost (kpoints.f90:749-766)."""
    pts = []
    for i in range(1, n[0] + 1):
        for j in range(1, n[1] + 1):
            for k in range(1, n[2] + 1):
                idx = np.array([i, j, k], float)
                pts.append(-0.5 + (idx - 1.0 + np.asarray(shift)) / np.asarray(n, float))
    pts = np.array(pts)
    return pts, np.full(len(pts), 2.0 / len(pts))


def cubic(a_bohr):
    return np.eye(3) * a_bohr


def hexagonal(a, c):
    return np.array([[a, 0, 0], [-0.5 * a, 0.5 * math.sqrt(3.0) * a, 0], [0, 0, c]])


DIAMOND_FRAC = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]], float)
DIAMOND_FRAC = np.vstack([DIAMOND_FRAC, DIAMOND_FRAC + 0.25])


def _pot_sites_from_atoms(pos, atom_type, type_to_pot):
    pt = np.array([type_to_pot[t] for t in atom_type], np.int32)
    order = np.argsort(pt, kind="stable")
    return pos[order].copy(), pt[order]


def _assemble(name, cell, pos, atom_type, types, pot_types, type_to_pot, kpts,
              kw, rng, gamma=False, cutoff=1e-16, rel=0, spin=1):
    ppos, ptype = _pot_sites_from_atoms(pos, atom_type, type_to_pot)
    pot_dim = sum(len(p.alphas) for p in pot_types)
    coeffs = rng.uniform(-1.0, 1.0, size=pot_dim)
    if spin == 2:   # a second, spin-split coefficient set
        coeffs = np.stack([coeffs, coeffs + rng.uniform(-0.1, 0.1, size=pot_dim)])
    return SystemInput(name=name, cell=cell, atom_pos=pos,
                       atom_type=np.asarray(atom_type, np.int32), types=types,
                       pot_types=pot_types, pot_pos=ppos, pot_type=ptype,
                       kpoints=kpts, kweights=kw, pot_coeffs=coeffs,
                       basis_cutoff=cutoff, gamma=gamma, rel=rel, spin=spin)


def diamond_c8(seed=1, kmesh=(4, 4, 4), n_alpha=20, n_pot=16):
    """C1: diamond C8, a = 3.567 A, 1s 2s 2p, 64 k."""
    rng = np.random.default_rng(seed)
    a = 3.567 * BOHR_PER_ANGSTROM
    cell = cubic(a)
    pos = DIAMOND_FRAC @ cell
    al = even_tempered(0.12, 1.0e6, n_alpha)
    t = make_type(rng, ELEM_C, al, (n_alpha, n_alpha - 2, 0, 0), [0, 0, 1])
    p = PotType(ELEM_C, 6.0, ALPHA_NUC, even_tempered(0.1, 1.0e5, n_pot))
    k, w = mp_mesh(kmesh)
    return _assemble("diamond_c8", cell, pos, [0] * 8, [t], [p], {0: 0}, k, w, rng)


def si64(seed=2, kmesh=(2, 2, 2), n_alpha=24, n_pot=16, reps=2, jitter=0.05):
    """C2: Si64 supercell, a = 5.431 A, s+p+d per Si, 8 k."""
    rng = np.random.default_rng(seed)
    a = 5.431 * BOHR_PER_ANGSTROM
    frac = []
    for i in range(reps):
        for j in range(reps):
            for k in range(reps):
                frac.append(DIAMOND_FRAC + np.array([i, j, k], float))
    frac = np.vstack(frac) / reps
    cell = cubic(a * reps)
    pos = frac @ cell + rng.normal(0.0, jitter * BOHR_PER_ANGSTROM, size=frac.shape)
    al = even_tempered(0.05, 1.0e7, n_alpha)
    t = make_type(rng, ELEM_SI, al, (n_alpha, n_alpha - 2, n_alpha - 4, 0), [0, 1, 2])
    p = PotType(ELEM_SI, 14.0, ALPHA_NUC, even_tempered(0.1, 1.0e5, n_pot))
    k, w = mp_mesh(kmesh)
    return _assemble("si64", cell, pos, [0] * len(pos), [t], [p], {0: 0}, k, w, rng)


def _sg152(xyz):
    """P3_1 21 (No. 152) general-position orbit, de-duplicated, in [0,1)."""
    x, y, z = xyz
    ops = [(x, y, z), (-y, x - y, z + 1 / 3), (-x + y, -x, z + 2 / 3),
           (y, x, -z), (x - y, -y, -z + 2 / 3), (-x, -x + y, -z + 1 / 3)]
    out = []
    for p in ops:
        p = np.mod(np.array(p, float), 1.0)
        if not any(np.allclose(np.mod(p - q + 0.5, 1.0) - 0.5, 0, atol=1e-6) for q in out):
            out.append(p)
    return out


def quartz(seed=3, kmesh=(8, 8, 8), n_alpha_si=24, n_alpha_o=20, n_pot=16):
    """C3: alpha-quartz SiO2, 9 atoms, Si s+p+d, O 1s 2s 2p, 512 k."""
    rng = np.random.default_rng(seed)
    cell = hexagonal(4.913 * BOHR_PER_ANGSTROM, 5.405 * BOHR_PER_ANGSTROM)
    si = _sg152((0.465, 0.0, 1.0 / 3.0))
    ox = _sg152((0.417, 0.278, 0.222))
    frac = np.array(si + ox)
    pos = frac @ cell
    tsi = make_type(rng, ELEM_SI, even_tempered(0.05, 1.0e7, n_alpha_si),
                    (n_alpha_si, n_alpha_si - 2, n_alpha_si - 4, 0), [0, 1, 2])
    to = make_type(rng, ELEM_O, even_tempered(0.15, 1.0e6, n_alpha_o),
                   (n_alpha_o, n_alpha_o - 2, 0, 0), [0, 0, 1])
    psi = PotType(ELEM_SI, 14.0, ALPHA_NUC, even_tempered(0.1, 1.0e5, n_pot))
    po = PotType(ELEM_O, 8.0, ALPHA_NUC, even_tempered(0.12, 1.0e5, n_pot))
    types = [0] * len(si) + [1] * len(ox)
    k, w = mp_mesh(kmesh)
    return _assemble("quartz", cell, pos, types, [tsi, to], [psi, po], {0: 0, 1: 1},
                     k, w, rng)


def small_test(seed=7, kmesh=(2, 2, 1), n_alpha=8, lmax=1, hexa=False,
               two_types=False, core=False, gamma=False, a_ang=3.2, n_pot=5, rel=0,
               spin=1):
    """Few-second oracle workloads for the parity tests."""
    rng = np.random.default_rng(seed)
    if hexa:
        cell = hexagonal(a_ang * BOHR_PER_ANGSTROM, 1.3 * a_ang * BOHR_PER_ANGSTROM)
    else:
        cell = cubic(a_ang * BOHR_PER_ANGSTROM)
    frac = np.array([[0.0, 0.0, 0.0], [0.27, 0.48, 0.61]])
    pos = frac @ cell + rng.normal(0, 0.03, size=frac.shape)
    al = even_tempered(0.18, 5.0e3, n_alpha)
    norb = [n_alpha, max(n_alpha - 1, 1), max(n_alpha - 2, 1), max(n_alpha - 3, 1)]
    norb = tuple(norb[l] if l <= lmax else 0 for l in range(4))
    vale = [0] + list(range(1, lmax + 1))
    core_l = [0] if core else []
    t0 = make_type(rng, ELEM_C, al, norb, vale, core_l)
    types = [t0]
    atom_types = [0, 0]
    pots = [PotType(ELEM_C, 6.0, ALPHA_NUC, even_tempered(0.15, 2.0e3, n_pot))]
    t2p = {0: 0}
    if two_types:
        al2 = even_tempered(0.22, 8.0e3, n_alpha - 1)
        norb2 = (n_alpha - 1, n_alpha - 2, 0, 0)
        types.append(make_type(rng, ELEM_O, al2, norb2, [0, 1], core_l))
        atom_types = [0, 1]
        pots.append(PotType(ELEM_O, 8.0, ALPHA_NUC, even_tempered(0.2, 3.0e3, n_pot - 1)))
        t2p = {0: 0, 1: 1}
    k, w = mp_mesh(kmesh)
    if gamma:
        k, w = np.zeros((1, 3)), np.array([2.0])
    return _assemble("small_test", cell, pos, atom_types, types, pots, t2p, k, w, rng,
                     gamma=gamma, rel=rel, spin=spin)


def asio2(seed=5, n_si=333, n_o=667, density=2.2, dmin=1.45, n_alpha_si=12, n_alpha_o=10,
          n_pot=8):
    """C5: amorphous SiO2 cube at `density` g/cm^3, Gamma only (real H and
    S).  Positions uniform from the seed with a `dmin` A minimum distance
    (minimum image) enforced by rejection.  Minimal basis: Si 1s 2s 2p 3s
    3p (9 orbitals) from `n_alpha_si` exponents 0.2..1e6; O 2s 2p valence
    (4 orbitals) with a 1s core state from `n_alpha_o` exponents 0.25..1e6.
    333 Si + 667 O: 5665 valence and 667 core orbitals."""
    rng = np.random.default_rng(seed)
    mass = n_si * 28.0855 + n_o * 15.999
    side = (mass * 1.66053906660e-24 / density * 1e24) ** (1.0 / 3.0)   # A
    pos = np.empty((0, 3))
    while len(pos) < n_si + n_o:
        p = rng.uniform(0.0, side, size=3)
        d = pos - p
        d -= side * np.round(d / side)
        if len(pos) == 0 or (d * d).sum(axis=1).min() >= dmin * dmin:
            pos = np.vstack([pos, p])
    cell = cubic(side * BOHR_PER_ANGSTROM)
    pos = pos * BOHR_PER_ANGSTROM
    tsi = make_type(rng, ELEM_SI, even_tempered(0.2, 1.0e6, n_alpha_si),
                    (n_alpha_si, n_alpha_si - 2, 0, 0), [0, 0, 1, 0, 1])
    to = make_type(rng, ELEM_O, even_tempered(0.25, 1.0e6, n_alpha_o),
                   (n_alpha_o, n_alpha_o - 2, 0, 0), [0, 1], core_l=[0])
    psi = PotType(ELEM_SI, 14.0, ALPHA_NUC, even_tempered(0.2, 1.0e4, n_pot))
    po = PotType(ELEM_O, 8.0, ALPHA_NUC, even_tempered(0.25, 1.0e4, n_pot))
    types = [0] * n_si + [1] * n_o
    k, w = np.zeros((1, 3)), np.array([2.0])
    return _assemble("asio2", cell, pos, types, [tsi, to], [psi, po], {0: 0, 1: 1}, k, w, rng,
                     gamma=True)


def asio2_1000(seed=5):
    """C5 at its full size: 333 Si + 667 O."""
    return asio2(seed=seed)


# conventional diamond cells per supercell, 8..216 atoms
BATCH_SUPERCELLS = [(1, 1, 1), (2, 1, 1), (3, 1, 1), (2, 2, 1), (3, 2, 1), (2, 2, 2),
                    (3, 3, 1), (3, 2, 2), (3, 3, 2), (3, 3, 3)]
BATCH_MESHES = [(1, 1, 1), (2, 1, 1), (2, 2, 1), (2, 2, 2), (3, 2, 2), (3, 3, 2), (3, 3, 3),
                (4, 3, 3), (4, 4, 3), (4, 4, 4)]
ATOMIC_MASS = {ELEM_C: 12.011, ELEM_O: 15.999, ELEM_SI: 28.0855}
ELEM_BASIS = {   # exponent range (bohr^-2), extended-basis exponent count, potential range
    ELEM_C: (0.12, 1.0e6, 20, 0.1),
    ELEM_O: (0.15, 1.0e6, 20, 0.12),
    ELEM_SI: (0.05, 1.0e7, 24, 0.1),
}


def batch_cell(rng, index, k_budget=512):
    """One cell of the uneven batch.  A diamond (one element) or zincblende
    (two elements on the two fcc sublattices) supercell of 8..216 atoms of
    C, Si or O, lattice constant set by a density drawn from 2.2..3.5
    g/cm^3, 0.1 A Gaussian jitter.  Basis chosen at random: minimal (1s 2s
    2p per atom, 4 fewer exponents) or extended (s + p + d, as the Si64
    workload).  k-points: a Monkhorst-Pack mesh of 1..64 points drawn from
    the meshes with nk * atoms <= k_budget (a k density that falls with the
    cell size); one point is the real Gamma build."""
    reps = BATCH_SUPERCELLS[rng.integers(len(BATCH_SUPERCELLS))]
    elems = [ELEM_C, ELEM_SI, ELEM_O]
    e1, e2 = elems[rng.integers(3)], elems[rng.integers(3)]
    density = rng.uniform(2.2, 3.5)
    m_avg = 0.5 * (ATOMIC_MASS[e1] + ATOMIC_MASS[e2])
    a = (8 * m_avg * 1.66053906660e-24 / density * 1e24) ** (1.0 / 3.0)   # A
    frac, sub = [], []
    for i in range(reps[0]):
        for j in range(reps[1]):
            for k in range(reps[2]):
                frac.append(DIAMOND_FRAC + np.array([i, j, k], float))
                sub.extend([0] * 4 + [1] * 4)
    frac = np.vstack(frac) / np.asarray(reps, float)
    cell = np.diag(np.asarray(reps, float) * a * BOHR_PER_ANGSTROM)
    pos = frac @ cell + rng.normal(0.0, 0.1 * BOHR_PER_ANGSTROM, size=frac.shape)
    extended = bool(rng.integers(2))
    uniq = sorted({e1, e2}, key=[e1, e2].index)
    types, pots = [], []
    for e in uniq:
        amin, amax, na, pmin = ELEM_BASIS[e]
        if extended:
            t = make_type(rng, e, even_tempered(amin, amax, na), (na, na - 2, na - 4, 0), [0, 1, 2])
        else:
            na -= 4
            t = make_type(rng, e, even_tempered(amin, amax, na), (na, na - 2, 0, 0), [0, 0, 1])
        types.append(t)
        pots.append(PotType(e, float(e), ALPHA_NUC, even_tempered(pmin, 1.0e5, 16)))
    atom_types = [uniq.index(e1 if s == 0 else e2) for s in sub]
    n_atoms = len(pos)
    meshes = [m for m in BATCH_MESHES if m[0] * m[1] * m[2] * n_atoms <= k_budget]
    mesh = meshes[rng.integers(len(meshes))]
    gamma = mesh == (1, 1, 1)
    if gamma:
        k, w = np.zeros((1, 3)), np.array([2.0])
    else:
        k, w = mp_mesh(mesh)
    name = f"batch{index:02d}"
    return _assemble(name, cell, pos, atom_types, types, pots,
                     {i: i for i in range(len(uniq))}, k, w, rng, gamma=gamma)


def uneven_batch(seed=4, n_cells=64):
    """C4: `n_cells` independent cells drawn from one seed (batch_cell)."""
    rng = np.random.default_rng(seed)
    return [batch_cell(rng, i) for i in range(n_cells)]


CONFIGS = {"diamond_c8": diamond_c8, "si64": si64, "quartz": quartz, "asio2_1000": asio2_1000}
