"""The seeded workload generators (no GPU): the amorphous SiO2 cube has the
stated composition, density, minimum interatomic distance and basis size."""
import numpy as np

from olcao_amd import configs
from olcao_amd.system import BOHR_PER_ANGSTROM


def test_amorphous_sio2_cube():
    s = configs.asio2_1000()
    n_si = int(np.sum(np.asarray(s.atom_type) == 0))
    assert len(s.atom_pos) == 1000 and n_si == 333
    side = s.cell[0, 0] / BOHR_PER_ANGSTROM
    assert np.allclose(s.cell, np.eye(3) * s.cell[0, 0])
    mass_g = (333 * 28.0855 + 667 * 15.999) * 1.66053906660e-24
    assert abs(mass_g / (side * 1e-8) ** 3 - 2.2) < 1e-9
    pos = np.asarray(s.atom_pos) / BOHR_PER_ANGSTROM
    d = pos[:, None, :] - pos[None, :, :]
    d -= side * np.round(d / side)
    r = np.sqrt((d * d).sum(-1)) + np.eye(len(pos)) * 1e9
    assert r.min() >= 1.45
    assert s.gamma and len(s.kpoints) == 1 and np.all(np.asarray(s.kpoints) == 0.0)
    assert s.vale_dim == 333 * 9 + 667 * 4 and s.core_dim == 667
    # the generator is deterministic in its seed
    t = configs.asio2_1000()
    assert np.array_equal(np.asarray(t.atom_pos), np.asarray(s.atom_pos))


def test_uneven_batch():
    cells = configs.uneven_batch()
    assert len(cells) == 64
    masses = configs.ATOMIC_MASS
    sizes, nks, kinds = set(), set(), set()
    for s in cells:
        s.validate()
        n = s.n_atoms
        assert 8 <= n <= 216 and n % 8 == 0
        assert 1 <= s.nk <= 64 and s.nk * n <= 512
        assert s.gamma == (s.nk == 1)
        if not s.gamma:
            assert abs(np.sum(s.kweights) - 2.0) < 1e-12
        elems = [s.types[t].element for t in s.atom_type]
        assert set(elems) <= {configs.ELEM_C, configs.ELEM_SI, configs.ELEM_O}
        vol = abs(np.linalg.det(s.cell)) / BOHR_PER_ANGSTROM ** 3 * 1e-24   # cm^3
        rho = sum(masses[e] for e in elems) * 1.66053906660e-24 / vol
        assert 2.2 - 1e-9 <= rho <= 3.5 + 1e-9
        per_atom = {t.num_vale_states for t in s.types}
        assert per_atom <= {5, 9}
        sizes.add(n)
        nks.add(s.nk)
        kinds.add(per_atom.pop())
    # the batch spans the stated ranges
    assert min(sizes) == 8 and max(sizes) == 216 and 1 in nks and 64 in nks
    assert kinds == {5, 9}
    t = configs.uneven_batch()
    assert all(np.array_equal(a.atom_pos, b.atom_pos) for a, b in zip(cells, t))
