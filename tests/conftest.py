import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def block_rel_err(a, b):
    """max |a - b| / max |b|: relative error against the matrix (or block)
    scale, the measure used for every integral parity check."""
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.abs(b).max()
    if scale == 0.0:
        return float(np.abs(a).max())
    return float(np.abs(a - b).max() / scale)


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", "prims_golden.npz"))


def mode_a_terms(o, s):
    """The oracle's setup-time term matrices as the solve and the density
    traces see them (orthogonalised against the core when there is one):
    dict S, NP, KE, MV (None when rel = 0), V (list over potDim)."""
    import oracle
    So = o.integral(oracle.OL)
    S = oracle.ortho_dataset(*So, So[1], overlap=True)

    def op(kind, term=0):
        m = o.integral(kind, term=term)
        return oracle.ortho_dataset(*m, So[1])
    return dict(S=S, NP=op(oracle.NP), KE=op(oracle.KE),
                MV=op(oracle.MV) if s.rel else None,
                V=[op(oracle.EP, j) for j in range(s.pot_dim)])


def mode_a_hamiltonian(t, coef):
    import oracle
    return oracle.assemble_packed(t["NP"], t["KE"], t["V"], coef, t["MV"])


def fermi_occupancy(evals, kweights, n_el, spin, kT=0.01):
    """Populations [nk][spin][numStates] from Fermi smearing of the
    eigenvalues [spin][nk][numStates] (equal populations on degenerate
    levels, so the density does not depend on the eigenvector basis)."""
    import numpy as np
    w = np.asarray(evals, np.float64).reshape(spin, len(kweights), -1)
    flat = np.sort(w.reshape(-1))
    n_occ = min(int(round(n_el / 2)) * len(kweights) * spin, flat.size - 1)   # levels below mu
    mu = 0.5 * (flat[n_occ - 1] + flat[n_occ])
    f = np.asarray(kweights)[None, :, None] / spin / (1.0 + np.exp((w - mu) / kT))
    return np.ascontiguousarray(f.transpose(1, 0, 2))
