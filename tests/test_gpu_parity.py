"""Parity of the HIP path (through the C-ABI) with the oracle and with the
reference's golden vectors.

Tolerances (stated per the path's contract): integrals <= 1e-10 relative
to the matrix scale (block_rel_err); eigenvalues <= 1e-6 eV; Boys values
equal the reference's to the cancellation noise of its closed form
(F_0..F_4 <= 1e-10, F_5/F_6 <= 2e-6 relative just above T = 0.1).
"""
import os
import sys

import numpy as np
import pytest

import oracle
from olcao_amd import configs
from tests.conftest import block_rel_err

pytestmark = pytest.mark.gpu
HARTREE_EV = 27.211386245988
NSH = [1, 4, 9, 16]
OPS = {"overlap": 0, "kinetic": 1, "electron3c": 2, "nuclear3c": 3, "massvel": 4}


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    from olcao_amd import gpu as g
    return g


@pytest.mark.parametrize("name", list(OPS))
def test_device_primitives_match_reference_golden(gpu, golden, name):
    _log(f"prims {name}")
    cls, inp, ref = golden[name + "_cls"], golden[name + "_in"], golden[name + "_sh"]
    out = gpu.prim_batch(OPS[name], inp[:, :3], inp[:, 3:6], inp[:, 6:9], inp[:, 9:12], cls[:, 2])
    worst = 0.0
    for i in range(len(cls)):
        na, nb = NSH[cls[i, 0]], NSH[cls[i, 1]]
        e = block_rel_err(out[i][:na, :nb], ref[i][:na, :nb])
        worst = max(worst, e)
        tol = 1e-10 if (name != "nuclear3c" or cls[i, 0] + cls[i, 1] <= 4) else 5e-6
        if name == "massvel":   # the reference's own cancellation (see test_oracle)
            tol = 3e-13 + 1e-15 * (inp[i, :2].max() / inp[i, :2].min()) ** 2
        assert e < tol, (name, i, cls[i], e)
    _log(f"  worst {worst:.2e}")


def test_device_boys_matches_reference(gpu, golden):
    T, ref = golden["boys_T"], golden["boys_F"]
    F = gpu.boys_batch(T)
    rel = np.abs(F - ref) / np.abs(ref)
    assert rel[:, :5].max() < 1e-10
    assert rel.max() < 2e-6


def test_device_primitives_random_vs_oracle(gpu):
    rng = np.random.default_rng(9)
    n = 4000
    a = np.exp(rng.uniform(np.log(0.05), np.log(1e5), (n, 3)))
    A = rng.uniform(-3, 3, (n, 3))
    B = A + rng.normal(0, 0.6, (n, 3))
    C = A + rng.normal(0, 0.8, (n, 3))
    la = rng.integers(0, 3, n)
    lb = rng.integers(0, 3, n)
    sw = (1 << la) + 16 * (1 << lb)
    for op in range(5):
        g = gpu.prim_batch(op, a, A, B, C, sw)
        o = oracle.prim_batch(op, a, A, B, C, sw)
        errs = [block_rel_err(g[i][:NSH[la[i]], :NSH[lb[i]]], o[i][:NSH[la[i]], :NSH[lb[i]]])
                for i in range(n) if np.abs(o[i]).max() > 1e-250]
        assert max(errs) < 1e-10, (op, max(errs))


def _pair(gpu, s):
    return gpu.OlcaoGPU(s), oracle.Oracle(s)


SMALL = {
    "cubic_sp": dict(),
    "hex_spd": dict(lmax=2, hexa=True),
    "two_types": dict(two_types=True),
    "spdf": dict(lmax=3, n_alpha=6),
    "core": dict(core=True),
    "gamma_spd": dict(gamma=True, lmax=2),
    "rel_spin2": dict(lmax=2, rel=1, spin=2),
    "core_gamma_rel": dict(core=True, gamma=True, rel=1, two_types=True),
}


def _ortho_close(a, b, tol):
    """block_rel_err allowing the 1e-10 zeroing of the orthogonalised
    packing to fall differently on components within rounding of 1e-10."""
    a, b = np.asarray(a), np.asarray(b)
    d = np.abs(a - b)
    edge = (np.minimum(np.abs(a), np.abs(b)) == 0) & (np.maximum(np.abs(a), np.abs(b)) < 2e-10)
    d = np.where(edge, 0.0, d)
    return float(d.max() / max(np.abs(b).max(), 1e-300)) < tol


@pytest.mark.parametrize("case", list(SMALL))
def test_setup_matches_oracle(gpu, case):
    s = configs.small_test(**SMALL[case])
    g, o = _pair(gpu, s)
    cells, sizes, k = g.setup()
    assert len(sizes) == len(o.cellSizes)
    assert np.allclose(sizes, o.cellSizes, rtol=1e-13)
    assert np.allclose(np.sort(np.round(cells, 9), axis=0), np.sort(np.round(o.cellDims, 9), axis=0))
    if not s.gamma:
        assert np.allclose(k, o.kcart, rtol=1e-14, atol=1e-15)
    g.close()


@pytest.mark.parametrize("case", list(SMALL))
def test_integral_sweeps_match_oracle(gpu, case):
    """Every sweep against the oracle: raw coreVale/coreCore, and valeVale as
    the reference writes it (orthogonalised against the core when there is
    one); per spin for the fused Hamiltonian."""
    _log(f"sweeps {case}")
    s = configs.small_test(**SMALL[case])
    g, o = _pair(gpu, s)
    tol = 1e-10 if case != "spdf" else 1e-8
    So = o.integral(oracle.OL)
    coefs = np.asarray(s.pot_coeffs).reshape(s.spin, -1)
    checks = [("OL", g.gaussOverlapOL, lambda: So, True),
              ("KE", g.gaussOverlapKE, lambda: o.integral(oracle.KE), False),
              ("NP", g.gaussOverlapNP, lambda: o.integral(oracle.NP), False)]
    if s.rel:
        checks.append(("MV", g.gaussOverlapMV, lambda: o.integral(oracle.MV), False))
    for j in range(s.pot_dim):
        checks.append((f"EP{j}", lambda j=j: g.elecPotGaussOverlap(j),
                       lambda j=j: o.integral(oracle.EP, term=j), False))
    hams = g.gaussOverlapHamPSCF(s.pot_coeffs)
    hams = [hams] if s.spin == 1 else list(hams)
    for q in range(s.spin):
        checks.append((f"HAM{q}", lambda q=q: hams[q],
                       lambda q=q: o.integral(oracle.HAM, coef=coefs[q]), False))
    for name, fg, fo, isol in checks:
        mg, mo = fg(), fo()
        if s.core_dim:
            ref = oracle.ortho_dataset(mo[0], mo[1], mo[2], So[1], overlap=isol)
            assert _ortho_close(mg[0], ref, tol), (case, name)
            assert block_rel_err(mg[1], mo[1]) < tol, (case, name, "coreVale")
            assert block_rel_err(mg[2], mo[2]) < tol, (case, name, "coreCore")
        else:
            e = block_rel_err(mg[0], mo[0])
            assert e < tol, (case, name, e)
    g.close()


@pytest.mark.parametrize("case", ["two_types", "core", "core_gamma_rel", "hex_spd"])
def test_dipole_and_momentum_sweeps_match_oracle(gpu, case):
    """The dipole (about a fixed centre) and momentum matrices, every
    Cartesian component: valeVale orthogonalised against the core, coreVale
    and coreCore raw, against the oracle to 1e-10 of each matrix's scale;
    complex and Gamma builds, core states, s-p-d bases."""
    s = configs.small_test(**SMALL[case])
    g, o = _pair(gpu, s)
    center = np.array([0.7, -1.3, 2.1])
    dm, mm = g.gaussOverlapDM(center), g.gaussOverlapMM()
    S_cv = o.integral(oracle.OL)[1]
    for name, dev, op in (("DM", dm, oracle.DM), ("MM", mm, oracle.MM)):
        for comp in range(3):
            vv, cv, cc = o.integral(op, term=comp, coef=center)
            ref = oracle.ortho_dataset(vv, cv, cc, S_cv)
            e = block_rel_err(dev[comp][0], ref)
            _log(f"  {case} {name}[{comp}]: {e:.2e}")
            assert _ortho_close(dev[comp][0], ref, 1e-10), (case, name, comp, e)
            if cv is not None:
                assert block_rel_err(dev[comp][1], cv) < 1e-10
                assert block_rel_err(dev[comp][2], cc) < 1e-10
    g.close()


def test_device_k_overlap_primitive_matches_reference_golden(gpu):
    """KOverlap2CIntg (gaussIntegrals.f90:129408) on the device (real part op
    7, imaginary part op 8) against the reference's own blocks: 10 cases per
    class in {s,p,d,f}^2, |deltaK| from 0 to 4 inverse bohr, within 1e-10 of
    each block's scale (the device uses Hermite E coefficients, the reference
    a binomial expansion about P)."""
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "prims_k_golden.npz"))
    cls, inp = d["k_cls"], d["k_in"]
    ref = d["k_re"] + 1j * d["k_im"]
    a = np.column_stack([inp[:, :2], np.zeros(len(inp))])
    re = gpu.prim_batch(7, a, inp[:, 2:5], inp[:, 5:8], inp[:, 8:11], cls[:, 2])
    im = gpu.prim_batch(8, a, inp[:, 2:5], inp[:, 5:8], inp[:, 8:11], cls[:, 2])
    worst = 0.0
    for i in range(len(cls)):
        na, nb = NSH[cls[i, 0]], NSH[cls[i, 1]]
        e = block_rel_err((re[i] + 1j * im[i])[:na, :nb], ref[i][:na, :nb])
        worst = max(worst, e)
        assert e < 1e-10, (i, cls[i], e)
    _log(f"  k-overlap primitive worst {worst:.2e}")


@pytest.mark.parametrize("case", ["cubic_sp", "core", "hex_spd", "two_types"])
def test_k_overlap_sweep_matches_oracle(gpu, case):
    """gaussKOverlap (integrals3Terms.F90:956-1384) per axis: valeVale is the
    full orthogonalised matrix (ortho3Terms opcode 8, no zeroing), coreVale
    and coreCore raw, against the oracle to 1e-10 of each matrix's scale;
    numAxialKPoints differs per axis so each axis has its own deltaK."""
    s = configs.small_test(**SMALL[case])
    g, o = _pair(gpu, s)
    num_axial = (2, 3, 1)
    dev = g.gaussKOverlap(num_axial)
    ref = o.k_overlap(num_axial)
    S_cv = o.integral(oracle.OL)[1]
    for axis in range(3):
        vv, cv, cc = ref[axis]
        full = oracle.ortho_k_overlap(vv, cv, cc, S_cv)
        e = block_rel_err(dev[axis][0], full)
        _log(f"  {case} KO[{axis}]: {e:.2e}")
        assert e < 1e-10, (case, axis, e)
        if cv is not None:
            assert block_rel_err(dev[axis][1], cv) < 1e-10
            assert block_rel_err(dev[axis][2], cc) < 1e-10
    g.close()


def test_k_overlap_refused_for_gamma(gpu):
    """The k-overlap matrices are complex: a Gamma context refuses them with
    an input error rather than returning real parts."""
    g = gpu.OlcaoGPU(configs.small_test(**SMALL["gamma_spd"]))
    with pytest.raises(Exception):
        g.gaussKOverlap((1, 1, 1))
    g.close()


@pytest.mark.skip(reason="the first version of this test (derivative ops in the parity-primitive "
                  "kernel, whose f-f instance it pushed to a new scratch maximum) went silent on the "
                  "GPU with no cause shown on the CPU; those ops are withdrawn and the derivative "
                  "primitives are checked through the force sweep against the oracle")
def test_device_force_primitives_match_reference_golden(gpu):
    """The force integrals' primitives on the device: d/dB of the kinetic
    (ops 10-12), nuclear (13-15) and three-Gaussian (16-18) blocks per
    Cartesian component, against the reference's dkinetic2CIntg,
    dnuclear3CIntgCB and delectron3CIntgCB (gaussIntegrals.f90:65512,
    :88076, :112289) on 8 cases per class, within 1e-10 of each block's
    scale; the three-centre f-f class is refused on the device (it needs
    F_7) and left out."""
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "prims_d_golden.npz"))
    cls, inp, ref = d["d_cls"], d["d_in"], d["d_sh"]
    worst = 0.0
    for op, base in ((10, 10), (11, 13), (12, 16)):
        m = (cls[:, 0] == op) & ~((op != 10) & (cls[:, 1] == 3) & (cls[:, 2] == 3))
        for x in range(3):
            out = gpu.prim_batch(base + x, inp[m, :3], inp[m, 3:6], inp[m, 6:9], inp[m, 9:12],
                                 cls[m, 3])
            for o_, r, c in zip(out, ref[m][:, x], cls[m]):
                na, nb = NSH[c[1]], NSH[c[2]]
                e = block_rel_err(o_[:na, :nb], r[:na, :nb])
                worst = max(worst, e)
                assert e < 1e-10, (op, x, c, e)
    _log(f"  force primitives worst {worst:.2e}")


FORCE_CASES = ["two_types", "core", "gamma_spd", "rel_spin2", "core_gamma_rel"]


@pytest.mark.parametrize("case", FORCE_CASES)
def test_force_integrals_match_oracle(gpu, case):
    """computeForceIntg (forces.F90:60-898) on the device against the
    oracle, every spin and Cartesian component: valeVale saved in full in
    complex builds (the upper triangle at Gamma), coreVale and coreCore,
    within 1e-10 of each matrix's scale."""
    _log(f"force integrals {case}")
    s = configs.small_test(**SMALL[case])
    g, o = _pair(gpu, s)
    E = -7.25
    dev = g.computeForceIntg(s.pot_coeffs, E)
    ref = o.force_integrals(E)
    for q in range(s.spin):
        for x in range(3):
            for part in range(3):
                if ref[q][x][part] is None:
                    continue
                e = block_rel_err(dev[q][x][part], ref[q][x][part])
                _log(f"  {case} F[{q}][{x}][{part}]: {e:.2e}")
                assert e < 1e-10, (case, q, x, part, e)
    g.close()


def test_force_integrals_refuse_f_orbitals(gpu):
    """A basis with f orbitals is refused (the f-f nuclear derivative needs
    F_7)."""
    s = configs.small_test(**SMALL["spdf"])
    g = gpu.OlcaoGPU(s)
    with pytest.raises(gpu.OlcaoError):
        g.computeForceIntg(s.pot_coeffs, -1.0)
    g.close()


@pytest.mark.parametrize("case", ["core", "rel_spin2", "core_gamma_rel"])
def test_solve_core_rel_spin_matches_oracle(gpu, case):
    """Eigenvalues of every spin (Mode B) and of Mode A after the core
    orthogonalisation and with the mass-velocity term, against LAPACK on the
    oracle's matrices."""
    s = configs.small_test(**SMALL[case])
    g, o = _pair(gpu, s)
    coefs = np.asarray(s.pot_coeffs).reshape(s.spin, -1)
    So = o.integral(oracle.OL)
    S = oracle.ortho_dataset(*So, So[1], overlap=True)
    wo = []
    for q in range(s.spin):
        Ho = o.integral(oracle.HAM, coef=coefs[q])
        wo.append(oracle.solve(oracle.ortho_dataset(*Ho, So[1]), S)[0])
    wo = np.array(wo)
    n = g.vale_dim
    wb = np.asarray(g.scf_iteration(s.pot_coeffs, n)).reshape(s.spin, -1, n)
    err = np.abs(wb - wo.reshape(wb.shape)).max() * HARTREE_EV
    _log(f"  {case} Mode B eigenvalue error {err:.2e} eV")
    assert err < 1e-6
    g.setupSCF()
    wa = np.asarray(g.secularEqnSCF(s.pot_coeffs, n)).reshape(wb.shape)
    err_a = np.abs(wa - wo.reshape(wb.shape)).max() * HARTREE_EV
    _log(f"  {case} Mode A eigenvalue error {err_a:.2e} eV")
    assert err_a < 1e-6
    g.close()


def test_amorphous_sio2_cell_matches_oracle(gpu):
    """The amorphous SiO2 workload's construction (Gamma only, real matrices,
    two element types, O 1s core states) on a 24-atom cell: the overlap,
    kinetic, nuclear and fused Hamiltonian sweeps (valeVale orthogonalised
    against the core, raw coreVale and coreCore) and the Mode-B eigenvalues
    against the oracle."""
    _log("asio2 24-atom cell")
    s = configs.asio2(seed=11, n_si=8, n_o=16)
    assert s.gamma and s.core_dim == 16
    g, o = _pair(gpu, s)
    So = o.integral(oracle.OL)
    checks = [("OL", g.gaussOverlapOL, lambda: So, True),
              ("KE", g.gaussOverlapKE, lambda: o.integral(oracle.KE), False),
              ("NP", g.gaussOverlapNP, lambda: o.integral(oracle.NP), False),
              ("HAM", lambda: g.gaussOverlapHamPSCF(s.pot_coeffs),
               lambda: o.integral(oracle.HAM, coef=s.pot_coeffs), False)]
    for name, fg, fo, isol in checks:
        mg, mo = fg(), fo()
        ref = oracle.ortho_dataset(mo[0], mo[1], mo[2], So[1], overlap=isol)
        assert _ortho_close(mg[0], ref, 1e-10), name
        assert block_rel_err(mg[1], mo[1]) < 1e-10, (name, "coreVale")
        assert block_rel_err(mg[2], mo[2]) < 1e-10, (name, "coreCore")
    S = oracle.ortho_dataset(*So, So[1], overlap=True)
    Ho = o.integral(oracle.HAM, coef=s.pot_coeffs)
    wo = oracle.solve(oracle.ortho_dataset(*Ho, So[1]), S)[0]
    w = np.asarray(g.scf_iteration(s.pot_coeffs, g.vale_dim)).reshape(-1)
    err = np.abs(w - np.asarray(wo).reshape(-1)).max() * HARTREE_EV
    _log(f"  eigenvalue error {err:.2e} eV")
    assert err < 1e-6
    g.close()


def test_amorphous_sio2_1000_full_size(gpu):
    """The 1000-atom amorphous SiO2 cube at full size (n = 5665 real,
    667 core states; the solve takes the library path): the real Gamma
    matrices, ascending finite eigenvalues, eigenpair residuals of the
    lowest and highest states |H x - w S x| within 1e-9 of the matrix scale,
    S-orthonormality within 1e-8, and the mixed form within 1 meV of FP64 on
    the occupied bands."""
    _log("asio2_1000 full")
    s = configs.asio2_1000()
    g = gpu.OlcaoGPU(s)
    n = g.vale_dim
    assert n == 5665 and s.core_dim == 667
    S = np.asarray(g.gaussOverlapOL()[0])
    H = np.asarray(g.gaussOverlapHamPSCF(s.pot_coeffs)[0])
    assert np.isrealobj(S) and np.isrealobj(H)
    S, H = S.reshape(n, n), H.reshape(n, n)
    w, v = g.secularEqnPSCF(n, vectors=True)
    w, v = np.asarray(w).reshape(-1), np.asarray(v).reshape(n, n)
    assert np.all(np.isfinite(w)) and np.all(np.diff(w) >= 0.0)
    Sf = np.triu(S) + np.triu(S, 1).T
    Hf = np.triu(H) + np.triu(H, 1).T
    sel = np.r_[0:8, n - 8:n]
    X = np.real(v[:, sel])
    R = Hf @ X - (Sf @ X) * w[sel]
    assert np.abs(R).max() < 1e-9 * np.abs(Hf).max() * max(1.0, np.abs(X).max())
    assert np.abs(X.T @ Sf @ X - np.eye(len(sel))).max() < 1e-8
    n_el = sum(s.pot_types[t].Z for t in s.pot_type) - 2 * s.core_dim
    n_occ = int(round(n_el / 2))
    w64 = np.asarray(g.scf_iteration(s.pot_coeffs, n_occ)).reshape(-1)
    g.close()
    gm = gpu.OlcaoGPU(s, precision=gpu.PREC_MIXED)
    wm = np.asarray(gm.scf_iteration(s.pot_coeffs, n_occ)).reshape(-1)
    _log(f"  timings {gm.timings()}")
    gm.close()
    err = np.abs(wm - w64).max() * HARTREE_EV * 1e3
    _log(f"  mixed vs FP64 occupied bands {err:.3e} meV")
    assert err < 1.0


BATCH_ORACLE_CELLS = (2, 6, 20, 59)   # the batch's 8-atom cells: minimal and extended
                                     # bases, Gamma (real) and 4 / 8 k-points


def test_uneven_batch_cells_match_oracle(gpu):
    """Cells of the uneven batch that the oracle finishes in seconds: S(k),
    H(k) and the Mode-B eigenvalues against the oracle, then the same cells
    through the batched call against their single-context iterations."""
    cells = configs.uneven_batch()
    ctxs = []
    for i in BATCH_ORACLE_CELLS:
        s = cells[i]
        _log(f"batch cell {i}: {s.n_atoms} atoms, n={s.vale_dim}, nk={s.nk}, gamma={s.gamma}")
        g, o = _pair(gpu, s)
        Sg = g.gaussOverlapOL()[0]
        Hg = g.gaussOverlapHamPSCF(s.pot_coeffs)[0]
        So = o.integral(oracle.OL)[0]
        Ho = o.integral(oracle.HAM, coef=s.pot_coeffs)[0]
        assert block_rel_err(Sg, So) < 1e-10, i
        assert block_rel_err(Hg, Ho) < 1e-10, i
        wo = np.asarray(oracle.solve(Ho, So)[0]).reshape(-1)
        w = np.asarray(g.scf_iteration(s.pot_coeffs, g.vale_dim)).reshape(-1)
        err = np.abs(w - wo).max() * HARTREE_EV
        _log(f"  eigenvalue error {err:.2e} eV")
        assert err < 1e-6, i
        ctxs.append((g, w))
    wb, ms = gpu.scf_iteration_batch([g for g, _ in ctxs], [g.vale_dim for g, _ in ctxs], 4)
    for (g, w), b in zip(ctxs, wb):
        assert np.abs(np.asarray(b).reshape(-1) - w).max() * HARTREE_EV < 1e-9
        g.close()
    assert np.all(ms > 0)


def test_uneven_batch_equals_single_cells(gpu):
    """All 64 cells of the uneven batch (8..216 atoms, n = 40..1296, 1..64
    k-points; every eigensolver path) in one batched call, twice in a row in
    each build mode (every cell on its own streams; the three-centre work of
    all cells in shared launches): each cell's eigenvalues equal its
    single-context Mode-B iteration's to 1e-12 of the cell's spectrum scale
    and 1e-6 eV (the library's BLAS calls use no atomics, so they are
    normally bit-identical), and they are finite and ascending."""
    cells = configs.uneven_batch()
    _log("uneven batch: creating 64 contexts")
    ctxs = [gpu.OlcaoGPU(s) for s in cells]
    ns = [g.vale_dim for g in ctxs]
    single = []
    for i, (g, n) in enumerate(zip(ctxs, ns)):
        single.append(np.asarray(g.scf_iteration(g.sysin.pot_coeffs, n)))
        _log(f"  cell {i} n={n} nk={g.nk}: {g.timings()['total_ms']:.2f} ms")
    for rep, mode in enumerate((gpu.BATCH_CELLS, gpu.BATCH_CELLS, gpu.BATCH_SHARED,
                                gpu.BATCH_SHARED)):
        wb, ms = gpu.scf_iteration_batch(ctxs, ns, 8, mode)
        worst, worst_rel, exact = 0.0, 0.0, 0
        for i, (w1, w2) in enumerate(zip(single, wb)):
            w2 = np.asarray(w2)
            assert w2.shape == w1.shape
            assert np.all(np.isfinite(w2)) and np.all(np.diff(w2, axis=-1) >= 0.0)
            d = float(np.abs(w2 - w1).max())
            if d != 0.0:
                _log(f"    cell {i} (n={ns[i]}, nk={ctxs[i].nk}, gamma={ctxs[i].gamma}) "
                     f"differs by {d * HARTREE_EV:.2e} eV")
            worst = max(worst, d)
            worst_rel = max(worst_rel, d / np.abs(w1).max())
            exact += int(d == 0.0)
        _log(f"  batch pass {rep} (mode {mode}): max |batch - single| {worst * HARTREE_EV:.2e} eV "
             f"({worst_rel:.1e} of the spectrum scale), {exact} of {len(single)} cells "
             f"bit-identical, latency p50 {np.percentile(ms, 50):.1f} ms "
             f"p99 {np.percentile(ms, 99):.1f} ms")
        assert worst_rel < 1e-12 and worst * HARTREE_EV < 1e-6
    for g in ctxs:
        g.close()


def test_batch_of_mixed_kinds_equals_single_cells(gpu):
    """The batched call over cells of every kind its build distinguishes:
    spin 1 and 2 (a second three-centre launch over the spin-2 cells), rel
    0 and 1 (kinetic blocks added by the finish or beforehand), core states,
    Gamma, f shells, untabulated and tabulated three-centre kernel forms,
    FP64 and mixed precision, in one call with the three-centre work in
    shared launches (twice in a row) and with every cell on its own streams;
    each cell's eigenvalues equal its single-context iteration's (the same
    kernels on the same data, in the same order) to 1e-12 of its spectrum
    scale."""
    specs = [(SMALL["rel_spin2"], gpu.PREC_FP64), (SMALL["core_gamma_rel"], gpu.PREC_FP64),
             (SMALL["spdf"], gpu.PREC_MIXED), (SMALL["core"], gpu.PREC_MIXED),
             (dict(core=True, spin=2, lmax=2, kmesh=(2, 2, 1)), gpu.PREC_FP64),
             (SMALL["hex_spd"], gpu.PREC_FP64), (SMALL["two_types"], gpu.PREC_MIXED)]
    ctxs = [gpu.OlcaoGPU(configs.small_test(**kw), precision=p) for kw, p in specs]
    ns = [g.vale_dim for g in ctxs]
    single = [np.asarray(g.scf_iteration(g.sysin.pot_coeffs, n)) for g, n in zip(ctxs, ns)]
    for rep, mode in enumerate((gpu.BATCH_SHARED, gpu.BATCH_SHARED, gpu.BATCH_CELLS)):
        wb, _ = gpu.scf_iteration_batch(ctxs, ns, 4, mode)
        for i, (w1, w2) in enumerate(zip(single, wb)):
            w2 = np.asarray(w2)
            d = float(np.abs(w2 - w1).max())
            _log(f"  batch {rep} (mode {mode}) cell {i} (spin {ctxs[i].spin}, n={ns[i]}): {d:.2e} Ha")
            assert w2.shape == w1.shape
            assert d <= 1e-12 * np.abs(w1).max(), (i, d)
    for g in ctxs:
        g.close()


VALE_CASES = {"cubic_sp": dict(lmax=1, kmesh=(2, 2, 1)),
              "core_gamma": dict(core=True, gamma=True, two_types=True),
              "rel_spin2": dict(lmax=1, rel=1, spin=2, kmesh=(2, 1, 1))}


@pytest.mark.parametrize("case", list(VALE_CASES))
def test_vale_charge_matches_oracle(gpu, case):
    """Valence density traces (makeValenceRho) after a Mode A solve on the
    device against the oracle's from its own terms and solve, with the
    same Fermi-smeared populations: electron energy, charge, nuclear,
    kinetic and mass-velocity traces and potRho to 1e-10 of their scale."""
    from tests.conftest import fermi_occupancy, mode_a_hamiltonian, mode_a_terms
    _log(f"vale charge {case}")
    s = configs.small_test(**VALE_CASES[case])
    g, o = _pair(gpu, s)
    t = mode_a_terms(o, s)
    coefs = np.asarray(s.pot_coeffs).reshape(s.spin, -1)
    ws, vs = [], []
    for q in range(s.spin):
        w, v = oracle.solve(mode_a_hamiltonian(t, coefs[q]), t["S"])
        ws.append(np.asarray(w).reshape(s.nk, -1))
        vs.append(np.asarray(v).reshape(s.nk, s.vale_dim, -1))
    n_el = sum(s.pot_types[p].Z for p in s.pot_type) - 2 * s.core_dim
    occ = fermi_occupancy(ws, s.kweights, n_el, s.spin)
    ee, tr, pr = oracle.vale_charge(ws, vs, occ, t["S"], t["NP"], t["KE"], t["V"], t["MV"])
    g.setupSCF()
    g.secularEqnSCF(s.pot_coeffs, g.vale_dim)
    d = g.makeValenceRho(occ)
    scale = max(1.0, np.abs(ee).max())
    assert np.abs(d["electron_energy"] - ee).max() < 1e-10 * scale
    got = np.stack([d["charge"], d["nuclear"], d["kinetic"], d["mass_velocity"]], axis=1)
    assert np.abs(got - tr).max() < 1e-10 * max(1.0, np.abs(tr).max())
    assert np.abs(d["pot_rho"] - pr).max() < 1e-10 * max(1.0, np.abs(pr).max())
    g.close()


def test_vale_charge_si64_invariants(gpu):
    """Si64 at full size (576 orbitals, 8 k): after a Mode A solve the
    charge trace equals the summed populations and the Hamiltonian traces
    sum_j c_j potRho_j + KE + NP equal sum f * eigenvalue."""
    import time
    from tests.conftest import fermi_occupancy
    _log("vale charge si64")
    s = configs.si64()
    g = gpu.OlcaoGPU(s)
    g.setupSCF()
    w = g.secularEqnSCF(s.pot_coeffs, g.vale_dim)
    n_el = sum(s.pot_types[p].Z for p in s.pot_type)
    occ = fermi_occupancy([w], s.kweights, n_el, 1)
    g.makeValenceRho(occ)
    t0 = time.perf_counter()
    d = g.makeValenceRho(occ)
    _log(f"  makeValenceRho {1e3 * (time.perf_counter() - t0):.2f} ms")
    total = occ.sum()
    assert abs(d["charge"][0] - total) < 1e-9 * total
    e_sum = np.asarray(s.pot_coeffs) @ d["pot_rho"][0] + d["kinetic"][0] + d["nuclear"][0]
    assert abs(e_sum - d["electron_energy"][0]) < 1e-9 * abs(d["electron_energy"][0])
    g.close()


def _check_k_overlap_file(lib, g, o, path, kg):
    """olcao_write_k_overlap_h5 into an integral file (kg None) or a PSCF
    k-point group: atomKOverlap and atomKOverlapPlusG hold, per axis and
    k-point, the full orthogonalised k-overlap as real/imag (valeDim,
    valeDim) datasets equal to the oracle's to 1e-10, status 1; a Gamma
    context is refused with OLCAO_ERR_UNSUPPORTED."""
    import ctypes as C

    from tests import h5read
    num_axial = np.array([2, 3, 1], np.int32)
    lib.olcao_write_k_overlap_h5.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p, C.c_char_p]
    rc = lib.olcao_write_k_overlap_h5(g.ctx, num_axial.ctypes.data,
                                      None if kg is None else kg.encode(), path.encode())
    if g.gamma:
        assert rc == 7
        return
    assert rc == 0
    top = "/atomIntgGroup/" if kg is None else f"/{kg}/atomIntgGroup/"
    S_cv = o.integral(oracle.OL)[1]
    refs = [oracle.ortho_k_overlap(*t, S_cv) for t in o.k_overlap(num_axial)]
    f = h5read.File(path)
    for group in ("atomKOverlap", "atomKOverlapPlusG"):
        assert f.status(top + group) == 1, group
        for x in range(3):
            for k in range(g.nk):
                for part, val in (("real", refs[x][k].real), ("imag", refs[x][k].imag)):
                    data, chunked, nfilt = f.read(f"{top}{group}/{x + 1:07d}/{part}{k + 1:07d}")
                    assert data.shape == (g.vale_dim, g.vale_dim) and chunked and nfilt >= 1
                    e = block_rel_err(data, val.T)
                    assert e < 1e-10, (group, x, k, part, e)
    f.close()


@pytest.mark.parametrize("case", [dict(core=True, kmesh=(2, 1, 1), spin=2),
                                  dict(core=True, gamma=True, rel=1, two_types=True)])
def test_scf_integral_file(gpu, tmp_path, case):
    """The setup's integral file in the reference's HDF5 layout: every term
    group and "%07d" dataset, packed (components, n(n+1)/2) datasets equal to
    the oracle's packed terms to 1e-10, coreVale real/imag (coreDim,
    valeDim) datasets, chunked and deflated, status = 1; then the eigenpairs
    of a Mode A solve added to the same file (eigenValues/<k><s>,
    eigenVectors/{real,imag}<k><s>, status 0) equal to what the solve
    returned."""
    import ctypes as C
    import os

    from tests import h5read
    from tests.conftest import ROOT, mode_a_terms
    if not h5read.available():
        pytest.skip("no HDF5 C library on this host")
    _log(f"integral file {case}")
    s = configs.small_test(**case)
    g, o = _pair(gpu, s)
    t = mode_a_terms(o, s)
    cv_raw = o.integral(oracle.OL)[1]
    g.setupSCF()
    path = str(tmp_path / "scf-integrals.hdf5")
    lib = C.CDLL(os.path.join(ROOT, "olcao_amd", "libolcao_h5.so"))
    lib.olcao_write_scf_integrals_h5.argtypes = [C.c_void_p, C.c_char_p]
    assert lib.olcao_write_scf_integrals_h5(g.ctx, path.encode()) == 0
    f = h5read.File(path)
    n, nk, comps = g.vale_dim, g.nk, 1 if s.gamma else 2
    groups = [("atomOverlap", t["S"]), ("atomKEOverlap", t["KE"]), ("atomNPOverlap", t["NP"])]
    groups += [("atomMVOverlap", t["MV"])] if s.rel else []
    groups += [(f"atomPotOverlap/{j + 1:07d}", v) for j, v in enumerate(t["V"])]
    for group, ref in groups:
        assert f.status("/atomIntgGroup/" + group) == 1, group
        for k in range(nk):
            data, chunked, nfilt = f.read(f"/atomIntgGroup/{group}/{k + 1:07d}")
            assert data.shape == (n * (n + 1) // 2, comps) and chunked and nfilt >= 1
            p = oracle.pack_upper(ref if s.gamma else ref[k])
            exp = p[:, None] if s.gamma else np.stack([p.real, p.imag], axis=1)
            assert _ortho_close(data, exp, 1e-10), (group, k)
    # every group, dataset and status attribute the reference's reader opens
    # (accessSCFIntegralHDF5, hdf5SCFIntg.F90:563-760), the terms outside
    # this path (and MV without rel) present with status 0
    top = "/atomIntgGroup/"
    for k in range(nk):
        kn = f"{k + 1:07d}"
        paths = [f"{top}{g}/{kn}" for g in ("atomOverlap", "atomKEOverlap", "atomMVOverlap",
                                             "atomNPOverlap")]
        paths += [f"{top}atomPotOverlap/{j + 1:07d}/{kn}" for j in range(s.pot_dim)]
        for x in range(1, 4):
            paths += [f"{top}{g}/{x:07d}/{kn}" for g in ("atomDMOverlap", "atomMMOverlap")]
            paths += [f"{top}{g}/{x:07d}/{part}{kn}" for g in ("atomKOverlap", "atomKOverlapPlusG")
                      for part in ("real", "imag")]
        paths += [f"{top}atomOverlapCV/{part}{kn}" for part in ("real", "imag")[:comps]]
        for name in paths:
            assert f.exists(name), name
    for group in ("atomDMOverlap", "atomMMOverlap", "atomKOverlap", "atomKOverlapPlusG"):
        assert f.status(top + group) == 0, group
    data, _, _ = f.read(f"{top}atomKOverlap/0000001/real0000001")
    assert data.shape == (n, n) and not data.any()
    assert f.status(top + "atomMVOverlap") == (1 if s.rel else 0)
    f.close()
    # the dipole and momentum groups filled in, as setupSCF does for
    # properties (OLCAO.F90:290-302)
    center = np.array([0.3, 0.2, -0.4])
    lib.olcao_write_scf_vector_terms_h5.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_char_p]
    assert lib.olcao_write_scf_vector_terms_h5(g.ctx, 3, center.ctypes.data, path.encode()) == 0
    f = h5read.File(path)
    S_cv = o.integral(oracle.OL)[1]
    for group, op in (("atomDMOverlap", oracle.DM), ("atomMMOverlap", oracle.MM)):
        assert f.status(top + group) == 1, group
        for x in range(3):
            ref = oracle.ortho_dataset(*o.integral(op, term=x, coef=center), S_cv)
            for k in range(nk):
                data, chunked, nfilt = f.read(f"{top}{group}/{x + 1:07d}/{k + 1:07d}")
                p = oracle.pack_upper(ref if s.gamma else ref[k])
                exp = p[:, None] if s.gamma else np.stack([p.real, p.imag], axis=1)
                assert data.shape == exp.shape and chunked and nfilt >= 1
                assert _ortho_close(data, exp, 1e-10), (group, x, k)
    for k in range(nk):
        cvk = cv_raw if s.gamma else cv_raw[k]
        for part, val in (("real", np.real(cvk)), ("imag", np.imag(cvk)))[:comps]:
            data, chunked, nfilt = f.read(f"/atomIntgGroup/atomOverlapCV/{part}{k + 1:07d}")
            assert data.shape == (n, s.core_dim) and chunked and nfilt >= 1
            assert block_rel_err(data, val.T) < 1e-10
    f.close()
    # the k-overlap groups, as setupSCF fills them with doMTOP = 1
    _check_k_overlap_file(lib, g, o, path, None)
    ns = n - 1
    w, v = g.secularEqnSCF(s.pot_coeffs, ns, vectors=True)
    w = np.asarray(w).reshape(s.spin, nk, ns)
    v = np.asarray(v).reshape(s.spin, nk, n, ns)
    lib.olcao_write_scf_eigen_h5.argtypes = [C.c_void_p, C.c_int, C.c_char_p]
    assert lib.olcao_write_scf_eigen_h5(g.ctx, ns, path.encode()) == 0
    f = h5read.File(path)
    assert f.exists("/atomIntgGroup/atomOverlap/0000001")   # the integral groups stay
    for q in range(s.spin):
        for k in range(nk):
            kq = f"{k + 1:07d}{q + 1:07d}"
            ev, chunked, nfilt = f.read(f"/eigenValues/{kq}")
            assert ev.shape == (ns,) and chunked and nfilt >= 1
            assert np.array_equal(ev, w[q, k])
            parts = [("", np.real(v[q, k]))] if s.gamma else \
                [("real", np.real(v[q, k])), ("imag", np.imag(v[q, k]))]
            for part, val in parts:
                data, chunked, nfilt = f.read(f"/eigenVectors/{part}{kq}")
                assert data.shape == (ns, n) and chunked and nfilt >= 1
                assert np.array_equal(data, val.T)
            assert f.dataset_status(f"/eigenVectors/{parts[0][0]}{kq}") == 0
    f.close()
    g.close()


@pytest.mark.parametrize("case", [dict(core=True, kmesh=(2, 2, 1)),
                                  dict(core=True, gamma=True, rel=1, two_types=True)])
def test_pscf_file(gpu, tmp_path, case):
    """The PSCF file of a Mode-B solve in the reference's layout, under its
    k-point group: every group, dataset and status attribute the reference
    reader opens (accessPSCFIntegralHDF5, hdf5PSCFIntg.F90:455-672); S(k)
    and H(k) packed and orthogonalised as the oracle's to 1e-10, coreVale
    raw, the eigenpairs equal to what the solve returned; a second write
    replaces the group; spin 2 refused as the reference stops there."""
    import ctypes as C
    import os

    from tests import h5read
    from tests.conftest import ROOT
    if not h5read.available():
        pytest.skip("no HDF5 C library on this host")
    s = configs.small_test(**case)
    g, o = _pair(gpu, s)
    n, nk, comps, ns = g.vale_dim, g.nk, 1 if s.gamma else 2, g.vale_dim - 1
    coef = np.ascontiguousarray(s.pot_coeffs, np.float64)
    g.gaussOverlapOL()
    g.gaussOverlapHamPSCF(coef)
    w, v = g.secularEqnPSCF(ns, vectors=True)
    w = np.asarray(w).reshape(nk, ns)
    v = np.asarray(v).reshape(nk, n, ns)
    lib = C.CDLL(os.path.join(ROOT, "olcao_amd", "libolcao_h5.so"))
    lib.olcao_write_pscf_h5.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_char_p, C.c_char_p]
    path = str(tmp_path / "gs_pscf-fb.hdf5")
    kg = "00002_00002_00001"
    for _ in range(2):
        assert lib.olcao_write_pscf_h5(g.ctx, coef.ctypes.data, ns, kg.encode(), path.encode()) == 0
    f = h5read.File(path)
    top = f"/{kg}/atomIntgGroup/"
    So = o.integral(oracle.OL)
    Ho = o.integral(oracle.HAM, coef=coef)
    refs = {"atomOverlap": oracle.ortho_dataset(*So, So[1], overlap=True),
            "atomHamOverlap": oracle.ortho_dataset(*Ho, So[1], overlap=False)}
    for group, ref in refs.items():
        assert f.status(top + group) == 1, group
        for k in range(nk):
            data, chunked, nfilt = f.read(f"{top}{group}/{k + 1:07d}")
            p = oracle.pack_upper(ref if s.gamma else ref[k])
            exp = p[:, None] if s.gamma else np.stack([p.real, p.imag], axis=1)
            assert data.shape == exp.shape and chunked and nfilt >= 1
            assert _ortho_close(data, exp, 1e-10), (group, k)
    for k in range(nk):
        kn = f"{k + 1:07d}"
        paths = [f"{top}atomOverlap/{kn}", f"{top}atomHamOverlap/{kn}"]
        for x in range(1, 4):
            paths += [f"{top}{gr}/{x:07d}/{kn}" for gr in ("atomDMOverlap", "atomMMOverlap")]
            paths += [f"{top}{gr}/{x:07d}/{part}{kn}" for gr in ("atomKOverlap", "atomKOverlapPlusG")
                      for part in ("real", "imag")]
        paths += [f"{top}atomOverlapCV/{part}{kn}" for part in ("real", "imag")[:comps]]
        for name in paths:
            assert f.exists(name), name
        cvk = So[1] if s.gamma else So[1][k]
        for part, val in (("real", np.real(cvk)), ("imag", np.imag(cvk)))[:comps]:
            data, _, _ = f.read(f"{top}atomOverlapCV/{part}{kn}")
            assert block_rel_err(data, val.T) < 1e-10
        ev, _, _ = f.read(f"/{kg}/eigenValues/{kn}0000001")
        assert np.array_equal(ev, w[k])
        parts = [("", np.real(v[k]))] if s.gamma else [("real", np.real(v[k])), ("imag", np.imag(v[k]))]
        for part, val in parts:
            data, _, _ = f.read(f"/{kg}/eigenVectors/{part}{kn}0000001")
            assert data.shape == (ns, n) and np.array_equal(data, val.T)
        assert f.dataset_status(f"/{kg}/eigenVectors/{parts[0][0]}{kn}0000001") == 0
    for group in ("atomDMOverlap", "atomMMOverlap", "atomKOverlap", "atomKOverlapPlusG"):
        assert f.status(top + group) == 0, group
    f.close()
    _check_k_overlap_file(lib, g, o, path, kg)
    g.close()
    s2 = configs.small_test(**SMALL["rel_spin2"])
    g2 = gpu.OlcaoGPU(s2)
    c2 = np.ascontiguousarray(s2.pot_coeffs, np.float64)
    g2.scf_iteration(c2, g2.vale_dim)
    assert lib.olcao_write_pscf_h5(g2.ctx, c2.ctypes.data, g2.vale_dim, kg.encode(),
                                   str(tmp_path / "spin2.hdf5").encode()) == 7
    g2.close()


def test_diamond_c8_full_parity(gpu):
    """First workload at full size: H(k), S(k) and all eigenvalues."""
    _log("diamond_c8 full")
    s = configs.diamond_c8()
    g, o = _pair(gpu, s)
    Sg = g.gaussOverlapOL()[0]
    Hg = g.gaussOverlapHamPSCF(s.pot_coeffs)[0]
    So = o.integral(oracle.OL)[0]
    Ho = o.integral(oracle.HAM, coef=s.pot_coeffs)[0]
    assert block_rel_err(Sg, So) < 1e-10
    assert block_rel_err(Hg, Ho) < 1e-10
    wg = g.secularEqnPSCF(g.vale_dim)
    wo, _ = oracle.solve(Ho, So)
    err_ev = np.abs(wg - wo).max() * HARTREE_EV
    _log(f"  max eigenvalue error {err_ev:.2e} eV")
    assert err_ev < 1e-6
    w2 = g.scf_iteration(s.pot_coeffs, g.vale_dim)
    assert np.abs(w2 - wo).max() * HARTREE_EV < 1e-6
    g.close()


def test_mode_a_equals_mode_b(gpu):
    s = configs.small_test(lmax=2, two_types=False)
    g = gpu.OlcaoGPU(s)
    g.setupSCF()
    wa = g.secularEqnSCF(s.pot_coeffs, g.vale_dim)
    wb = g.scf_iteration(s.pot_coeffs, g.vale_dim)
    assert np.abs(wa - wb).max() * HARTREE_EV < 1e-6
    # a zero coefficient adds its term unscaled (matrixSubs.F90:369-373)
    c = s.pot_coeffs.copy()
    c[1] = 0.0
    wz = g.secularEqnSCF(c, g.vale_dim)
    c1 = c.copy()
    c1[1] = 1.0
    w1 = g.secularEqnSCF(c1, g.vale_dim)
    assert np.abs(wz - w1).max() < 1e-12
    g.close()


def test_eigenvectors_and_gamma_solve(gpu):
    s = configs.small_test(gamma=True, lmax=2)
    g, o = _pair(gpu, s)
    S = g.gaussOverlapOL()[0]
    H = g.gaussOverlapHamPSCF(s.pot_coeffs)[0]
    w, v = g.secularEqnPSCF(4, vectors=True)
    wo, vo = oracle.solve(o.integral(oracle.HAM, coef=s.pot_coeffs)[0], o.integral(oracle.OL)[0])
    assert np.abs(w[0] - wo[:4]).max() * HARTREE_EV < 1e-6
    Sf = np.triu(S) + np.triu(S, 1).T
    Hf = np.triu(H) + np.triu(H, 1).T
    assert np.abs(Hf @ v - Sf @ v * w[0]).max() < 1e-9
    g.close()


def test_drop_in_solvers(gpu):
    from scipy.linalg import eigh
    rng = np.random.default_rng(2)
    n, nb = 37, 5
    Z = rng.normal(size=(nb, n, n)) + 1j * rng.normal(size=(nb, n, n))
    B = Z @ np.conj(Z.transpose(0, 2, 1)) + n * np.eye(n)
    A = rng.normal(size=(nb, n, n)) + 1j * rng.normal(size=(nb, n, n))
    A = A + np.conj(A.transpose(0, 2, 1))
    w, V = gpu.solveZHEGV(np.triu(A), np.triu(B))
    for k in range(nb):
        assert np.allclose(w[k], eigh(A[k], B[k], eigvals_only=True), atol=1e-9)
        assert np.abs(A[k] @ V[k] - B[k] @ V[k] * w[k]).max() < 1e-8
    Ar, Br = A.real, (Z.real @ Z.real.transpose(0, 2, 1)) + n * np.eye(n)
    wr, _ = gpu.solveDSYGV(np.triu(Ar), np.triu(Br))
    for k in range(nb):
        assert np.allclose(wr[k], eigh(Ar[k], Br[k], eigvals_only=True), atol=1e-9)


def _pencil(rng, n, nb, cplx, degenerate):
    """Random Hermitian pencils (A, B), B positive definite; with
    degenerate, A has every eigenvalue (of B^-1 A) three times over."""
    def rnd(*shape):
        x = rng.normal(size=shape)
        return x + 1j * rng.normal(size=shape) if cplx else x
    Z = rnd(nb, n, n)
    B = Z @ np.conj(Z.transpose(0, 2, 1)) / n + np.eye(n)
    if degenerate:
        L = np.linalg.cholesky(B)
        Q, _ = np.linalg.qr(rnd(nb, n, n))
        lam = np.repeat(rng.normal(size=(nb, (n + 2) // 3)), 3, axis=1)[:, :n]
        C = Q @ (lam[:, :, None] * np.conj(Q.transpose(0, 2, 1)))
        A = L @ C @ np.conj(L.transpose(0, 2, 1))
        A = 0.5 * (A + np.conj(A.transpose(0, 2, 1)))
    else:
        A = rnd(nb, n, n)
        A = A + np.conj(A.transpose(0, 2, 1))
    return A, B


@pytest.mark.parametrize("n,nb,cplx,degenerate", [(100, 3, True, False), (331, 2, True, True),
                                                  (576, 9, True, False), (200, 2, False, False),
                                                  (640, 2, False, True), (648, 5, True, False),
                                                  (720, 3, True, False), (768, 2, False, True),
                                                  (768, 5, True, True), (864, 3, False, False),
                                                  (896, 2, True, False)])
def test_tridiagonal_path_solvers(gpu, n, nb, cplx, degenerate):
    """Sizes past the one-workgroup LDS solver take the cooperative
    tridiagonal reduction: eigenvalues against LAPACK (scipy eigh) within
    1e-10 of the spectrum's scale, residuals |A x - w B x| and
    |X^H B X - I| within 1e-9 (relative), for random and for threefold
    degenerate spectra, complex and real, and batches larger than the
    matrices in flight at once (576 x 9)."""
    from scipy.linalg import eigh
    _log(f"tridiagonal path n={n} nb={nb} cplx={cplx} degenerate={degenerate}")
    A, B = _pencil(np.random.default_rng(n + nb), n, nb, cplx, degenerate)
    solve = gpu.solveZHEGV if cplx else gpu.solveDSYGV
    w, X = solve(np.triu(A), np.triu(B))
    for k in range(nb):
        wr = eigh(A[k], B[k], eigvals_only=True)
        scale = np.abs(wr).max()
        assert np.abs(w[k] - wr).max() < 1e-10 * scale, (k, np.abs(w[k] - wr).max())
        R = A[k] @ X[k] - B[k] @ X[k] * w[k]
        assert np.abs(R).max() < 1e-9 * scale * np.abs(B[k]).max()
        O = np.conj(X[k].T) @ B[k] @ X[k]
        assert np.abs(O - np.eye(n)).max() < 1e-9


def _check_pencil_solution(A, B, w, X, n, nb):
    from scipy.linalg import eigh
    for k in range(nb):
        wr = eigh(A[k], B[k], eigvals_only=True)
        scale = np.abs(wr).max()
        assert np.abs(w[k] - wr).max() < 1e-10 * scale, (k, np.abs(w[k] - wr).max())
        R = A[k] @ X[k] - B[k] @ X[k] * w[k]
        assert np.abs(R).max() < 1e-9 * scale * np.abs(B[k]).max()
        O = np.conj(X[k].T) @ B[k] @ X[k]
        assert np.abs(O - np.eye(n)).max() < 1e-9


@pytest.mark.parametrize("n,nb,cplx,degenerate", [(100, 2, True, False), (331, 2, True, True),
                                                  (576, 3, True, False), (200, 2, False, False),
                                                  (640, 2, False, True), (896, 2, True, False)])
def test_blocked_reduction_small_sizes(gpu, monkeypatch, n, nb, cplx, degenerate):
    """The global-memory blocked reduction (panels of 32 columns, trailing
    her2k/syr2k) forced where the LDS-resident one would run
    (OLCAO_EIGENSOLVER=blocked), so that it meets the divide and conquer of
    the smaller sizes: eigenvalues against LAPACK (scipy eigh) within 1e-10
    of the scale, residuals and B-orthonormality within 1e-9."""
    monkeypatch.setenv("OLCAO_EIGENSOLVER", "blocked")
    _log(f"blocked reduction n={n} nb={nb} cplx={cplx} degenerate={degenerate}")
    A, B = _pencil(np.random.default_rng(7 * n + nb), n, nb, cplx, degenerate)
    solve = gpu.solveZHEGV if cplx else gpu.solveDSYGV
    w, X = solve(np.triu(A), np.triu(B))
    _check_pencil_solution(A, B, w, X, n, nb)


@pytest.mark.parametrize("n,nb,cplx,degenerate", [(1000, 2, False, False), (1080, 2, True, False),
                                                  (1296, 3, True, True), (2000, 1, False, True),
                                                  (2600, 1, True, False), (3001, 1, False, False)])
def test_large_device_solvers(gpu, n, nb, cplx, degenerate):
    """Sizes past the LDS-resident reduction: the blocked reduction and the
    divide and conquer with its merges past 896 in global memory (the
    batch's n = 1080 and 1296 cells, and real sizes on the way to the
    1000-atom cell's 5665): the same bounds against LAPACK."""
    _log(f"large device solver n={n} nb={nb} cplx={cplx} degenerate={degenerate}")
    A, B = _pencil(np.random.default_rng(n + 5 * nb), n, nb, cplx, degenerate)
    solve = gpu.solveZHEGV if cplx else gpu.solveDSYGV
    w, X = solve(np.triu(A), np.triu(B))
    _check_pencil_solution(A, B, w, X, n, nb)


def _tridiagonal_cases(rng):
    w21 = np.abs(np.arange(21) - 10.0)
    yield "glued Wilkinson", np.tile(w21, 7), np.concatenate([np.ones(20), [1e-8]] * 7)[:-1]
    yield "graded", 10.0 ** rng.uniform(-3, 3, 150), 10.0 ** rng.uniform(-3, 1, 149)
    yield "tiny couplings", np.repeat(rng.normal(size=50), 3), np.full(149, 1e-15)
    yield "wide spread", np.concatenate([rng.normal(size=300), 1e3 * rng.random(276)]), rng.normal(size=575)


def test_tridiagonal_divide_and_conquer_cases(gpu):
    """Tridiagonal A with B = I through the real drop-in solver: the
    reduction returns A's own tridiagonal form, so the divide and conquer
    eigensolver meets its hard cases directly: glued Wilkinson matrices
    (close pairs, rotation deflation), graded entries, negligible couplings
    with repeated diagonal entries (small-z deflation) and a wide spectrum.
    Eigenvalues against scipy within 1e-12 of the scale; residual and
    orthogonality within 1e-11."""
    from scipy.linalg import eigh_tridiagonal
    rng = np.random.default_rng(7)
    for name, d, e in _tridiagonal_cases(rng):
        n = len(d)
        _log(f"tridiagonal case {name} n={n}")
        A = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        w, X = gpu.solveDSYGV(np.triu(A)[None], np.eye(n)[None])
        wr = eigh_tridiagonal(d, e, eigvals_only=True)
        scale = np.abs(wr).max()
        assert np.abs(w[0] - wr).max() < 1e-12 * scale, (name, np.abs(w[0] - wr).max())
        assert np.abs(A @ X[0] - X[0] * w[0]).max() < 1e-11 * scale, name
        assert np.abs(X[0].T @ X[0] - np.eye(n)).max() < 1e-11, name


def test_blocked_cholesky_reports_indefinite_overlap(gpu):
    """Sizes on the tridiagonal path factor S with the blocked Cholesky: an
    overlap whose leading minor fails in the second block row (column 71) is
    NOT_SPD, as zhegv reports info = n + 71."""
    n = 150
    B = np.eye(n, dtype=complex)
    B[70, 70] = -1.0
    A = np.diag(np.arange(n, dtype=float)).astype(complex)
    with pytest.raises(gpu.OlcaoError) as e:
        gpu.solveZHEGV(A[None], B[None])
    assert e.value.code == 5


def test_error_paths(gpu):
    # indefinite overlap -> NOT_SPD, the condition on which the reference stops
    with pytest.raises(gpu.OlcaoError) as e:
        gpu.solveZHEGV(np.eye(3, dtype=complex), np.diag([1.0, -1.0, 1.0]).astype(complex))
    assert e.value.code == 5
    # non-ascending exponents -> INPUT
    s = configs.small_test()
    s.types[0].alphas = s.types[0].alphas[::-1].copy()
    with pytest.raises((gpu.OlcaoError, ValueError)):
        gpu.OlcaoGPU(s)
    # solve before build -> STATE
    g = gpu.OlcaoGPU(configs.small_test())
    with pytest.raises(gpu.OlcaoError) as e:
        g.secularEqnPSCF(2)
    assert e.value.code == 8
    g.close()


def test_si64_full_size_properties(gpu):
    """Benchmark workload at full size: S(k) positive definite (the solve
    succeeds), H(-k) = conj H(k) over the shifted 2x2x2 mesh, Mode A
    (setup terms, then H = NP + KE + sum_j c_j V_j) equals Mode B (the fused
    rebuild) on the same coefficients to 1e-6 eV over all states."""
    _log("si64 full")
    s = configs.si64()
    g = gpu.OlcaoGPU(s)
    H = g.gaussOverlapHamPSCF(s.pot_coeffs)[0]
    S = g.gaussOverlapOL()[0]
    k = s.kpoints
    for a in range(len(k)):
        b = np.where(np.all(np.abs(k + k[a]) < 1e-12, axis=1))[0][0]
        assert block_rel_err(H[b], np.conj(H[a])) < 1e-11
        assert block_rel_err(S[b], np.conj(S[a])) < 1e-11
    n = g.vale_dim
    wb = g.scf_iteration(s.pot_coeffs, n)
    assert np.all(np.isfinite(wb))
    _log(f"  timings {g.timings()}")
    g.setupSCF()
    wa = g.secularEqnSCF(s.pot_coeffs, n)
    err = np.abs(wa - wb).max() * HARTREE_EV
    _log(f"  Mode A vs Mode B max |dE| {err:.2e} eV")
    assert err < 1e-6
    g.close()


def _pair_blocks(s, pidx):
    """(row slice, column slice) of the atom pairs i <= j numbered row-major
    over i, as the oracle's pair sampling numbers them."""
    off = np.cumsum([0] + [s.types[t].num_vale_states for t in s.atom_type])
    pairs = [(i, j) for i in range(s.n_atoms) for j in range(i, s.n_atoms)]
    return [(slice(off[pairs[p][0]], off[pairs[p][0] + 1]),
             slice(off[pairs[p][1]], off[pairs[p][1] + 1])) for p in pidx]


@pytest.mark.parametrize("make,stride,offset", [("si64", 97, 5), ("quartz", 4, 1)])
def test_full_size_sampled_pairs_match_oracle(gpu, make, stride, offset):
    """The benchmark (Si64, 576 orbitals, 8 k) and quartz (512 k) at full
    size against the oracle on a sample of atom pairs (every stride-th pair
    i <= j from offset; the oracle leaves the other blocks zero): S(k) and
    the fused H(k) blocks of those pairs for every k, and the mixed form's H
    blocks, within 1e-10 (FP64) and 1e-6 (mixed) of the sampled scale."""
    _log(f"{make} sampled pairs, stride {stride}")
    s = getattr(configs, make)()
    g = gpu.OlcaoGPU(s)
    Sg = g.gaussOverlapOL()[0]
    Hg = g.gaussOverlapHamPSCF(s.pot_coeffs)[0]
    g.close()
    gm = gpu.OlcaoGPU(s, precision=gpu.PREC_MIXED)
    Hm = gm.gaussOverlapHamPSCF(s.pot_coeffs)[0]
    gm.close()
    o = oracle.Oracle(s)
    oracle.Oracle.set_pair_sample(stride, offset)
    try:
        So = o.integral(oracle.OL)[0]
        Ho = o.integral(oracle.HAM, coef=s.pot_coeffs)[0]
    finally:
        oracle.Oracle.set_pair_sample(1, 0)
    npairs = s.n_atoms * (s.n_atoms + 1) // 2
    blocks = _pair_blocks(s, range(offset, npairs, stride))

    def take(M):
        return np.concatenate([M[:, r, c].reshape(M.shape[0], -1) for r, c in blocks], axis=1)
    es, eh, em = (block_rel_err(take(Sg), take(So)), block_rel_err(take(Hg), take(Ho)),
                  block_rel_err(take(Hm), take(Ho)))
    _log(f"  {len(blocks)} pairs: S {es:.2e}, H {eh:.2e}, mixed H {em:.2e}")
    assert es < 1e-10 and eh < 1e-10
    assert em < 1e-6


def band_energy(w, kweights, n_occ):
    """Band-energy proxy of the total energy: sum_k w_k sum_{n < n_occ} e_nk
    (the k weights carry the spin factor 2)."""
    return float(np.sum(np.asarray(kweights)[:, None] * np.asarray(w)[:, :n_occ]))


@pytest.mark.parametrize("make", ["si64", "diamond_c8", "quartz", "asio2_1000"])
def test_mixed_precision_form_tolerances(gpu, make):
    """The chosen (mixed-precision) form against the FP64 form on the full
    workloads: occupied-band eigenvalues within 1 meV, band energy within
    1e-5 Ha/atom and the fused Hamiltonian within 1e-6 of the matrix scale.
    The error over all states is reported: the top of these spectra lies at
    1e4-1e6 Ha, where the FP32 exponential's ~1e-7 relative error in the
    potential terms can exceed 1 meV in absolute terms."""
    s = getattr(configs, make)()
    g64 = gpu.OlcaoGPU(s)
    gmx = gpu.OlcaoGPU(s, precision=gpu.PREC_MIXED)
    n = g64.vale_dim
    w64 = g64.scf_iteration(s.pot_coeffs, n)
    wmx = gmx.scf_iteration(s.pot_coeffs, n)
    n_el = sum(s.pot_types[t].Z for t in s.pot_type)
    n_occ = min(n, int(round(n_el / 2)))
    err = np.abs(w64[:, :n_occ] - wmx[:, :n_occ]).max() * HARTREE_EV * 1e3
    err_all = np.abs(w64 - wmx).max() * HARTREE_EV * 1e3
    de = abs(band_energy(w64, s.kweights, n_occ) - band_energy(wmx, s.kweights, n_occ)) / s.n_atoms
    H64 = g64.gaussOverlapHamPSCF(s.pot_coeffs)[0]
    Hmx = gmx.gaussOverlapHamPSCF(s.pot_coeffs)[0]
    eh = block_rel_err(Hmx, H64)
    _log(f"  {make}: mixed vs FP64 occupied-band max |dE| {err:.3e} meV (all states "
         f"{err_all:.3e} meV), band energy {de:.2e} Ha/atom, H rel err {eh:.2e}; times "
         f"{gmx.timings()['total_ms']:.2f} vs {g64.timings()['total_ms']:.2f} ms")
    assert err < 1.0
    assert de < 1e-5
    assert eh < 1e-6
    g64.close()
    gmx.close()


@pytest.mark.parametrize("case", ["core", "core_gamma_rel", "quartz"])
def test_bloch_kernels_agree(gpu, case):
    """The Bloch sum on the matrix cores against the vector-ALU kernel on
    the same device blocks, within each form: FP64 MFMA within 1e-14 of the
    matrix scale (also when the mixed form takes it) (the FP64 MFMA adds its four products per step in cell
    order, as the vector kernel's FMAs do, so the two agree bit for bit);
    the mixed form's exact BF16 slice products within 1e-11, a tenth of the
    integral parity bar (their truncation is about 2^-48 of each block
    column's largest entry per cell); overlap, fused Hamiltonian and the
    core blocks, complex and Gamma builds, and quartz's 512 k with up to
    ~900 cells per pair."""
    s = configs.quartz() if case == "quartz" else configs.small_test(**SMALL[case])
    res = {}
    for name, prec, bloch in (("valu", gpu.PREC_FP64, gpu.BLOCH_VALU),
                              ("mfma", gpu.PREC_FP64, gpu.BLOCH_MFMA),
                              ("valu_mixed", gpu.PREC_MIXED, gpu.BLOCH_VALU),
                              ("split", gpu.PREC_MIXED, gpu.BLOCH_MFMA),
                              ("mfma64_mixed", gpu.PREC_MIXED, gpu.BLOCH_MFMA64)):
        g = gpu.OlcaoGPU(s, precision=prec, bloch=bloch)
        res[name] = (g.gaussOverlapOL(), g.gaussOverlapHamPSCF(s.pot_coeffs))
        g.close()
    for name, ref, tol in (("mfma", "valu", 1e-14), ("split", "valu_mixed", 1e-11),
                           ("mfma64_mixed", "valu_mixed", 1e-14)):
        for a, b in zip(res[name], res[ref]):
            for x, y in zip(a, b):
                if y is None:
                    continue
                e = block_rel_err(x, y)
                _log(f"  {case} {name}: {e:.2e}")
                assert _ortho_close(x, y, tol), (case, name, e)


@pytest.mark.parametrize("make", ["diamond_c8", "quartz", "core_spin2", "si64", "asio2_1000"])
def test_graph_replay_matches_eager(gpu, make):
    """Graph mode: the first iteration runs eagerly and captures the
    iteration's device work; each later one replays it (asio2_1000: the
    blocked reduction, the large divide-and-conquer merges and the split-K
    GEMMs of one n = 5665 matrix inside the graph).  Three replays with
    other eager work and synchronisations between them, the last two with
    new coefficients, each against an eager context's iteration on the same
    coefficients (bit for bit: the same kernels in the same order); then a
    reallocation (an eigenvector copy) forces an eager run and a new capture,
    again equal to the eager result."""
    if make == "core_spin2":
        s = configs.small_test(core=True, spin=2, lmax=2, kmesh=(2, 2, 1))
    else:
        s = getattr(configs, make)()
    ge = gpu.OlcaoGPU(s)
    gg = gpu.OlcaoGPU(s, graph=True)
    n = ge.vale_dim
    c0 = np.asarray(s.pot_coeffs, np.float64)
    w = gg.scf_iteration(c0, n)
    state, why = gg.graph_state()
    _log(f"  {make}: graph state {state} {why}")
    assert state == 1, why
    assert np.array_equal(w, ge.scf_iteration(c0, n))
    for rep, cf in enumerate((c0, c0 * 1.01, c0 * 0.97)):
        wg = gg.scf_iteration(cf, n)
        assert gg.timings()["graph_replay"]
        we = ge.scf_iteration(cf, n)          # eager work between the replays
        ge.gaussOverlapOL()
        d = float(np.abs(wg - we).max())
        _log(f"  {make} replay {rep}: max |graph - eager| {d:.2e} Ha, "
             f"{gg.timings()['total_ms']:.2f} ms vs {ge.timings()['total_ms']:.2f} ms eager")
        assert d <= 1e-12 * np.abs(we).max()
    gg.secularEqnPSCF(2, vectors=True)           # allocates: the graph is stale
    assert gg.graph_state()[0] == 2
    wg = gg.scf_iteration(c0, n)
    assert not gg.timings()["graph_replay"] and gg.graph_state()[0] == 1
    assert np.array_equal(wg, ge.scf_iteration(c0, n))
    wg = gg.scf_iteration(c0, n)
    assert gg.timings()["graph_replay"]
    assert float(np.abs(wg - ge.scf_iteration(c0, n)).max()) <= 1e-12 * np.abs(wg).max()
    ge.close()
    gg.close()


def test_mixed_precision_uneven_batch_tolerances(gpu):
    """The uneven batch (64 cells) in the mixed form against the FP64 form,
    both through the batched call: every cell's occupied-band eigenvalues
    within 1 meV and its band energy within 1e-5 Ha/atom."""
    cells = configs.uneven_batch()
    _log("uneven batch, mixed vs FP64")
    g64 = [gpu.OlcaoGPU(s) for s in cells]
    gmx = [gpu.OlcaoGPU(s, precision=gpu.PREC_MIXED) for s in cells]
    ns = [g.vale_dim for g in g64]
    w64, _ = gpu.scf_iteration_batch(g64, ns, 8)
    wmx, _ = gpu.scf_iteration_batch(gmx, ns, 8)
    worst_ev, worst_e = 0.0, 0.0
    for i, s in enumerate(cells):
        a = np.asarray(w64[i]).reshape(s.nk, -1)
        b = np.asarray(wmx[i]).reshape(s.nk, -1)
        n_el = sum(s.pot_types[t].Z for t in s.pot_type)
        n_occ = min(ns[i], int(round(n_el / 2)))
        ev = np.abs(a[:, :n_occ] - b[:, :n_occ]).max() * HARTREE_EV * 1e3
        de = abs(band_energy(a, s.kweights, n_occ) - band_energy(b, s.kweights, n_occ)) / s.n_atoms
        worst_ev, worst_e = max(worst_ev, ev), max(worst_e, de)
        assert ev < 1.0 and de < 1e-5, (i, ev, de)
    _log(f"  worst cell: occupied bands {worst_ev:.3e} meV, band energy {worst_e:.2e} Ha/atom")
    for g in g64 + gmx:
        g.close()


@pytest.mark.parametrize("make", ["diamond_c8", "quartz", "si64"])
def test_lds_eigensolver_matches_library(gpu, make):
    """The device eigensolvers against rocSOLVER's zpotrf/ztrsm/zheevd on
    the same device-resident H(k), S(k): the batched in-LDS solver (one
    workgroup per k: Cholesky, reduction, Jacobi) for diamond and quartz, the
    cooperative tridiagonal path for Si64 (n = 576): all eigenvalues within
    1e-6 eV."""
    s = getattr(configs, make)()
    g1 = gpu.OlcaoGPU(s)
    g2 = gpu.OlcaoGPU(s, eigensolver=gpu.EIG_LIBRARY)
    n = g1.vale_dim
    w1 = g1.scf_iteration(s.pot_coeffs, n)
    w2 = g2.scf_iteration(s.pot_coeffs, n)
    err = np.abs(w1 - w2).max() * HARTREE_EV
    _log(f"  {make}: LDS vs rocSOLVER max |dE| {err:.2e} eV; solve "
         f"{g1.timings()['solve_ms']:.2f} vs {g2.timings()['solve_ms']:.2f} ms")
    assert err < 1e-6
    g1.close()
    g2.close()
