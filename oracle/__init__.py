"""CPU oracle for the OLCAO integral -> H(k)/S(k) -> eigenvalue path.

TEST INFRASTRUCTURE ONLY.  This package restates the reference algorithm on
the CPU (plain C in olcao_oracle.c for the integral sweeps, numpy/LAPACK
here for the small derived setup, orthogonalisation, packing, assembly and
the generalised eigensolve).  It is the checker for the HIP product path and
the timed CPU baseline; only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg import it.  The product package (olcao_amd) never does.

Reference files restated here (file:line):
  basis        src/uolcao/basis.f90:55-243          renormalizeBasis
  lattice      src/uolcao/lattice.f90:174-237       getRecipCellVectors
               src/uolcao/lattice.f90:310-443       initializeLattice (C)
  k-points     src/uolcao/kpoints.f90:424-454       convertKPointsToXYZ
  ortho        src/uolcao/intgOrtho.F90:23-211      (coreDim > 0)
  packing      src/uolcao/matrixSubs.F90:339-534    pack / readPackedMatrixAccum
  assembly     src/uolcao/secularEqn.F90:147-170    H = NP + KE + sum c_j V_j
  solve        src/uolcao/interfaces.F90:73-200     zhegv/dsygv(1,'V','U')

Parity status of each function is stated in its docstring; every part is
pinned by tests/test_oracle.py (golden vectors from the compiled reference
primitives, closed forms, brute force and library cross-checks).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libolcao_ref_prims.so")

OL, KE, NP, EP, HAM, MV, DM, MM, KO, FORCE = 0, 1, 2, 3, 4, 5, 6, 7, 8, 9
SMALL_THRESH = 1.0e-8

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        _lib = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        _lib.orc_boys.argtypes = [C.c_double, dp]
        _lib.orc_prim_batch.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp,
                                        C.POINTER(C.c_int), dp]
        _lib.orc_neglig_limit.restype = C.c_double
        _lib.orc_neglig_limit.argtypes = [dp, C.c_double, C.c_double, C.c_double]
        _lib.orc_make_lattice.argtypes = [dp, C.c_double, C.c_int, dp, dp]
        _lib.orc_nearest_lattice_vector.argtypes = [dp, dp, dp]
        _lib.orc_integral.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, dp, dp, dp, C.c_int]
        _lib.orc_last_counts.argtypes = [C.POINTER(C.c_longlong)] * 3
        _lib.orc_set_pair_sample.argtypes = [C.c_int, C.c_int]
        _lib.orc_set_three_center_range_scale.argtypes = [C.c_double]
        _lib.orc_primk_batch.argtypes = [C.c_int, dp, dp, dp, dp, C.POINTER(C.c_int), dp, dp]
        _lib.orc_primd_batch.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, C.POINTER(C.c_int), dp]
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


# ---------------------------------------------------------------------------
# primitives
def boys(T):
    F = np.zeros(7)
    lib().orc_boys(float(T), _dp(F))
    return F


def prim_batch(op, a, A, B, C_, sw):
    """Primitive blocks sh(16,16) (column-major, returned as [n,16,16] with
    [i, row, col]) for op 0 overlap, 1 kinetic, 2 three-centre overlap,
    3 nuclear (gaussIntegrals.f90:26, :4422, :11624, :16223)."""
    a = np.ascontiguousarray(a, np.float64).reshape(-1, 3)
    n = a.shape[0]
    A = np.ascontiguousarray(A, np.float64).reshape(n, 3)
    B = np.ascontiguousarray(B, np.float64).reshape(n, 3)
    C_ = np.ascontiguousarray(C_, np.float64).reshape(n, 3)
    sw = np.ascontiguousarray(sw, np.int32).reshape(n)
    out = np.zeros((n, 16, 16))
    lib().orc_prim_batch(op, n, _dp(a), _dp(A), _dp(B), _dp(C_), _ip(sw), _dp(out))
    return out.transpose(0, 2, 1)


def primk_batch(a, A, B, dK, sw):
    """k-overlap blocks (KOverlap2CIntg, gaussIntegrals.f90:129408): complex
    [n, 16, 16] ([i, row, col]) of exp(-i dK . r) between the Gaussians."""
    a = np.ascontiguousarray(a, np.float64).reshape(-1, 2)
    n = a.shape[0]
    A = np.ascontiguousarray(A, np.float64).reshape(n, 3)
    B = np.ascontiguousarray(B, np.float64).reshape(n, 3)
    dK = np.ascontiguousarray(dK, np.float64).reshape(n, 3)
    sw = np.ascontiguousarray(sw, np.int32).reshape(n)
    re = np.zeros((n, 16, 16))
    im = np.zeros((n, 16, 16))
    lib().orc_primk_batch(n, _dp(a), _dp(A), _dp(B), _dp(dK), _ip(sw), _dp(re), _dp(im))
    return (re + 1j * im).transpose(0, 2, 1)


def primd_batch(op, a, A, B, C_, sw):
    """Force-derivative blocks d/dB (computeForceIntg's primitives): op 10
    dkinetic2CIntg (gaussIntegrals.f90:65512), 11 dnuclear3CIntgCB (:88076),
    12 delectron3CIntgCB (:112289); [n, 3, 16, 16] ([i, component, row,
    col]).  A nuclear block of la + lb = 6 raises (it needs F_7)."""
    a = np.ascontiguousarray(a, np.float64).reshape(-1, 3)
    n = a.shape[0]
    A = np.ascontiguousarray(A, np.float64).reshape(n, 3)
    B = np.ascontiguousarray(B, np.float64).reshape(n, 3)
    C_ = np.ascontiguousarray(C_, np.float64).reshape(n, 3)
    sw = np.ascontiguousarray(sw, np.int32).reshape(n)
    out = np.zeros((n, 3, 16, 16))
    if lib().orc_primd_batch(op, n, _dp(a), _dp(A), _dp(B), _dp(C_), _ip(sw), _dp(out)):
        raise ValueError("nuclear derivative of an f-f block (needs F_7)")
    return out.transpose(0, 1, 3, 2)


def k_overlap_deltas(sysin, num_axial):
    """The three deltaK vectors of gaussKOverlap: recipVectors(:, axis) /
    numAxialKPoints(axis) (integrals3Terms.F90:1230-1236)."""
    b = recip_vectors(sysin.cell)
    return [b[x] / float(num_axial[x]) for x in range(3)]


# ---------------------------------------------------------------------------
# derived setup
ANG_NORM = np.array([
    math.sqrt(1 / (4 * math.pi)),
    *[math.sqrt(3 / (4 * math.pi))] * 3,
    *[math.sqrt(15 / (4 * math.pi))] * 3,
    math.sqrt(15 / (16 * math.pi)), math.sqrt(5 / (16 * math.pi)),
    math.sqrt(105 / (4 * math.pi)), math.sqrt(105 / (16 * math.pi)),
    math.sqrt(35 / (32 * math.pi)), math.sqrt(35 / (32 * math.pi)),
    math.sqrt(7 / (16 * math.pi)),
    math.sqrt(21 / (32 * math.pi)), math.sqrt(21 / (32 * math.pi))])


def renormalize_basis(sysin):
    """lmAlphaIndex, lmIndex and basis(alpha, state) per type
    (basis.f90:55-243): states run over valence radial functions in input
    order, each expanded over m (lmIndex = l^2 + m), then core; the
    coefficient of each state is radial coefficient x angNorm(lmIndex)."""
    maxA = max(len(t.alphas) for t in sysin.types)
    maxS = max(t.num_vale_states + t.num_core_states for t in sysin.types)
    nT = len(sysin.types)
    lmai = np.zeros((nT, maxA), np.int32)
    lmi = np.zeros((nT, maxS), np.int32)
    basis = np.zeros((nT, maxS, maxA))
    for i, t in enumerate(sysin.types):
        no = t.num_orb_alphas
        for j in range(len(t.alphas)):
            if j < no[3]:
                lmai[i, j] = 16
            elif j < no[2]:
                lmai[i, j] = 9
            elif j < no[1]:
                lmai[i, j] = 4
            else:
                lmai[i, j] = 1
        s = 0
        for l, coef in list(zip(t.vale_l, t.vale_coef)) + list(zip(t.core_l, t.core_coef)):
            na = no[l]
            for k in range(2 * l + 1):
                lmi[i, s] = l * l + k + 1
                basis[i, s, :na] = np.asarray(coef)[:na] * ANG_NORM[lmi[i, s] - 1]
                s += 1
    return lmai, lmi, basis, maxA, maxS


def recip_vectors(cell):
    """b_i = 2 pi (a_j x a_k)/V (lattice.f90:174-237); rows are b_i."""
    a = np.asarray(cell, float)
    vol = np.dot(a[0], np.cross(a[1], a[2]))
    b = np.array([np.cross(a[1], a[2]), np.cross(a[2], a[0]), np.cross(a[0], a[1])])
    return 2.0 * math.pi * b / vol


def kpoints_cartesian(sysin):
    """k_cart = sum_a f_a b_a (kpoints.f90:424-454)."""
    b = recip_vectors(sysin.cell)
    return np.ascontiguousarray(sysin.kpoints @ b)


def make_lattice(sysin):
    """Sorted real-space cell list within the negligibility sphere
    (lattice.f90:310-443, :469-781), computed by the C restatement."""
    tau = -math.log(sysin.basis_cutoff)
    vec = np.ascontiguousarray(sysin.cell, np.float64).reshape(9)
    min_a = min(float(np.min(t.alphas)) for t in sysin.types)
    min_p = min(float(np.min(p.alphas)) for p in sysin.pot_types)
    nl = lib().orc_neglig_limit(_dp(vec), tau, min_a, min_p)
    cap = 1 << 20
    dims = np.zeros((cap, 3))
    sizes = np.zeros(cap)
    n = lib().orc_make_lattice(_dp(vec), nl, cap, _dp(dims), _dp(sizes))
    if n < 0:
        raise RuntimeError("cell list too large")
    return dims[:n].copy(), sizes[:n].copy(), nl


def nearest_lattice_vector(cell, v):
    vec = np.ascontiguousarray(cell, np.float64).reshape(9)
    v = np.ascontiguousarray(v, np.float64)
    out = np.zeros(3)
    lib().orc_nearest_lattice_vector(_dp(vec), _dp(v), _dp(out))
    return out


class _Sys(C.Structure):
    _fields_ = [
        ("nAtoms", C.c_int), ("pos", C.c_void_p), ("atomType", C.c_void_p),
        ("nTypes", C.c_int), ("maxAlphas", C.c_int), ("maxStates", C.c_int),
        ("nAlpha", C.c_void_p), ("alpha", C.c_void_p), ("lmAlphaIndex", C.c_void_p),
        ("nVale", C.c_void_p), ("nCore", C.c_void_p), ("lmIndex", C.c_void_p),
        ("basis", C.c_void_p), ("element", C.c_void_p),
        ("nPotTypes", C.c_int), ("maxPotAlphas", C.c_int),
        ("potNAlpha", C.c_void_p), ("potAlpha", C.c_void_p), ("potZ", C.c_void_p),
        ("potNucAlpha", C.c_void_p), ("nPotSites", C.c_int), ("potPos", C.c_void_p),
        ("potSiteType", C.c_void_p), ("vec", C.c_double * 9), ("nCells", C.c_int),
        ("cellDims", C.c_void_p), ("cellSizes", C.c_void_p), ("nk", C.c_int),
        ("kpt", C.c_void_p), ("logThresh", C.c_double), ("gamma", C.c_int),
        ("rel", C.c_int)]


class Oracle:
    """Oracle view of one system: derived setup plus the integral sweeps."""

    def __init__(self, sysin):
        sysin.validate()
        self.sysin = sysin
        f = sysin.flat()
        self.lmai, self.lmi, self.basis, maxA, maxS = renormalize_basis(sysin)
        self.cellDims, self.cellSizes, self.negligLimit = make_lattice(sysin)
        self.kcart = kpoints_cartesian(sysin)
        nT = len(sysin.types)
        self._keep = dict(
            pos=f["atomPos"], atype=f["atomType"],
            nalpha=f["typeNumAlphas"], alpha=np.ascontiguousarray(f["typeAlphas"]),
            lmai=np.ascontiguousarray(self.lmai), lmi=np.ascontiguousarray(self.lmi),
            basis=np.ascontiguousarray(self.basis),
            nvale=np.array([t.num_vale_states for t in sysin.types], np.int32),
            ncore=np.array([t.num_core_states for t in sysin.types], np.int32),
            elem=f["typeElement"], pna=f["potNumAlphas"],
            palpha=np.ascontiguousarray(f["potAlphas"]), pZ=f["potZ"],
            pnuc=f["potNucAlpha"], ppos=f["potPos"], ptype=f["potSiteType"],
            cells=np.ascontiguousarray(self.cellDims), sizes=self.cellSizes,
            kpt=self.kcart)
        k = self._keep
        s = _Sys()
        s.nAtoms = sysin.n_atoms
        s.pos, s.atomType = k["pos"].ctypes.data, k["atype"].ctypes.data
        s.nTypes, s.maxAlphas, s.maxStates = nT, maxA, maxS
        s.nAlpha, s.alpha = k["nalpha"].ctypes.data, k["alpha"].ctypes.data
        s.lmAlphaIndex, s.nVale = k["lmai"].ctypes.data, k["nvale"].ctypes.data
        s.nCore, s.lmIndex = k["ncore"].ctypes.data, k["lmi"].ctypes.data
        s.basis, s.element = k["basis"].ctypes.data, k["elem"].ctypes.data
        s.nPotTypes, s.maxPotAlphas = len(sysin.pot_types), f["maxPotAlphas"]
        s.potNAlpha, s.potAlpha = k["pna"].ctypes.data, k["palpha"].ctypes.data
        s.potZ, s.potNucAlpha = k["pZ"].ctypes.data, k["pnuc"].ctypes.data
        s.nPotSites = f["nPotSites"]
        s.potPos, s.potSiteType = k["ppos"].ctypes.data, k["ptype"].ctypes.data
        for i, v in enumerate(np.asarray(sysin.cell, float).reshape(9)):
            s.vec[i] = v
        s.nCells = len(self.cellSizes)
        s.cellDims, s.cellSizes = k["cells"].ctypes.data, k["sizes"].ctypes.data
        s.nk = len(self.kcart)
        s.kpt = k["kpt"].ctypes.data
        s.logThresh = -math.log(sysin.basis_cutoff)
        s.gamma = int(sysin.gamma)
        s.rel = int(getattr(sysin, "rel", 0))
        self._s = s
        self.n = sysin.vale_dim
        self.nc = sysin.core_dim
        self.nk = sysin.nk

    def integral(self, op, term=0, coef=None, nthreads=0):
        """(valeVale, coreVale, coreCore) of one integral sweep, before any
        core orthogonalisation.  Complex builds return arrays [k, row, col];
        the Gamma build returns real [row, col] (integrals.F90:151-3605).
        HAM takes one spin's coefficients [potDim]."""
        n, nc, nk = self.n, self.nc, self.nk
        cplx = 1 if self.sysin.gamma else 2
        vv = np.zeros(cplx * n * n * nk)
        cv = np.zeros(max(cplx * nc * n * nk, 1))
        cc = np.zeros(max(cplx * nc * nc * nk, 1))
        cf = np.ascontiguousarray(coef if coef is not None else np.zeros(1), np.float64)
        err = lib().orc_integral(C.byref(self._s), op, term, _dp(cf), _dp(vv), _dp(cv),
                                 _dp(cc), nthreads)
        if err:
            raise RuntimeError(f"oracle integral failed with code {err}")

        def shape(a, r, c):
            if cplx == 2:
                a = a[:2 * r * c * nk].view(np.complex128).reshape(nk, c, r)
                return a.transpose(0, 2, 1).copy()
            return a[:r * c].reshape(c, r).T.copy()
        return shape(vv, n, n), shape(cv, nc, n) if nc else None, \
            shape(cc, nc, nc) if nc else None

    def k_overlap(self, num_axial, nthreads=0):
        """gaussKOverlap (integrals3Terms.F90:956-1384) before
        orthogonalisation: per axis (vv, cv, cc) complex [k, row, col], the
        sweep of the blocks' real parts plus the sweep of minus their
        imaginary parts under the momentum's -i Bloch factor.  Complex
        builds only (the reference's sweep is complex throughout)."""
        if self.sysin.gamma:
            raise ValueError("the k-overlap sweep is complex; the Gamma build has none")
        out = []
        for dk in k_overlap_deltas(self.sysin, num_axial):
            a = self.integral(KO, 0, coef=dk, nthreads=nthreads)
            b = self.integral(KO, 1, coef=dk, nthreads=nthreads)
            out.append(tuple(None if x is None else x + y for x, y in zip(a, b)))
        return out

    def force_integrals(self, total_energy, nthreads=0):
        """computeForceIntg (forces.F90:60-898): per spin and Cartesian
        component, (vv, cv, cc) of the ket-derivative matrices with that
        spin's potential coefficients and -2 totalEnergy times the momentum
        block; complex builds hold the full valeVale (the block and its
        dagger, a diagonal atom block as its own dagger), the Gamma build the
        upper triangle; no core orthogonalisation."""
        s = self.sysin
        if (np.asarray(self.lmai) == 16).any():
            raise ValueError("force integrals of f orbitals: the f-f nuclear derivative needs F_7")
        coefs = np.asarray(s.pot_coeffs, np.float64).reshape(s.spin, -1)
        out = []
        for q in range(s.spin):
            cf = np.concatenate([coefs[q], [float(total_energy)]])
            out.append([self.integral(FORCE, x, coef=cf, nthreads=nthreads) for x in range(3)])
        return out

    @staticmethod
    def set_pair_sample(stride, offset=0):
        """Sample every stride-th atom pair in later sweeps (1 = all)."""
        lib().orc_set_pair_sample(int(stride), int(offset))

    @staticmethod
    def set_three_center_range_scale(f):
        """Test-only: scale the three-centre image range (threeAlphaDist
        and its twins) by f in later sweeps; 1 is the reference's range."""
        lib().orc_set_three_center_range_scale(float(f))

    def last_counts(self):
        v = [C.c_longlong() for _ in range(3)]
        lib().orc_last_counts(*[C.byref(x) for x in v])
        return tuple(x.value for x in v)


# ---------------------------------------------------------------------------
# orthogonalisation, packing, assembly, solve
def ortho_overlap(S_vv, S_cv, S_cc):
    """S' = S_vv - 2 S_cv^H S_cv + S_cv^H S_cc S_cv on the upper triangle
    (intgOrtho.F90:23-46 zherk alpha = -2, :78-113, :115-211)."""
    A = np.conj(S_cv).swapaxes(-1, -2) @ S_cc @ S_cv
    out = S_vv - 2.0 * (np.conj(S_cv).swapaxes(-1, -2) @ S_cv) + A
    return np.triu(out) if out.ndim == 2 else np.stack([np.triu(o) for o in out])


def ortho_k_overlap(O_vv, O_cv, O_cc, S_cv):
    """ortho3Terms opcode 8/9 (integrals3Terms.F90:1388-1608): the full
    matrix O_vv - S_cv^H O_cv - O_cv^H S_cv + S_cv^H O_cc S_cv
    (valeCoreCoreVale fullFlag 1, coreValeCoreCore, makeValeVale fullFlag 1,
    intgOrtho.F90:48-76, :78-211), with no packing and no zeroing; O_vv as
    swept when there are no core states."""
    if O_cv is None:
        return O_vv
    H = lambda x: np.conj(x).swapaxes(-1, -2)
    return O_vv - H(S_cv) @ O_cv - H(O_cv) @ S_cv + H(S_cv) @ O_cc @ S_cv


def ortho_operator(O_vv, O_cv, O_cc, S_cv):
    """O' = O_vv - S_cv^H O_cv - O_cv^H S_cv + S_cv^H O_cc S_cv
    (intgOrtho.F90:48-76 zher2k, :78-211)."""
    H = lambda x: np.conj(x).swapaxes(-1, -2)
    out = O_vv - H(S_cv) @ O_cv - H(O_cv) @ S_cv + H(S_cv) @ O_cc @ S_cv
    return np.triu(out) if out.ndim == 2 else np.stack([np.triu(o) for o in out])


def ortho_dataset(vv, cv, cc, S_cv, overlap=False):
    """valeVale as the reference writes it after orthogonalisation against
    the core (orthoOL / ortho, integrals.F90:3608-3910): the projected upper
    triangle with components below 1e-10 zeroed (makeValeVale packing,
    intgOrtho.F90:142-190); unchanged when there is no core."""
    if cv is None or cv.shape[-2] == 0:
        return vv
    out = ortho_overlap(vv, cv, cc) if overlap else ortho_operator(vv, cv, cc, S_cv)
    if np.iscomplexobj(out):
        re, im = out.real.copy(), out.imag.copy()
        re[np.abs(re) < 1e-10] = 0.0
        im[np.abs(im) < 1e-10] = 0.0
        return re + 1j * im
    return np.where(np.abs(out) < 1e-10, 0.0, out)


def pack_upper(M, zero_small=False):
    """Column-packed upper triangle, index i(i-1)/2 + j for column i, row
    j <= i (matrixSubs.F90:506-534); with zero_small, components below 1e-10
    are zeroed as the orthogonalised packers do (intgOrtho.F90:142-190)."""
    n = M.shape[-1]
    cols, rows = [], []
    for i in range(n):
        cols.extend([i] * (i + 1))
        rows.extend(range(i + 1))
    p = M[..., np.array(rows), np.array(cols)]
    if zero_small:
        if np.iscomplexobj(p):
            re, im = p.real.copy(), p.imag.copy()
            re[np.abs(re) < 1e-10] = 0.0
            im[np.abs(im) < 1e-10] = 0.0
            p = re + 1j * im
        else:
            p = np.where(np.abs(p) < 1e-10, 0.0, p)
    return p


def unpack_upper(p, n):
    """Upper triangle into a zeroed matrix (unpackMatrix fullFlag=0,
    matrixSubs.F90:405-433)."""
    shape = p.shape[:-1] + (n, n)
    M = np.zeros(shape, dtype=p.dtype)
    idx = 0
    for i in range(n):
        M[..., :i + 1, i] = p[..., idx:idx + i + 1]
        idx += i + 1
    return M


def assemble_packed(NP, KE, V, coeffs, MV=None):
    """H = NP + KE (+ MV) + sum_j c_j V_j on packed storage
    (secularEqn.F90:147-170, readPackedMatrixAccum matrixSubs.F90:339-375:
    a multFactor of exactly 0.0 adds the term unscaled)."""
    H = NP + KE
    if MV is not None:
        H = H + MV
    for j, c in enumerate(coeffs):
        H = H + (V[j] if c == 0.0 else V[j] * c)
    return H


def solve(H, S, num_states=None):
    """Generalised eigenproblem H x = e S x from the upper triangles, LAPACK
    zhegv/dsygv(itype 1, 'V', 'U') as solveZHEGV/solveDSYGV
    (interfaces.F90:73-120, :152-200).  Returns eigenvalues (ascending) and
    eigenvectors; raises on info != 0 as the reference stops."""
    from scipy.linalg import lapack
    H = np.asarray(H)
    single = H.ndim == 2
    Hs = H[None] if single else H
    Ss = np.asarray(S)[None] if single else np.asarray(S)
    evals, evecs = [], []
    for h, s in zip(Hs, Ss):
        if np.iscomplexobj(h) or np.iscomplexobj(s):
            w, v, info = lapack.zhegv(np.asfortranarray(h, np.complex128),
                                      np.asfortranarray(s, np.complex128),
                                      itype=1, jobz="V", uplo="U")
        else:
            w, v, info = lapack.dsygv(np.asfortranarray(h), np.asfortranarray(s),
                                      itype=1, jobz="V", uplo="U")
        if info != 0:
            raise RuntimeError(f"zhegv error code={info}")
        if num_states is not None:
            w, v = w[:num_states], v[:, :num_states]
        evals.append(w)
        evecs.append(v)
    if single:
        return evals[0], evecs[0]
    return np.array(evals), np.array(evecs)


def matrix_element_mult(m1, m2):
    """Tr[M rho] of two packed Hermitian matrices as matrixElementMult sums
    them (matrixSubs.F90:66-112): off-diagonal pairs re*re + im*im, diagonal
    real parts halved, the whole doubled.  Real packed matrices follow
    matrixElementMultGamma (:143-179): twice the sum of all products minus
    the diagonal products."""
    n = int(round((math.sqrt(8 * m1.shape[-1] + 1) - 1) / 2))
    diag = np.cumsum(np.arange(1, n + 1)) - 1
    if np.iscomplexobj(m1) or np.iscomplexobj(m2):
        prod = m1.real * m2.real + m1.imag * m2.imag
        off = prod.sum() - prod[diag].sum()
        return 2.0 * (off + (m1.real[diag] * m2.real[diag]).sum() / 2.0)
    return 2.0 * float(np.sum(m1 * m2)) - float(np.sum(m1[diag] * m2[diag]))


def vale_charge(evals, evecs, occupancy, S, NP, KE, V, MV=None):
    """makeValenceRho with inSCF = 1 (valeCharge.F90:34-526) on the CPU.

    evals: [spin][nk][numStates]; evecs: [spin][nk][n][numStates] (columns
    are states; real at Gamma); occupancy: [nk][spin][numStates]
    (electronPopulation order, :160-169).  S, NP, KE, MV, V[j]: the term
    matrices as [nk][n][n] upper triangles (or [n][n] at Gamma).

    Per k: states whose populations sum below smallThresh in magnitude are
    skipped (:214-216); electronEnergy += f * eigenvalue (:218-219); the
    upper triangle of rho gets f x x^H (zher / dsyr, :228-232, :273-277); rho
    is packed (:263-265) and, at spin 2, rewritten as up + down and up - down
    (:306-314); then each trace accumulates matrixElementMult of the packed
    term matrix against the packed density (:330-418).  Returns
    (electronEnergy [spin], traces [spin][4] = charge, nuclear, kinetic,
    mass velocity, potRho [spin][potDim])."""
    occ = np.asarray(occupancy, np.float64)
    nk, spin, ns = occ.shape
    gamma = not np.iscomplexobj(np.asarray(evecs))
    mats = lambda M: (np.asarray(M)[None] if gamma else np.asarray(M))   # noqa: E731
    S, NP, KE = mats(S), mats(NP), mats(KE)
    MV = mats(MV) if MV is not None else None
    V = [mats(v) for v in V]
    n = S.shape[-1]
    ee = np.zeros(spin)
    traces = np.zeros((spin, 4))
    pot_rho = np.zeros((spin, len(V)))
    for k in range(nk):
        rho = np.zeros((spin, n, n), dtype=np.float64 if gamma else np.complex128)
        for j in range(ns):
            f = occ[k, :, j]
            if np.sum(np.abs(f)) < SMALL_THRESH:
                continue
            ee += f * np.array([evals[q][k][j] for q in range(spin)])
            for q in range(spin):
                x = np.asarray(evecs[q][k])[:, j]
                rho[q] += np.triu(f[q] * np.outer(x, np.conj(x)))
        packed = np.array([pack_upper(rho[q]) for q in range(spin)])
        if spin == 2:
            packed = np.array([packed[0] + packed[1], packed[0] - packed[1]])
        for t in range(spin):
            traces[t, 0] += matrix_element_mult(pack_upper(S[k]), packed[t])
            traces[t, 1] += matrix_element_mult(pack_upper(NP[k]), packed[t])
            traces[t, 2] += matrix_element_mult(pack_upper(KE[k]), packed[t])
            if MV is not None:
                traces[t, 3] += matrix_element_mult(pack_upper(MV[k]), packed[t])
            for jj, v in enumerate(V):
                pot_rho[t, jj] += matrix_element_mult(pack_upper(v[k]), packed[t])
    return ee, traces, pot_rho
