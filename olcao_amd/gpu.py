"""ctypes binding of libolcao_gpu.so with the reference's procedure names.

Each method of `OlcaoGPU` stands in for one Fortran procedure that
src/uolcao/OLCAO.F90 calls on the integral -> eigenvalue path:

    gaussOverlapOL        integrals.F90:151    (overlap S(k))
    gaussOverlapKE        integrals.F90:656    (kinetic)
    gaussOverlapNP        integrals.F90:1531   (nuclear attraction)
    elecPotGaussOverlap   integrals.F90:2357   (one matrix per potential term)
    gaussOverlapHamPSCF   integrals.F90:2860   (fused Hamiltonian)
    secularEqnPSCF        secularEqn.F90:331   (solve all k)
    setupSCF / secularEqnSCF  OLCAO.F90:248-279 / secularEqn.F90:37
    solveZHEGV / solveDSYGV   interfaces.F90:73 / :152

Matrices come back as numpy arrays [k, row, col] (complex) or [row, col]
(real, Gamma build) holding the upper triangle, as the reference's valeVale
does.  Errors raise `OlcaoError` carrying the library's message, matching
the reference's `stop`.  There is no CPU fallback: if the shared library or
a GPU is missing the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libolcao_gpu.so")

STATUS = {0: "OK", 1: "HIP", 2: "ALLOC", 3: "INPUT", 4: "CELLS", 5: "NOT_SPD",
          6: "SOLVER", 7: "UNSUPPORTED", 8: "STATE", 9: "NO_DEVICE"}
PREC_FP64, PREC_MIXED = 0, 1
BATCH_CELLS, BATCH_SHARED = 0, 1


class OlcaoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"olcao status {code} ({STATUS.get(code, '?')}): {msg}")
        self.code = code


class _System(C.Structure):
    _fields_ = [
        ("nAtoms", C.c_int), ("atomPos", C.c_void_p), ("atomType", C.c_void_p),
        ("nTypes", C.c_int), ("maxAlphas", C.c_int), ("maxRadial", C.c_int),
        ("typeNumAlphas", C.c_void_p), ("typeAlphas", C.c_void_p),
        ("typeNumOrbAlphas", C.c_void_p), ("typeElement", C.c_void_p),
        ("typeNumValeRadial", C.c_void_p), ("typeNumCoreRadial", C.c_void_p),
        ("typeValeL", C.c_void_p), ("typeCoreL", C.c_void_p),
        ("typeValeCoef", C.c_void_p), ("typeCoreCoef", C.c_void_p),
        ("nPotTypes", C.c_int), ("maxPotAlphas", C.c_int),
        ("potNumAlphas", C.c_void_p), ("potAlphas", C.c_void_p), ("potZ", C.c_void_p),
        ("potNucAlpha", C.c_void_p), ("potElement", C.c_void_p),
        ("nPotSites", C.c_int), ("potPos", C.c_void_p), ("potSiteType", C.c_void_p),
        ("cellVectors", C.c_double * 9), ("nk", C.c_int), ("kpoints", C.c_void_p),
        ("basisCutoff", C.c_double), ("gamma", C.c_int), ("rel", C.c_int), ("spin", C.c_int)]


class _Options(C.Structure):
    _fields_ = [("precision", C.c_int), ("eigensolver", C.c_int), ("bloch", C.c_int),
                ("graph", C.c_int), ("chunkItems", C.c_int), ("segPairs", C.c_int),
                ("reserved", C.c_int * 2)]


EIG_AUTO, EIG_LIBRARY, EIG_BLOCKED = 0, 1, 2
BLOCH_AUTO, BLOCH_VALU, BLOCH_MFMA, BLOCH_MFMA64 = 0, 1, 2, 3


_lib = None
FUNCTIONS = {
    "olcao_ctx_create": [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)],
    "olcao_ctx_destroy": [C.c_void_p],
    "olcao_last_error": [C.c_void_p],
    "olcao_get_dims": [C.c_void_p, C.c_void_p],
    "olcao_get_flags": [C.c_void_p, C.c_void_p],
    "olcao_get_eigen": [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p],
    "olcao_get_setup": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "olcao_gauss_overlap": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p],
    "olcao_gauss_overlap_np": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "olcao_elec_pot_gauss_overlap": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p],
    "olcao_gauss_overlap_dm": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "olcao_gauss_overlap_mm": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "olcao_gauss_k_overlap": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "olcao_force_integrals": [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                              C.c_void_p],
    "olcao_gauss_overlap_ham_pscf": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "olcao_secular_eqn_pscf": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p],
    "olcao_setup_scf": [C.c_void_p],
    "olcao_get_scf_term": [C.c_void_p, C.c_int, C.c_int, C.c_void_p],
    "olcao_secular_eqn_scf": [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p],
    "olcao_scf_iteration": [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p],
    "olcao_vale_charge": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "olcao_scf_iteration_batch": [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_int, C.c_void_p, C.c_void_p],
    "olcao_scf_iteration_batch_mode": [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_int, C.c_int, C.c_void_p, C.c_void_p],
    "olcao_get_timings": [C.c_void_p, C.c_void_p],
    "olcao_get_graph_state": [C.c_void_p, C.c_void_p, C.c_void_p],
    "olcao_get_class_pairs": [C.c_void_p, C.c_void_p],
    "olcao_get_work_counts": [C.c_void_p, C.c_void_p],
    "olcao_solve_zhegv": [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p],
    "olcao_solve_dsygv": [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p],
    "olcao_solve_zhegv_noalloc": [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p],
    "olcao_prim_batch": [C.c_int, C.c_int] + [C.c_void_p] * 6,
    "olcao_boys_batch": [C.c_int, C.c_void_p, C.c_void_p],
}


def lib():
    """Load the in-tree shared library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -m olcao_amd.build` (no CPU fallback exists)")
        _lib = C.CDLL(LIB_PATH)
        for name, args in FUNCTIONS.items():
            f = getattr(_lib, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib.olcao_ctx_destroy.restype = None
        _lib.olcao_last_error.restype = C.c_char_p
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _check(code, ctx=None):
    if code != 0:
        msg = lib().olcao_last_error(ctx).decode()
        raise OlcaoError(code, msg)


def solveZHEGV(A, B):
    """Batched drop-in for solveZHEGV (interfaces.F90:73-120).  A, B:
    [nb, n, n] complex (upper triangles used).  Returns (w [nb, n],
    eigenvectors [nb, n, n] with eigenvector s in [:, :, s])."""
    A = np.asarray(A, np.complex128)
    single = A.ndim == 2
    A = A[None] if single else A
    B = np.asarray(B, np.complex128)
    B = B[None] if single else B
    nb, n = A.shape[0], A.shape[1]
    Af = np.ascontiguousarray(A.transpose(0, 2, 1))   # column-major per matrix
    Bf = np.ascontiguousarray(B.transpose(0, 2, 1))
    w = np.zeros((nb, n))
    _check(lib().olcao_solve_zhegv(n, nb, _p(Af), _p(Bf), _p(w)))
    V = Af.transpose(0, 2, 1)
    return (w[0], V[0]) if single else (w, V)


def solveDSYGV(A, B):
    """Batched drop-in for solveDSYGV (interfaces.F90:152-200)."""
    A = np.asarray(A, np.float64)
    single = A.ndim == 2
    A = A[None] if single else A
    B = np.asarray(B, np.float64)
    B = B[None] if single else B
    nb, n = A.shape[0], A.shape[1]
    Af = np.ascontiguousarray(A.transpose(0, 2, 1))
    Bf = np.ascontiguousarray(B.transpose(0, 2, 1))
    w = np.zeros((nb, n))
    _check(lib().olcao_solve_dsygv(n, nb, _p(Af), _p(Bf), _p(w)))
    V = Af.transpose(0, 2, 1)
    return (w[0], V[0]) if single else (w, V)


def prim_batch(op, a, A, B, C_, sw):
    """Primitive blocks on the device, [n, 16, 16] with [i, row, col]."""
    a = np.ascontiguousarray(a, np.float64).reshape(-1, 3)
    n = a.shape[0]
    arrs = [np.ascontiguousarray(x, np.float64).reshape(n, 3) for x in (A, B, C_)]
    sw = np.ascontiguousarray(sw, np.int32).reshape(n)
    out = np.zeros((n, 16, 16))
    _check(lib().olcao_prim_batch(op, n, _p(a), _p(arrs[0]), _p(arrs[1]), _p(arrs[2]),
                                  _p(sw), _p(out)))
    return out.transpose(0, 2, 1)


def boys_batch(T):
    T = np.ascontiguousarray(T, np.float64).reshape(-1)
    F = np.zeros((len(T), 7))
    _check(lib().olcao_boys_batch(len(T), _p(T), _p(F)))
    return F


class OlcaoGPU:
    """One device context for one system (setupSCF's derived state)."""

    def __init__(self, sysin, precision=PREC_FP64, eigensolver=EIG_AUTO, bloch=BLOCH_AUTO,
                 graph=False, chunk_items=0, seg_pairs=0, tuning=None):
        """tuning: a dict of autotuned tile shapes (chunk_items, seg_pairs,
        bloch), or "table" for the entry of olcao_amd/tuning.json nearest to
        this system (olcao_amd.autotune.lookup); it overrides chunk_items,
        seg_pairs and bloch."""
        if tuning is not None:
            from olcao_amd import autotune
            t = autotune.lookup(sysin, precision) if tuning == "table" else tuning
            if t:
                chunk_items = t.get("chunk_items", chunk_items)
                seg_pairs = t.get("seg_pairs", seg_pairs)
                bloch = t.get("bloch", bloch)
        self.tuning = dict(chunk_items=chunk_items, seg_pairs=seg_pairs, bloch=bloch)
        self.sysin = sysin
        f = sysin.flat()
        self._keep = f
        s = _System()
        s.nAtoms = f["nAtoms"]
        s.atomPos, s.atomType = _p(f["atomPos"]), _p(f["atomType"])
        s.nTypes, s.maxAlphas, s.maxRadial = f["nTypes"], f["maxAlphas"], f["maxRadial"]
        for k in ("typeNumAlphas", "typeAlphas", "typeNumOrbAlphas", "typeElement",
                  "typeNumValeRadial", "typeNumCoreRadial", "typeValeL", "typeCoreL",
                  "typeValeCoef", "typeCoreCoef", "potNumAlphas", "potAlphas", "potZ",
                  "potNucAlpha", "potElement", "potPos", "potSiteType", "kpoints"):
            f[k] = np.ascontiguousarray(f[k])
            setattr(s, k, _p(f[k]))
        s.nPotTypes, s.maxPotAlphas = f["nPotTypes"], f["maxPotAlphas"]
        s.nPotSites = f["nPotSites"]
        for i, v in enumerate(f["cellVectors"]):
            s.cellVectors[i] = v
        s.nk, s.basisCutoff, s.gamma = f["nk"], f["basisCutoff"], f["gamma"]
        s.rel, s.spin = f["rel"], f["spin"]
        self._sys = s
        opt = _Options()
        opt.precision = precision
        opt.eigensolver = eigensolver
        opt.bloch = bloch
        opt.graph = 1 if graph else 0
        opt.chunkItems = int(chunk_items)
        opt.segPairs = int(seg_pairs)
        ctx = C.c_void_p()
        code = lib().olcao_ctx_create(C.byref(s), C.byref(opt), C.byref(ctx))
        _check(code, None)
        self.ctx = ctx
        dims = np.zeros(8, np.int64)
        _check(lib().olcao_get_dims(self.ctx, _p(dims)), self.ctx)
        (self.vale_dim, self.core_dim, self.nk, self.pot_dim, self.n_cells,
         self.n_pair_cells, self.n_prim_pairs, self.n_pairs) = [int(x) for x in dims]
        self.gamma = bool(sysin.gamma)
        self.spin = int(sysin.spin)

    def close(self):
        if getattr(self, "ctx", None):
            lib().olcao_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- setup inspection --------------------------------------------------
    def setup(self):
        cells = np.zeros((self.n_cells, 3))
        sizes = np.zeros(self.n_cells)
        k = np.zeros((self.nk, 3))
        _check(lib().olcao_get_setup(self.ctx, _p(cells), _p(sizes), _p(k)), self.ctx)
        return cells, sizes, k

    # ---- output buffers ----------------------------------------------------
    def _bufs(self, nspin=1):
        n, nc, nk = self.vale_dim, self.core_dim, self.nk
        cp = 1 if self.gamma else 2
        vv = np.zeros(nspin * cp * n * n * nk)
        cv = np.zeros(max(nspin * cp * nc * n * nk, 1))
        cc = np.zeros(max(nspin * cp * nc * nc * nk, 1))
        return vv, cv, cc

    def _shape(self, vv, cv, cc, nspin=1):
        n, nc, nk = self.vale_dim, self.core_dim, self.nk
        cp = 1 if self.gamma else 2

        def sh(a, r, c, q):
            a = a[q * cp * r * c * nk:(q + 1) * cp * r * c * nk]
            if not self.gamma:
                a = a.view(np.complex128).reshape(nk, c, r)
                return a.transpose(0, 2, 1).copy()
            return a.reshape(c, r).T.copy()
        out = [(sh(vv, n, n, q), (sh(cv, nc, n, q) if nc else None),
                (sh(cc, nc, nc, q) if nc else None)) for q in range(nspin)]
        return out[0] if nspin == 1 else tuple(out)

    def _coeffs(self, pot_coeffs):
        c = np.ascontiguousarray(pot_coeffs, np.float64).reshape(-1)
        if c.size != self.spin * self.pot_dim:
            raise ValueError(f"pot_coeffs must hold spin x potDim = {self.spin} x "
                             f"{self.pot_dim} values")
        return c

    # ---- integral procedures -----------------------------------------------
    def gaussOverlapOL(self):
        """valeVale (orthogonalised when coreDim > 0), raw coreVale, coreCore."""
        vv, cv, cc = self._bufs()
        _check(lib().olcao_gauss_overlap(self.ctx, 0, _p(vv), _p(cv), _p(cc)), self.ctx)
        return self._shape(vv, cv, cc)

    def gaussOverlapKE(self):
        vv, cv, cc = self._bufs()
        _check(lib().olcao_gauss_overlap(self.ctx, 1, _p(vv), _p(cv), _p(cc)), self.ctx)
        return self._shape(vv, cv, cc)

    def gaussOverlapMV(self):
        """Negated mass-velocity matrices (rel = 1), as gaussOverlapMV stores."""
        vv, cv, cc = self._bufs()
        _check(lib().olcao_gauss_overlap(self.ctx, 2, _p(vv), _p(cv), _p(cc)), self.ctx)
        return self._shape(vv, cv, cc)

    def gaussOverlapNP(self):
        vv, cv, cc = self._bufs()
        _check(lib().olcao_gauss_overlap_np(self.ctx, _p(vv), _p(cv), _p(cc)), self.ctx)
        return self._shape(vv, cv, cc)

    def elecPotGaussOverlap(self, term):
        vv, cv, cc = self._bufs()
        _check(lib().olcao_elec_pot_gauss_overlap(self.ctx, int(term), _p(vv), _p(cv), _p(cc)),
               self.ctx)
        return self._shape(vv, cv, cc)

    def gaussOverlapDM(self, center):
        """Dipole-moment matrices about `center`: three (vv, cv, cc) tuples,
        x, y, z (gaussOverlapDM, integrals3Terms.F90:64)."""
        vv, cv, cc = self._bufs(3)
        ctr = np.ascontiguousarray(center, np.float64).reshape(3)
        _check(lib().olcao_gauss_overlap_dm(self.ctx, _p(ctr), _p(vv), _p(cv), _p(cc)), self.ctx)
        return self._shape(vv, cv, cc, 3)

    def gaussOverlapMM(self):
        """Momentum matrices (-i nabla): three (vv, cv, cc) tuples
        (gaussOverlapMM, integrals3Terms.F90:511)."""
        vv, cv, cc = self._bufs(3)
        _check(lib().olcao_gauss_overlap_mm(self.ctx, _p(vv), _p(cv), _p(cc)), self.ctx)
        return self._shape(vv, cv, cc, 3)

    def gaussKOverlap(self, num_axial):
        """k-overlap matrices <phi_i(k)| exp(-i dK . r) |phi_j(k)>, dK =
        b_axis / num_axial[axis]: three (vv, cv, cc) tuples, axis-major
        (gaussKOverlap, integrals3Terms.F90:956; full matrices, orthogonalised
        without zeroing as ortho3Terms opcode 8 does).  Complex contexts
        only."""
        vv, cv, cc = self._bufs(3)
        na = np.ascontiguousarray(num_axial, np.int32).reshape(3)
        _check(lib().olcao_gauss_k_overlap(self.ctx, _p(na), _p(vv), _p(cv), _p(cc)), self.ctx)
        return self._shape(vv, cv, cc, 3)

    def computeForceIntg(self, pot_coeffs, total_energy):
        """computeForceIntg (forces.F90:60-898): per spin a list of three
        (vv, cv, cc) tuples, one per Cartesian component of the ket
        derivative (full matrices in a complex build, upper triangle at
        Gamma; not orthogonalised).  Bases with f orbitals are refused."""
        c = self._coeffs(pot_coeffs)
        vv, cv, cc = self._bufs(3 * self.spin)
        _check(lib().olcao_force_integrals(self.ctx, _p(c), float(total_energy), _p(vv), _p(cv),
                                           _p(cc)), self.ctx)
        flat = self._shape(vv, cv, cc, 3 * self.spin)
        return [flat[3 * q:3 * q + 3] for q in range(self.spin)]

    def gaussOverlapHamPSCF(self, pot_coeffs):
        """One (vv, cv, cc) per spin; a single tuple when spin = 1."""
        c = self._coeffs(pot_coeffs)
        vv, cv, cc = self._bufs(self.spin)
        _check(lib().olcao_gauss_overlap_ham_pscf(self.ctx, _p(c), _p(vv), _p(cv), _p(cc)),
               self.ctx)
        return self._shape(vv, cv, cc, self.spin)

    # ---- solve ---------------------------------------------------------------
    def _evecs_shape(self, ev, ns):
        n, nk = self.vale_dim, self.nk
        out = []
        per = (1 if self.gamma else 2) * n * ns * nk
        for q in range(self.spin):
            e = ev[q * per:(q + 1) * per]
            if self.gamma:
                out.append(e.reshape(ns, n).T.copy())
            else:
                out.append(e.view(np.complex128).reshape(nk, ns, n).transpose(0, 2, 1).copy())
        return out[0] if self.spin == 1 else np.array(out)

    def _evals(self, w):
        return w[0] if self.spin == 1 else w

    def secularEqnPSCF(self, num_states, vectors=False):
        """Eigenvalues [nk, numStates] ([spin, nk, numStates] when spin = 2)."""
        n, nk = self.vale_dim, self.nk
        w = np.zeros((self.spin, nk, num_states))
        ev = (np.zeros(self.spin * (1 if self.gamma else 2) * n * num_states * nk)
              if vectors else None)
        _check(lib().olcao_secular_eqn_pscf(self.ctx, num_states, _p(w), _p(ev)), self.ctx)
        return (self._evals(w), self._evecs_shape(ev, num_states)) if vectors else self._evals(w)

    def setupSCF(self):
        _check(lib().olcao_setup_scf(self.ctx), self.ctx)

    def secularEqnSCF(self, pot_coeffs, num_states, vectors=False):
        c = self._coeffs(pot_coeffs)
        n, nk = self.vale_dim, self.nk
        w = np.zeros((self.spin, nk, num_states))
        ev = (np.zeros(self.spin * (1 if self.gamma else 2) * n * num_states * nk)
              if vectors else None)
        _check(lib().olcao_secular_eqn_scf(self.ctx, _p(c), num_states, _p(w), _p(ev)),
               self.ctx)
        return (self._evals(w), self._evecs_shape(ev, num_states)) if vectors else self._evals(w)

    def scf_iteration(self, pot_coeffs, num_states):
        """Mode B iteration: overlap sweep, fused Hamiltonian sweep, solve."""
        c = self._coeffs(pot_coeffs)
        w = np.zeros((self.spin, self.nk, num_states))
        _check(lib().olcao_scf_iteration(self.ctx, _p(c), num_states, _p(w)), self.ctx)
        return self._evals(w)

    def makeValenceRho(self, occupancy):
        """Valence density traces from the last solve (makeValenceRho,
        valeCharge.F90:34).  occupancy: [nk, spin, numStates] populations.
        Returns dict(electron_energy [spin], charge, nuclear, kinetic,
        mass_velocity [spin] (0 = total, 1 = difference), pot_rho [spin,
        potDim])."""
        occ = np.ascontiguousarray(occupancy, np.float64).reshape(self.nk, self.spin, -1)
        ns = occ.shape[2]
        ee = np.zeros(self.spin)
        tr = np.zeros((self.spin, 4))
        pr = np.zeros((self.spin, self.pot_dim))
        _check(lib().olcao_vale_charge(self.ctx, ns, _p(occ), _p(ee), _p(tr), _p(pr)), self.ctx)
        return dict(electron_energy=ee, charge=tr[:, 0], nuclear=tr[:, 1], kinetic=tr[:, 2],
                    mass_velocity=tr[:, 3], pot_rho=pr)

    def work_counts(self):
        """Three-centre evaluations of the last sweep per class (la, lb):
        array [4, 4, 2] of (three-Gaussian overlap, nuclear)."""
        c = np.zeros(32, np.uint64)
        _check(lib().olcao_get_work_counts(self.ctx, _p(c)), self.ctx)
        return c.reshape(4, 4, 2)

    def timings(self):
        t = np.zeros(8)
        _check(lib().olcao_get_timings(self.ctx, _p(t)), self.ctx)
        return dict(overlap_ms=t[0], hamiltonian_ms=t[1], contract_bloch_ms=t[2],
                    solve_ms=t[3], total_ms=t[4], three_center_ms=t[5], graph_replay=t[6] > 0,
                    trd_ms=t[7])

    def class_pairs(self):
        """Screened primitive pairs per class, [4, 4] (la, lb)."""
        v = np.zeros(16, np.int64)
        _check(lib().olcao_get_class_pairs(self.ctx, _p(v)), self.ctx)
        return v.reshape(4, 4)

    def graph_state(self):
        """(state, reason): 0 graph mode off, 1 a captured graph is current,
        2 waiting for the next iteration to capture, 3 capture refused
        (reason says why)."""
        st = C.c_int(0)
        why = C.c_char_p()
        _check(lib().olcao_get_graph_state(self.ctx, C.byref(st), C.byref(why)), self.ctx)
        return st.value, (why.value or b"").decode()


def scf_iteration_batch(cells, num_states, workers=8, mode=BATCH_CELLS):
    """Batched Mode-B iteration over independent cells (OlcaoGPU contexts,
    one per cell) in one call: olcao_scf_iteration_batch_mode.  num_states:
    one count per cell; mode: BATCH_CELLS (each cell's build on its own
    streams) or BATCH_SHARED (the cells' three-centre work in shared
    launches).  Returns (eigenvalues per cell as scf_iteration gives them,
    each cell's latency in ms within the batch)."""
    n = len(cells)
    coefs = [c._coeffs(c.sysin.pot_coeffs) for c in cells]
    ws = [np.zeros((c.spin, c.nk, ns)) for c, ns in zip(cells, num_states)]
    ctxs = (C.c_void_p * n)(*[c.ctx for c in cells])
    cp = (C.c_void_p * n)(*[_p(a) for a in coefs])
    wp = (C.c_void_p * n)(*[_p(a) for a in ws])
    ns = np.asarray(num_states, np.int32)
    ms = np.zeros(n)
    bad = np.full(1, -1, np.int32)
    code = lib().olcao_scf_iteration_batch_mode(n, C.cast(ctxs, C.c_void_p),
                                                C.cast(cp, C.c_void_p), _p(ns),
                                                C.cast(wp, C.c_void_p), workers, mode, _p(ms),
                                                _p(bad))
    if code != 0:
        msg = lib().olcao_last_error(None).decode()
        raise OlcaoError(code, f"cell {int(bad[0])}: {msg}")
    return [c._evals(w) for c, w in zip(cells, ws)], ms
