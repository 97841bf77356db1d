"""olcao_solve_zhegv_noalloc (no device allocation; a fixed per-GPU staging
array) against olcao_solve_zhegv on the same batches, through the C-ABI:
two calls in a row, the second with a smaller input, each equal to the
allocating entry point's eigenvalues and eigenvectors (the same LDS solver
kernel on the same inputs: bit for bit); sizes beyond the array or the
solver are refused with a status."""
import ctypes as C

import numpy as np
import pytest

HARTREE_EV = 27.211386245988


def _pencil(rng, n, nb):
    """Hermitian A and Hermitian positive definite B, column-major per matrix."""
    Z = rng.normal(size=(nb, n, n)) + 1j * rng.normal(size=(nb, n, n))
    A = (Z + np.conj(Z.transpose(0, 2, 1))) / 2
    Y = rng.normal(size=(nb, n, n)) + 1j * rng.normal(size=(nb, n, n))
    B = Y @ np.conj(Y.transpose(0, 2, 1)) / n + np.eye(n)
    return (np.ascontiguousarray(A.transpose(0, 2, 1)), np.ascontiguousarray(B.transpose(0, 2, 1)))


def _solve(fn, n, nb, A, B):
    A, B = A.copy(), B.copy()
    w = np.zeros((nb, n))
    p = lambda a: a.ctypes.data_as(C.c_void_p)   # noqa: E731
    st = fn(n, nb, p(A), p(B), p(w))
    return st, w, A


@pytest.mark.gpu
def test_noalloc_solver_matches_allocating_solver():
    from olcao_amd import gpu
    lib = gpu.lib()
    rng = np.random.default_rng(11)
    for n, nb in ((57, 40), (33, 7)):          # the second call smaller than the first
        print(f"noalloc zhegv: n={n} nb={nb}", flush=True)
        A, B = _pencil(rng, n, nb)
        st0, w0, V0 = _solve(lib.olcao_solve_zhegv, n, nb, A, B)
        st1, w1, V1 = _solve(lib.olcao_solve_zhegv_noalloc, n, nb, A, B)
        assert st0 == 0 and st1 == 0, (st0, st1, lib.olcao_last_error(None))
        assert np.array_equal(w0, w1)
        assert np.array_equal(V0, V1)
        # and against LAPACK on the first matrix
        from scipy.linalg import eigh
        a = A[0].view(np.complex128).reshape(n, n).T
        b = B[0].view(np.complex128).reshape(n, n).T
        ref = eigh(np.triu(a) + np.conj(np.triu(a, 1)).T, np.triu(b) + np.conj(np.triu(b, 1)).T,
                   eigvals_only=True)
        assert np.abs(w1[0] - ref).max() < 1e-10 * np.abs(ref).max()


@pytest.mark.gpu
def test_noalloc_solver_refuses_sizes_beyond_its_limits():
    from olcao_amd import gpu
    lib = gpu.lib()
    A = np.zeros(2 * 71 * 71)
    B = np.zeros(2 * 71 * 71)
    w = np.zeros(71)
    p = lambda a: a.ctypes.data_as(C.c_void_p)   # noqa: E731
    assert lib.olcao_solve_zhegv_noalloc(71, 1, p(A), p(B), p(w)) == 7       # UNSUPPORTED
    assert b"n <= 70" in lib.olcao_last_error(None)
    assert lib.olcao_solve_zhegv_noalloc(57, 10 ** 6, p(A), p(B), p(w)) == 3  # INPUT
    assert b"MiB scratch" in lib.olcao_last_error(None)
