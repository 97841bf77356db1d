"""The C-ABI library loads without a GPU and exports every entry point that
include/olcao_gpu.h declares; host-side validation paths answer with status
codes (no compute runs here)."""
import ctypes as C
import os
import re

import pytest

from olcao_amd import configs, gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "olcao_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(olcao_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = gpu.lib()
    names = declared_functions()
    assert len(names) >= 18
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(gpu.FUNCTIONS)


def test_context_creation_reports_missing_gpu_or_succeeds():
    try:
        g = gpu.OlcaoGPU(configs.small_test())
    except gpu.OlcaoError as e:
        assert e.code == 9          # OLCAO_ERR_NO_DEVICE on a host without GPU
        return
    g.close()


def test_null_arguments_are_rejected():
    lib = gpu.lib()
    assert lib.olcao_ctx_create(None, None, None) == 3
    assert lib.olcao_get_dims(None, None) == 3
    assert lib.olcao_scf_iteration(None, None, 1, None) == 3
    bad = C.c_int(7)
    assert lib.olcao_scf_iteration_batch(0, None, None, None, None, 4, None, C.byref(bad)) == 3
    assert bad.value == -1
    assert lib.olcao_scf_iteration_batch_mode(0, None, None, None, None, 4, 1, None,
                                              C.byref(bad)) == 3


def test_fortran_facade_source_names_reference_procedures():
    src = open(os.path.join(ROOT, "olcao_amd", "fortran", "olcao_gpu_facade.f90")).read()
    for proc in ("gaussOverlapOL", "gaussOverlapKE", "gaussOverlapNP",
                 "elecPotGaussOverlap", "gaussOverlapHamPSCF", "secularEqnPSCF",
                 "secularEqnSCF", "solveZHEGV", "solveDSYGV", "makeValenceRho"):
        assert re.search(r"subroutine\s+" + proc + r"\b", src, re.I), proc


def test_hdf5_writer_library_exports_its_symbol():
    """libolcao_h5.so (built where the HDF5 C library is installed) loads and
    exports what include/olcao_h5.h declares."""
    path = os.path.join(ROOT, "olcao_amd", "libolcao_h5.so")
    if not os.path.exists("/opt/conda/include/hdf5.h"):
        pytest.skip("no HDF5 C library on this host")
    lib = C.CDLL(path)
    src = open(os.path.join(ROOT, "include", "olcao_h5.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set(re.findall(r"\b(olcao_[a-z_0-9]+)\s*\(", src))
    assert names == {"olcao_write_scf_integrals_h5", "olcao_write_scf_eigen_h5",
                     "olcao_write_scf_vector_terms_h5", "olcao_write_pscf_h5",
                     "olcao_write_k_overlap_h5"}
    for name in names:
        assert hasattr(lib, name), name
    lib.olcao_write_scf_integrals_h5.argtypes = [C.c_void_p, C.c_char_p]
    assert lib.olcao_write_scf_integrals_h5(None, b"/nonexistent") == 3
    lib.olcao_write_k_overlap_h5.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p, C.c_char_p]
    assert lib.olcao_write_k_overlap_h5(None, None, None, b"/nonexistent") == 3
    lib.olcao_write_pscf_h5.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_char_p, C.c_char_p]
    assert lib.olcao_write_pscf_h5(None, None, 1, b"k", b"/nonexistent") == 3
