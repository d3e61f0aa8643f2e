"""The C-ABI library: it is built, loads without a GPU and exports every
function include/skkr_gpu.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "skkr_gpu.h")
LIB = os.path.join(ROOT, "skkr_budapest_amd", "libskkr_gpu.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(skkr_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_k_loop_entry_points():
    d = declared()
    for name in ("skkr_create", "skkr_destroy", "skkr_set_structure", "skkr_set_tinv",
                 "skkr_tau_bz", "skkr_dugo_batch", "skkr_tau2d_batch",
                 "skkr_surfgfd_batch", "skkr_gjinv_batch"):
        assert name in d


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_binding_lists_every_symbol():
    from skkr_budapest_amd import gpu
    assert sorted(gpu.SYMBOLS) == declared()


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_argument_validation_without_device():
    lib = ctypes.CDLL(LIB)
    lib.skkr_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    assert lib.skkr_create(None, None) == 1          # SKKR_ERR_ARG
    lib.skkr_error_string.restype = ctypes.c_char_p
    assert lib.skkr_error_string(6) == b"decimation exceeded itermax"


MLP_HEADER = os.path.join(ROOT, "include", "skkr_mlp.h")
MLP_LIB = os.path.join(ROOT, "skkr_budapest_amd", "libskkr_mlp.so")


def _declared_in(path):
    src = re.sub(r"/\*.*?\*/", "", open(path).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(skkr_[a-z0-9_]+)\s*\(", src)))


@pytest.mark.skipif(not os.path.exists(MLP_LIB), reason="library not built")
def test_mlp_library_exports_every_declared_symbol():
    d = _declared_in(MLP_HEADER)
    assert "skkr_mlp_train" in d and "skkr_mlp_loss_grad" in d
    lib = ctypes.CDLL(MLP_LIB)
    missing = [s for s in d if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(MLP_LIB), reason="library not built")
def test_mlp_argument_validation_without_device():
    lib = ctypes.CDLL(MLP_LIB)
    lib.skkr_mlp_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_ulonglong]
    h = ctypes.c_void_p()
    assert lib.skkr_mlp_create(ctypes.byref(h), 0, 64, 0, 1) == 1     # nin < 1
    assert lib.skkr_mlp_create(ctypes.byref(h), 8, 32, 0, 1) == 1     # nhid != 64
    assert lib.skkr_mlp_create(None, 8, 64, 0, 1) == 1


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_screening_entry_points_validate_arguments_without_device():
    lib = ctypes.CDLL(LIB)
    lib.skkr_set_screening.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_void_p]
    assert lib.skkr_set_screening(None, 3, None, None) == 1      # SKKR_ERR_ARG
    lib.skkr_tau_bz_phys.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int] + [ctypes.c_void_p] * 8
    assert lib.skkr_tau_bz_phys(None, 0.5, 0.1, 0, *([None] * 8)) == 1


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_structure_constant_entry_points_validate_arguments_without_device():
    from skkr_budapest_amd import gpu
    lib = ctypes.CDLL(LIB)
    lib.skkr_str_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    assert lib.skkr_str_create(None, None, None) == 1                      # SKKR_ERR_ARG
    h = ctypes.c_void_p()
    cv = (ctypes.c_double * 9)(0, 0, 0, 1, 0, 1, 2, 0, 2)
    A = ctypes.c_double * 2
    for (lmax, s, nextra, eta, code) in ((2, 2, 0, 1.0, 1),   # nextra < 1
                                         (2, 2, 1, 0.0, 1),   # eta <= 0
                                         (4, 2, 1, 1.0, 7),   # lmax > 3
                                         (3, 4, 3, 1.0, 7)):  # nl2 (2 nextra + 3) s > 512
        cfg = gpu.StrConfig(lmax, s, nextra, eta, A(5.0, 0.0), A(0.0, 5.0), 0, 0)
        assert lib.skkr_str_create(ctypes.byref(h), ctypes.byref(cfg), cv) == code
    lib.skkr_gstore_batch.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p]
    assert lib.skkr_gstore_batch(None, 0.5, 0.05, None, 1, None, None, None) == 1
    lib.skkr_error_string.restype = ctypes.c_char_p
    assert lib.skkr_error_string(8) == b"structure-constant shell sum not converged"
