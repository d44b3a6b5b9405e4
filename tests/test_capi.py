"""C-ABI boundary checks that need no GPU: the library loads, exports every entry
point include/tdx.h declares, and rejects invalid arguments with a status code
(the reference prints and returns instead, src/mttkrp.cu:75-77, src/ttm.cu:41-43,
src/tensor2matrix.cu:48-50)."""
import ctypes
import os
import re

import pytest

import tdx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "tdx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tdx_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(tdx.LIB_PATH)
    syms = header_symbols()
    assert len(syms) > 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(tdx.HEADER_SYMBOLS) == header_symbols()


def test_cudecompose_header_names():
    # the drop-in C++ header declares the reference's 22 host functions by name
    txt = open(os.path.join(ROOT, "include", "cudecompose.h")).read()
    for name in ["hadmard", "hadmard3", "kr", "tensor_product", "printTensor", "f2h", "ttm", "mttkrp", "t2m",
                 "tminust", "rse", "gencp", "gentucker", "gencp4", "gentucker4", "cp_als", "cp_als4",
                 "cp_tensorcore", "tucker_hosvd", "tucker_hosvd4", "tucker_tensorcore", "initMat"]:
        assert re.search(r"\b%s\s*\(" % name, txt), name


@pytest.mark.parametrize("call", [
    lambda L: L.tdx_t2m(None, None, 4, 2, 2, 2),
    lambda L: L.tdx_t2m(ctypes.c_void_p(16), ctypes.c_void_p(16), 0, 2, 2, 2),
    lambda L: L.tdx_mttkrp(ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(16),
                           4, 2, 2, 2, 2),
    lambda L: L.tdx_ttm(ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(16), 5, 2, 2, 2, 2),
    lambda L: L.tdx_kr(None, None, None, 2, 2, 2),
    lambda L: L.tdx_hadamard(ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(16), -1, 2),
])
def test_invalid_arguments_return_status(call):
    L = tdx.lib()
    assert call(L) == tdx.TDX_ERR_ARG
    assert L.tdx_last_status() == tdx.TDX_ERR_ARG


def test_status_strings():
    L = tdx.lib()
    for st in range(8):
        assert len(L.tdx_status_string(st)) > 0
