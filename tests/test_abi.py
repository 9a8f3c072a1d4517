"""The C ABI: libraries load, every declared symbol is exported, and the ctypes mirrors match the C layouts.
No compute call touches the GPU here."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return set(re.findall(r"WS_API\s+[\w\s\*]+?\b(ws[m]?_\w+)\s*\(", txt))


def test_weave_h_symbols_exported(W):
    A = W._abi
    lib = C.CDLL(os.path.join(A.LIB_DIR, "libweave_hip.so"))
    declared = _declared("weave.h")
    assert declared == set(A.WEAVE_H_SYMBOLS), declared ^ set(A.WEAVE_H_SYMBOLS)
    for s in declared:
        assert hasattr(lib, s), s


def test_weave_model_h_symbols_exported(W):
    A = W._abi
    lib = A.model_lib()
    declared = _declared("weave_model.h")
    assert declared == set(A.WEAVE_MODEL_H_SYMBOLS), declared ^ set(A.WEAVE_MODEL_H_SYMBOLS)
    for s in declared:
        assert hasattr(lib, s), s


def test_struct_layouts_match(W):
    A = W._abi
    out = (C.c_int64 * 32)()
    n = A.model_lib().wsm_abi_sizes(out, 32)
    assert n == len(A.ABI_STRUCTS)
    for i, name in enumerate(A.ABI_STRUCTS):
        assert out[i] == C.sizeof(getattr(A, name)), name


def test_hip_library_has_gfx950_code(W):
    """the device library carries a gfx950 code object (a build for the right target)"""
    data = open(os.path.join(W._abi.LIB_DIR, "libweave_hip.so"), "rb").read()
    assert b"gfx950" in data


def test_device_errors_are_loud_without_gpu(W):
    """on a machine without a GPU the device path raises instead of falling back to the CPU"""
    import torch  # noqa: F401  (only to probe visibility, no allocation)
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and _has_gpu():
        pytest.skip("a GPU is visible")
    with pytest.raises(W.solver.DeviceError):
        W.solver.Context(0)


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False
