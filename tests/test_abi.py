"""CPU-side checks of the C-ABI boundary: the libraries load, export every
function the public headers declare, and agree with the struct layouts."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

import rta_amd
from rta_amd import abi


def test_hip_library_exports_every_declared_symbol():
    lib = rta_amd.hip()
    declared = abi.declared_functions("rta_hip.h")
    assert len(declared) >= 30
    missing = [name for name in declared if not hasattr(lib, name)]
    assert not missing, missing
    # and the ctypes signature table covers them all
    assert set(declared) - {"rta_hip_last_error"} <= set(abi.HIP_SIGNATURES)


def test_host_library_exports_every_declared_symbol():
    lib = rta_amd.host()
    declared = abi.declared_functions("rta_host.h")
    assert declared == sorted(abi.HOST_SIGNATURES)
    for name in declared:
        assert hasattr(lib, name)


def test_layouts():
    assert abi.TRI_DTYPE.itemsize == 104
    assert abi.NODE_DTYPE.itemsize == 32
    assert abi.HIT_DTYPE.itemsize == 16
    assert abi.AABB_DTYPE.itemsize == 24


def test_argument_errors_need_no_gpu():
    # argument validation happens on the host before any HIP call
    with_err = rta_amd.hip().rta_lbvh_workspace_bytes(0, None)
    assert with_err == abi.RTA_E_INVALID
    sz = C.c_size_t(0)
    assert rta_amd.hip().rta_lbvh_sort_work_bytes(1000, C.byref(sz)) == 0 and sz.value > 8000
    assert rta_amd.hip().rta_lbvh_workspace_bytes(1000, C.byref(sz)) == 0
    assert sz.value > 1000 * (104 + 24 + 12 + 4 + 8)
    msg = rta_amd.hip().rta_hip_last_error()
    assert isinstance(msg, bytes)


def test_injected_fault_surfaces_as_an_error():
    """The fault hook: the armed entry point's next call returns the injected
    status before touching its arguments (so no GPU is needed), the Python
    layer raises it with the message, other entry points are unaffected, and
    the hook is one-shot."""
    lib = rta_amd.hip()
    fake = C.c_void_p(64)                         # never dereferenced on these paths
    assert lib.rta_hip_inject_fault(b"rta_lbvh_build", 0) == abi.RTA_E_INVALID
    assert lib.rta_hip_inject_fault(b"rta_lbvh_build", 701) == 0
    # another entry point: its own argument check, not the fault
    assert lib.rta_lbvh_workspace_bytes(0, None) == abi.RTA_E_INVALID
    try:
        abi.call("rta_lbvh_build", fake, 10, fake, fake, fake, C.c_size_t(1 << 20), None, None,
                 None)
    except abi.RtaError as e:
        assert e.code == 701 and e.func == "rta_lbvh_build"
        assert "injected fault" in str(e)
    else:
        raise AssertionError("the injected fault must raise")
    # one shot: the next call runs its normal argument check (n = 0)
    assert lib.rta_lbvh_build(fake, 0, fake, fake, fake, C.c_size_t(1 << 20), None, None,
                              None) == abi.RTA_E_INVALID
    assert b"n must be > 0" in lib.rta_hip_last_error()


def test_library_is_built_for_gfx950():
    data = open(abi.HIP_LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_no_cpu_fallback_when_library_missing(tmp_path, monkeypatch):
    monkeypatch.setattr(abi, "_hip", None)
    monkeypatch.setattr(abi, "HIP_LIB_PATH", str(tmp_path / "missing.so"))
    try:
        abi.hip()
    except RuntimeError as e:
        assert "not built" in str(e)
    else:
        raise AssertionError("loading a missing library must fail")
