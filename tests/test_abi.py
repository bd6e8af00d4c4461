"""The C-ABI library loads and exports exactly what include/*.h declares (no GPU needed)."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(cm_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert "cm_ising_run" in names and "cm_last_error" in names


def test_library_exports_every_declared_symbol():
    from critmc_amd._lib import SIGNATURES, lib

    L = lib()
    for name in sorted(_declared()):
        assert hasattr(L, name), f"{name} declared but not exported"
        assert name in SIGNATURES, f"{name} has no ctypes signature"
    assert set(SIGNATURES) <= _declared()
    assert L.cm_version() == 1


def test_argument_errors_are_reported_without_a_gpu():
    from critmc_amd._lib import CM_EINVAL, lib

    L = lib()
    dims = np.array([5, 4], dtype=np.int32)  # odd extent: no checkerboard
    betas = np.array([0.4])
    sched = np.array([[0, 1]], dtype=np.int64)
    sw = np.zeros(1, dtype=np.uint64)
    P = ctypes.c_void_p
    rc = L.cm_ising_run(P(0), 0, 2, dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 1,
                        betas.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 0.0, P(0), P(0),
                        P(0), sw.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                        sched.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 1, 10, P(0), P(0),
                        P(0), P(0), P(0), P(0), 1 << 20, 0, 0, 0, P(0), None, P(0))
    assert rc == CM_EINVAL
    assert b"even" in L.cm_last_error()


def test_library_refuses_cluster_codes():
    from critmc_amd._lib import CM_EINVAL, lib

    L = lib()
    dims = np.array([4, 4], dtype=np.int32)
    betas = np.array([0.4])
    sched = np.array([[3, 1]], dtype=np.int64)  # wolff
    sw = np.zeros(1, dtype=np.uint64)
    P = ctypes.c_void_p
    rc = L.cm_ising_run(P(0), 0, 2, dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 1,
                        betas.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 0.0, P(0), P(0),
                        P(0), sw.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                        sched.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 1, 10, P(0), P(0),
                        P(0), P(0), P(0), P(0), 1 << 20, 0, 0, 0, P(0), None, P(0))
    assert rc == CM_EINVAL
    assert b"checkerboard" in L.cm_last_error()


def test_copy_spans_checks_its_count_without_a_gpu():
    from critmc_amd._lib import CM_EINVAL, CM_OK, lib

    L = lib()
    assert L.cm_copy_spans(ctypes.c_void_p(0), -1, None) == CM_EINVAL
    assert b"negative" in L.cm_last_error()
    assert L.cm_copy_spans(ctypes.c_void_p(0), 0, None) == CM_OK  # nothing to copy, no launch
