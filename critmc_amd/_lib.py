"""ctypes binding of ``libcritmc_hip.so`` (the C-ABI declared in include/critmc_hip.h).

The library is the only compute path: if it is missing, or no GPU is present,
every operation raises instead of falling back to anything else.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libcritmc_hip.so"

CM_OK = 0
CM_EINVAL = -1
CM_EUNSUPPORTED = -2
CM_EHIP = -3

_c_i32 = ctypes.c_int32
_c_i64 = ctypes.c_int64
_c_u32 = ctypes.c_uint32
_c_u64 = ctypes.c_uint64
_c_f64 = ctypes.c_double
_c_f32 = ctypes.c_float
_vp = ctypes.c_void_p
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)

# name -> (argtypes, restype); must match include/critmc_hip.h
SIGNATURES: dict[str, tuple[list, object]] = {
    "cm_version": ([], _c_i32),
    "cm_last_error": ([], ctypes.c_char_p),
    "cm_philox_fill": ([_vp, _c_i64, _c_u32, _c_u32, _c_u32, _c_u32, _c_u32, _c_u32, _vp], _c_i32),
    "cm_philox_host": ([_u32p, _u32p, _u32p], None),
    "cm_copy_spans": ([_vp, _c_i32, _vp], _c_i32),
    "cm_device_info": ([_i32p, _i64p, _i32p, _i32p], _c_i32),
    "cm_ising_workspace_bytes": ([_c_i32], _c_i64),
    "cm_ising_tables": ([_c_f64, _c_f64, _c_i32, _u32p], _c_i32),
    "cm_ising_init": ([_vp, _c_i32, _i32p, _c_i32, _vp, _vp, _c_i32, _vp], _c_i32),
    "cm_ising_measure": ([_vp, _c_i32, _i32p, _c_i32, _vp, _vp, _vp], _c_i32),
    "cm_ising_convert": ([_vp, _vp, _c_i32, _i32p, _c_i32, _c_i32, _vp], _c_i32),
    "cm_ising_run": (
        [_vp, _c_i32, _c_i32, _i32p, _c_i32, _f64p, _c_f64, _vp, _vp, _vp, _u64p, _i64p, _c_i32,
         _c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _c_i64, _c_i32, _c_i32, _c_i32, _vp, _i32p, _vp],
        _c_i32,
    ),
    "cm_ising_prepare": (
        [_c_i32, _i32p, _c_i32, _f64p, _c_f64, _i64p, _c_i32, _vp, _c_i64, _i32p, _vp], _c_i32),
    "cm_ising_cluster_workspace_bytes": ([_c_i32, _i32p, _c_i32], _c_i64),
    "cm_ising_cluster_run": (
        [_vp, _c_i32, _i32p, _c_i32, _f64p, _c_f64, _vp, _vp, _vp, _u64p, _i64p, _c_i32, _c_i64,
         _vp, _vp, _vp, _vp, _vp, _vp, _c_i64, _vp],
        _c_i32,
    ),
    "cm_potts_workspace_bytes": ([_c_i32], _c_i64),
    "cm_potts_cluster_workspace_bytes": ([_c_i32, _i32p, _c_i32, _c_i32], _c_i64),
    "cm_potts_cluster_run": (
        [_vp, _c_i32, _i32p, _c_i32, _c_i32, _f64p, _vp, _vp, _vp, _i64p, _c_i32, _c_i64, _vp,
         _vp, _vp, _vp, _vp, _vp, _c_i64, _vp],
        _c_i32,
    ),
    "cm_potts_init": ([_vp, _c_i32, _i32p, _c_i32, _c_i32, _vp, _vp, _c_i32, _vp], _c_i32),
    "cm_potts_measure": ([_vp, _c_i32, _i32p, _c_i32, _c_i32, _vp, _vp, _vp], _c_i32),
    "cm_potts_run": (
        [_vp, _c_i32, _i32p, _c_i32, _c_i32, _f64p, _vp, _vp, _vp, _i64p, _c_i32, _c_i64, _vp,
         _vp, _vp, _vp, _vp, _vp, _c_i64, _c_i32, _c_i32, _vp, _vp],
        _c_i32,
    ),
    "cm_potts_prepare": ([_c_i32, _i32p, _c_i32, _f64p, _vp, _c_i64, _vp], _c_i32),
    "cm_ea_workspace_bytes": ([_c_i32, _c_i32, _c_i64], _c_i64),
    "cm_ea_multispin": ([_c_i32, _c_i32, _c_i32, _c_i32], _c_i32),
    "cm_det_exp": ([_c_f64], _c_f64),
    "cm_ea_init": ([_vp, _vp, _vp, _c_i32, _c_i32, _c_i32, _vp, _vp], _c_i32),
    "cm_ea_energy": ([_vp, _vp, _c_i32, _vp, _vp, _vp, _c_i32, _c_i32, _c_i32, _vp, _vp], _c_i32),
    "cm_ea_pt_run": (
        [_vp, _vp, _c_i32, _vp, _vp, _vp, _c_i32, _c_i32, _c_i32, _i32p, _f64p, _vp, _vp, _vp, _vp, _vp,
         _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_i32, _vp, _vp, _c_i32, _c_i32, _c_u64, _c_i64,
         _vp, _c_i64, _c_i32, _vp],
        _c_i32,
    ),
    "cm_set_option": ([ctypes.c_char_p, _c_i64, _i64p], _c_i32),
    "cm_table_workspace_bytes": ([_c_i64, _c_i32, _c_i32], _c_i64),
    "cm_ising_table_run": (
        [_vp, _c_i64, _c_i32, _vp, _c_i32, _f64p, _c_f64, _vp, _vp, _vp, _i64p, _c_i32, _c_i64,
         _vp, _vp, _vp, _vp, _vp, _vp, _c_i64, _vp], _c_i32),
    "cm_potts_table_run": (
        [_vp, _c_i64, _c_i32, _vp, _c_i32, _c_i32, _f64p, _vp, _vp, _vp, _i64p, _c_i32, _c_i64,
         _vp, _vp, _vp, _vp, _vp, _vp, _c_i64, _vp], _c_i32),
    "cm_table_measure": ([_vp, _c_i64, _c_i32, _vp, _c_i32, _c_i32, _vp, _vp, _vp], _c_i32),
    "cm_on_table_workspace_bytes": ([_c_i64, _c_i32], _c_i64),
    "cm_on_table_run": (
        [_vp, _c_i64, _c_i32, _vp, _c_i32, _c_i32, _f64p, _f64p, _vp, _vp, _vp, _i64p, _c_i32,
         _c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _c_i64, _vp], _c_i32),
    "cm_on_workspace_bytes": ([_c_i32, _i32p, _c_i32, _c_i32], _c_i64),
    "cm_on_init": ([_vp, _c_i32, _i32p, _c_i32, _c_i32, _vp, _vp, _vp], _c_i32),
    "cm_on_convert": ([_vp, _vp, _c_i32, _i32p, _c_i32, _c_i32, _c_i32, _vp], _c_i32),
    "cm_mbar_workspace_bytes": ([_c_i32, _c_i32], _c_i64),
    "cm_mbar_objective": ([_vp, _vp, _c_i64, _vp, _vp, _vp, _c_i32, _vp, _vp, _vp, _c_i64, _vp],
                          _c_i32),
    "cm_mbar_objective_noalloc": ([_vp, _vp, _c_i64, _vp, _vp, _vp, _c_i32, _vp, _vp, _vp],
                                  _c_i32),
    "cm_mbar_log_z": ([_vp, _vp, _vp, _c_i64, _vp, _c_i32, _vp, _vp, _c_i64, _vp], _c_i32),
    "cm_wl_run": ([_vp, _c_i32, _i32p, _c_i32, _c_i32, _vp, _vp, _vp, _c_i32, _vp, _vp, _vp, _vp,
                   _vp, _vp, _vp, _vp, _vp, _c_i64, _vp], _c_i32),
    "cm_muca_run": ([_vp, _c_i32, _i32p, _c_i32, _c_i32, _vp, _vp, _vp, _vp, _c_i32, _vp, _vp, _vp,
                     _vp, _vp, _c_i64, _c_i64, _vp, _vp, _vp], _c_i32),
    "cm_on_cluster_workspace_bytes": ([_c_i32, _i32p, _c_i32], _c_i64),
    "cm_on_run": (
        [_vp, _c_i32, _i32p, _c_i32, _c_i32, _vp, _vp, _vp, _vp, _vp, _i64p, _c_i32, _c_i64,
         _vp, _vp, _vp, _vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _c_i64, _vp],
        _c_i32,
    ),
}

_lib: ctypes.CDLL | None = None


class LibraryMissing(ImportError):
    pass


def lib() -> ctypes.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise LibraryMissing(
                f"{LIB_PATH} is not built; run `python -m critmc_amd.build` (needs hipcc)"
            )
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        _lib = handle
    return _lib


def check(rc: int, what: str = "") -> None:
    """Map a C return code to a Python exception carrying the library's message."""
    if rc == CM_OK:
        return
    msg = lib().cm_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == CM_EINVAL:
        raise ValueError(text)
    if rc == CM_EUNSUPPORTED:
        raise NotImplementedError(text)
    raise RuntimeError(text)


def set_option(name: str, value: int) -> int:
    """Set a library tuning option (see ``cm_set_option``); returns the old value."""
    old = ctypes.c_int64(0)
    check(lib().cm_set_option(name.encode(), int(value), ctypes.byref(old)), "cm_set_option")
    return int(old.value)
