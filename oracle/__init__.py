"""ORACLE -- test infrastructure, never part of the product path.

CPU restatements of the reference's algorithms used to check the HIP path:

* ``liboracle.so`` (C, built from the ``*.c`` files here by ``Makefile``):
  Philox4x32-10 and bit-exact replays of the checkerboard schedules.
* ``exact.py``: exact results (Kaufman torus, enumeration, O(n) rings).
* ``stats.py``: the Gamma-method error analysis used to judge pulls.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the checker.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None

_i8p = ctypes.POINTER(ctypes.c_int8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)


def build(force: bool = False) -> Path:
    """Compile ``liboracle.so`` with the Makefile next to this file."""
    if force and LIB_PATH.exists():
        LIB_PATH.unlink()
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


def _declare(L: ctypes.CDLL) -> None:
    L.orc_philox.argtypes = [_u32p, _u32p, _u32p]
    L.orc_philox.restype = None
    L.orc_ising_thresholds.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, _u64p, _u64p]
    L.orc_ising_thresholds.restype = None
    L.orc_ising_site_uniform.argtypes = [ctypes.c_uint32] * 3 + [ctypes.c_uint64, ctypes.c_int,
                                                                  ctypes.c_uint32, ctypes.c_int]
    L.orc_ising_site_uniform.restype = ctypes.c_uint32
    L.orc_ising_energy.argtypes = [_i8p, ctypes.c_int, _i32p, _i64p, _i64p]
    L.orc_ising_energy.restype = ctypes.c_int
    L.orc_ising_run.argtypes = [
        _i8p, ctypes.c_int, _i32p, ctypes.c_double, ctypes.c_double,
        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u64p,
        _i64p, ctypes.c_int, ctypes.c_int64, _i64p, _i64p, _i64p, _i64p, _i64p,
    ]
    L.orc_ising_run.restype = ctypes.c_int
    L.orc_ising_run_aux.argtypes = L.orc_ising_run.argtypes + [_i64p]
    L.orc_ising_run_aux.restype = ctypes.c_int
    L.orc_init_spins.argtypes = [_i8p, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_uint32]
    L.orc_init_spins.restype = None
    for name, decl in _EXTRA_DECLS.items():
        if hasattr(L, name):
            fn = getattr(L, name)
            fn.argtypes, fn.restype = decl


_EXTRA_DECLS: dict = {}


def _ptr(a: np.ndarray, typ):
    return a.ctypes.data_as(typ)


# ---------------------------------------------------------------------------
def philox(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox(_ptr(c, _u32p), _ptr(k, _u32p), _ptr(out, _u32p))
    return out


def ising_thresholds(beta: float, h: float, z: int) -> tuple[np.ndarray, np.ndarray]:
    met = np.zeros((2 * z + 1) * 2, dtype=np.uint64)
    hb = np.zeros(2 * z + 1, dtype=np.uint64)
    lib().orc_ising_thresholds(float(beta), float(h), int(z), _ptr(met, _u64p), _ptr(hb, _u64p))
    return met.reshape(2 * z + 1, 2), hb


def ising_energy(s: np.ndarray, dims) -> tuple[int, int]:
    s = np.ascontiguousarray(s, dtype=np.int8)
    d = np.ascontiguousarray(dims, dtype=np.int32)
    e = np.zeros(1, dtype=np.int64)
    m = np.zeros(1, dtype=np.int64)
    rc = lib().orc_ising_energy(_ptr(s, _i8p), len(d), _ptr(d, _i32p), _ptr(e, _i64p), _ptr(m, _i64p))
    if rc:
        raise ValueError(f"orc_ising_energy failed ({rc})")
    return int(e[0]), int(m[0])


def ising_init(n: int, key, stream: int = 0) -> np.ndarray:
    s = np.zeros(n, dtype=np.int8)
    lib().orc_init_spins(_ptr(s, _i8p), int(n), int(key[0]), int(key[1]), int(stream))
    return s


class IsingReplay:
    """One checkerboard Ising chain replayed on the CPU (mutates ``s`` in place)."""

    def __init__(self, s: np.ndarray, dims, beta: float, h: float, key, stream: int = 0,
                 sweep: int = 0):
        self.s = np.ascontiguousarray(s, dtype=np.int8).copy()
        self.dims = np.ascontiguousarray(dims, dtype=np.int32)
        self.beta = float(beta)
        self.h = float(h)
        self.key = (int(key[0]), int(key[1]))
        self.stream = int(stream)
        self.sweep = np.array([sweep], dtype=np.uint64)
        e, m = ising_energy(self.s, self.dims)
        self.e_b = np.array([e], dtype=np.int64)
        self.m = np.array([m], dtype=np.int64)

    def run(self, schedule, n_meas: int) -> dict[str, np.ndarray]:
        sched = np.ascontiguousarray(schedule, dtype=np.int64).reshape(-1, 2)
        e_out = np.zeros(n_meas, dtype=np.int64)
        m_out = np.zeros(n_meas, dtype=np.int64)
        acc = np.zeros(n_meas, dtype=np.int64)
        aux = np.zeros((n_meas, 4), dtype=np.int64)
        rc = lib().orc_ising_run_aux(
            _ptr(self.s, _i8p), len(self.dims), _ptr(self.dims, _i32p), self.beta, self.h,
            self.key[0], self.key[1], self.stream, _ptr(self.sweep, _u64p),
            _ptr(sched, _i64p), sched.shape[0], int(n_meas),
            _ptr(self.e_b, _i64p), _ptr(self.m, _i64p),
            _ptr(e_out, _i64p), _ptr(m_out, _i64p), _ptr(acc, _i64p), _ptr(aux, _i64p),
        )
        if rc:
            raise ValueError(f"orc_ising_run failed ({rc})")
        return {"Eb": e_out, "M": m_out, "accepted": acc, "aux": aux}


os.environ.setdefault("OMP_NUM_THREADS", str(min(16, os.cpu_count() or 1)))


_EXTRA_DECLS.update({
    "orc_ising_run_batch": ([_i8p, ctypes.c_int, ctypes.c_int, _i32p, _f64p, ctypes.c_double,
                             _u32p, _u32p, _u64p, _i64p, ctypes.c_int, ctypes.c_int64, _i64p,
                             _i64p, _i64p, _i64p, _i64p, ctypes.c_int], ctypes.c_int),
    "orc_ising_typewriter_batch": ([_i8p, ctypes.c_int, ctypes.c_int, _i32p, _f64p,
                                    ctypes.c_double, _u32p, ctypes.c_int64, _i64p, _i64p,
                                    ctypes.c_int], ctypes.c_int),
})


def ising_energy_batch(s: np.ndarray, dims) -> tuple[np.ndarray, np.ndarray]:
    e = np.zeros(s.shape[0], np.int64)
    m = np.zeros(s.shape[0], np.int64)
    for r in range(s.shape[0]):
        e[r], m[r] = ising_energy(s[r], dims)
    return e, m


def ising_batch_checkerboard(s: np.ndarray, dims, betas, h, keys, sids, sweeps, schedule,
                             n_meas: int, threads: int) -> dict[str, np.ndarray]:
    """Checkerboard replay of R chains with OpenMP (mutates ``s``, ``sweeps``)."""
    R = s.shape[0]
    d = np.ascontiguousarray(dims, np.int32)
    e, m = ising_energy_batch(s, dims)
    sched = np.ascontiguousarray(schedule, np.int64).reshape(-1, 2)
    eo = np.zeros((R, n_meas), np.int64)
    mo = np.zeros((R, n_meas), np.int64)
    ao = np.zeros((R, n_meas), np.int64)
    kk = np.ascontiguousarray(keys, np.uint32)
    ss = np.ascontiguousarray(sids, np.uint32)
    rc = lib().orc_ising_run_batch(
        _ptr(s, _i8p), R, len(d), _ptr(d, _i32p), _ptr(np.ascontiguousarray(betas, np.float64), _f64p),
        float(h), _ptr(kk, _u32p), _ptr(ss, _u32p), _ptr(sweeps, _u64p), _ptr(sched, _i64p),
        sched.shape[0], int(n_meas), _ptr(e, _i64p), _ptr(m, _i64p), _ptr(eo, _i64p),
        _ptr(mo, _i64p), _ptr(ao, _i64p), int(threads))
    if rc:
        raise ValueError(f"orc_ising_run_batch failed ({rc})")
    return {"Eb": eo, "M": mo, "accepted": ao}


def ising_batch_typewriter(s: np.ndarray, dims, betas, h, keys, n_sweeps: int,
                           threads: int) -> tuple[np.ndarray, np.ndarray]:
    R = s.shape[0]
    d = np.ascontiguousarray(dims, np.int32)
    e, m = ising_energy_batch(s, dims)
    kk = np.ascontiguousarray(keys, np.uint32)
    rc = lib().orc_ising_typewriter_batch(
        _ptr(s, _i8p), R, len(d), _ptr(d, _i32p), _ptr(np.ascontiguousarray(betas, np.float64), _f64p),
        float(h), _ptr(kk, _u32p), int(n_sweeps), _ptr(e, _i64p), _ptr(m, _i64p), int(threads))
    if rc:
        raise ValueError(f"orc_ising_typewriter_batch failed ({rc})")
    return e, m


_EXTRA_DECLS.update({
    "orc_potts_energy": ([_i8p, ctypes.c_int, _i32p, ctypes.c_int, _i64p, _i64p], ctypes.c_int),
    "orc_potts_run": ([_i8p, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_double, ctypes.c_uint32,
                       ctypes.c_uint32, ctypes.c_uint32, _u64p, _i64p, ctypes.c_int,
                       ctypes.c_int64, _i64p, _i64p, _i64p, _i64p, _i64p], ctypes.c_int),
    "orc_potts_run_aux": ([_i8p, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_double, ctypes.c_uint32,
                           ctypes.c_uint32, ctypes.c_uint32, _u64p, _i64p, ctypes.c_int,
                           ctypes.c_int64, _i64p, _i64p, _i64p, _i64p, _i64p, _i64p], ctypes.c_int),
    "orc_potts_init": ([_i8p, ctypes.c_int64, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                        ctypes.c_uint32], None),
})


def potts_energy(s: np.ndarray, dims, q: int) -> tuple[int, np.ndarray]:
    s = np.ascontiguousarray(s, dtype=np.int8)
    d = np.ascontiguousarray(dims, dtype=np.int32)
    e = np.zeros(1, np.int64)
    c = np.zeros(q, np.int64)
    rc = lib().orc_potts_energy(_ptr(s, _i8p), len(d), _ptr(d, _i32p), int(q), _ptr(e, _i64p),
                                _ptr(c, _i64p))
    if rc:
        raise ValueError(f"orc_potts_energy failed ({rc})")
    return int(e[0]), c


def potts_init(n: int, q: int, key, stream: int = 0) -> np.ndarray:
    s = np.zeros(n, dtype=np.int8)
    lib().orc_potts_init(_ptr(s, _i8p), int(n), int(q), int(key[0]), int(key[1]), int(stream))
    return s


class PottsReplay:
    """One checkerboard Potts chain replayed on the CPU."""

    def __init__(self, s, dims, q: int, beta: float, key, stream: int = 0, sweep: int = 0):
        self.s = np.ascontiguousarray(s, dtype=np.int8).copy()
        self.dims = np.ascontiguousarray(dims, dtype=np.int32)
        self.q = int(q)
        self.beta = float(beta)
        self.key = (int(key[0]), int(key[1]))
        self.stream = int(stream)
        self.sweep = np.array([sweep], dtype=np.uint64)
        e, c = potts_energy(self.s, self.dims, self.q)
        self.energy = np.array([e], np.int64)
        self.counts = c

    def run(self, schedule, n_meas: int) -> dict[str, np.ndarray]:
        sched = np.ascontiguousarray(schedule, dtype=np.int64).reshape(-1, 2)
        e_out = np.zeros(n_meas, np.int64)
        c_out = np.zeros((n_meas, self.q), np.int64)
        acc = np.zeros(n_meas, np.int64)
        aux = np.zeros((n_meas, 4), np.int64)
        rc = lib().orc_potts_run_aux(
            _ptr(self.s, _i8p), len(self.dims), _ptr(self.dims, _i32p), self.q, self.beta,
            self.key[0], self.key[1], self.stream, _ptr(self.sweep, _u64p), _ptr(sched, _i64p),
            sched.shape[0], int(n_meas), _ptr(self.energy, _i64p), _ptr(self.counts, _i64p),
            _ptr(e_out, _i64p), _ptr(c_out, _i64p), _ptr(acc, _i64p), _ptr(aux, _i64p))
        if rc:
            raise ValueError(f"orc_potts_run failed ({rc})")
        return {"E": e_out, "counts": c_out, "accepted": acc, "aux": aux}


_EXTRA_DECLS.update({
    "orc_det_exp": ([ctypes.c_double], ctypes.c_double),
    "orc_ea_energy": ([_i8p, _i8p, ctypes.c_int, ctypes.c_int], ctypes.c_int32),
    "orc_ea_energy_f64": ([_i8p, _f64p, ctypes.c_int, ctypes.c_int], ctypes.c_double),
    "orc_ea_init": ([_i8p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint32], None),
    "orc_ea_pt_run": ([_i8p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                       _f64p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, _u64p,
                       ctypes.c_void_p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _f64p,
                       ctypes.c_int, ctypes.c_void_p, _i32p, ctypes.c_uint64, ctypes.c_int64,
                       ctypes.c_int], ctypes.c_int),
})


def det_exp(x: float) -> float:
    return float(lib().orc_det_exp(float(x)))


class EaReplay:
    """One EA sample under parallel tempering, replayed on the CPU.

    ``spins``: [2, m, N] int8 start; ``J``: [d*N] forward-bond couplings (all
    +-1: the integer path, else FP64 as for Gaussian disorder); ``betas``:
    hot -> cold.  State arrays follow the device layout.
    """

    def __init__(self, spins, J, ndim: int, L: int, betas, key, code: int = 2,
                 sweeps_per_step: int = 1, multispin: bool | None = None):
        self.spins = np.ascontiguousarray(spins, dtype=np.int8).copy()
        J = np.asarray(J, dtype=np.float64)
        self.jtype = 0 if np.all(np.abs(J) == 1.0) else 1
        self.J = np.ascontiguousarray(J if self.jtype else J.astype(np.int8))
        self.d, self.L = int(ndim), int(L)
        self.betas = np.ascontiguousarray(betas, dtype=np.float64)
        self.m = len(self.betas)
        self.key = (int(key[0]), int(key[1]))
        self.code, self.sps = int(code), int(sweeps_per_step)
        m = self.m
        self.N = self.L ** self.d
        self.sweeps = np.zeros(2 * m, np.uint64)
        if self.jtype:
            self.energies = np.array([lib().orc_ea_energy_f64(_ptr(self.spins[r, c], _i8p),
                                                              _ptr(self.J, _f64p), self.d, self.L)
                                      for r in range(2) for c in range(m)], dtype=np.float64)
        else:
            self.energies = np.array([lib().orc_ea_energy(_ptr(self.spins[r, c], _i8p),
                                                          _ptr(self.J, _i8p), self.d, self.L)
                                      for r in range(2) for c in range(m)], dtype=np.int32)
        self.perm = np.tile(np.arange(m, dtype=np.int32), 2)
        self.kofc = self.perm.copy()
        self.labels = np.zeros(2 * m, np.int32)
        self.step = 0
        # uniform mapping: the multi-spin one where the device takes a batch of
        # this sample alone multi-spin coded (+-J, m <= 32, L <= 32 in 2D or
        # L <= 16 in 3D); a mixed batch passes the batch's choice
        if multispin is None:
            multispin = (self.jtype == 0 and m <= 32
                         and self.L <= (32 if self.d == 2 else 16))
        self.multispin = bool(multispin)

    def run(self, n_steps: int, n_bins: int, record: bool = True) -> dict[str, np.ndarray]:
        m = self.m
        trips = np.zeros(2, np.int32)
        hist = np.zeros(2 * m * 2, np.int32)
        sacc = np.zeros(2 * (m - 1), np.int32)
        stry = np.zeros(2 * (m - 1), np.int32)
        acc = np.zeros(n_bins * m * 7, np.float64)
        es = np.zeros(2 * m * n_steps, np.float64 if self.jtype else np.int32)
        qs = np.zeros(m * n_steps, np.int32)
        rc = lib().orc_ea_pt_run(
            _ptr(self.spins, _i8p), self.J.ctypes.data, self.jtype, self.d, self.L, m,
            _ptr(self.betas, _f64p), self.code, self.sps, self.key[0], self.key[1],
            _ptr(self.sweeps, _u64p), self.energies.ctypes.data, _ptr(self.perm, _i32p),
            _ptr(self.kofc, _i32p), _ptr(self.labels, _i32p), _ptr(trips, _i32p), _ptr(hist, _i32p),
            _ptr(sacc, _i32p), _ptr(stry, _i32p), _ptr(acc, _f64p), int(n_bins), es.ctypes.data,
            _ptr(qs, _i32p), int(self.step), int(n_steps), int(self.multispin))
        if rc:
            raise ValueError(f"orc_ea_pt_run failed ({rc})")
        self.step += n_steps
        return {"acc": acc.reshape(n_bins, m, 7), "trips": trips, "label_hist": hist.reshape(2, m, 2),
                "swap_acc": sacc.reshape(2, m - 1), "swap_try": stry.reshape(2, m - 1),
                "E": es.reshape(2, m, n_steps), "q": qs.reshape(m, n_steps)}


def ea_init(m: int, N: int, key) -> np.ndarray:
    s = np.zeros(2 * m * N, np.int8)
    lib().orc_ea_init(_ptr(s, _i8p), int(m), int(N), int(key[0]), int(key[1]))
    return s.reshape(2, m, N)


_EXTRA_DECLS.update({
    "orc_on_half_sweep": ([_f64p, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                           ctypes.c_double, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32,
                           ctypes.c_uint32, ctypes.c_uint64], ctypes.c_int64),
    "orc_on_measure": ([_f64p, ctypes.c_int, _i32p, ctypes.c_int, _f64p, _f64p, _f64p, _f64p],
                       ctypes.c_int),
    "orc_on_run": ([_f64p, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                    ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u64p, _i64p, ctypes.c_int,
                    ctypes.c_int64, _f64p, _f64p, _f64p, _f64p, _i64p], ctypes.c_int),
    "orc_on_run_aux": ([_f64p, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u64p, _i64p, ctypes.c_int,
                        ctypes.c_int64, _f64p, _f64p, _f64p, _f64p, _i64p, _i64p], ctypes.c_int),
})


class OnReplay:
    """FP64 replay of one O(n) chain (``s``: [N, n])."""

    def __init__(self, s, dims, beta: float, step: float, key, stream: int = 0, sweep: int = 0):
        self.s = np.ascontiguousarray(s, dtype=np.float64).copy()
        self.n = self.s.shape[1]
        self.dims = np.ascontiguousarray(dims, dtype=np.int32)
        self.beta, self.step = float(beta), float(step)
        self.key = (int(key[0]), int(key[1]))
        self.stream = int(stream)
        self.sweep = np.array([sweep], np.uint64)

    def half_sweep(self, code: int, colour: int) -> int:
        return int(lib().orc_on_half_sweep(
            _ptr(self.s, _f64p), len(self.dims), _ptr(self.dims, _i32p), self.n, int(code),
            int(colour), self.beta, self.step, self.key[0], self.key[1], self.stream,
            int(self.sweep[0])))

    def run(self, schedule, n_meas: int) -> dict[str, np.ndarray]:
        sched = np.ascontiguousarray(schedule, np.int64).reshape(-1, 2)
        d = len(self.dims)
        e = np.zeros(n_meas)
        m = np.zeros((n_meas, self.n))
        sc = np.zeros((n_meas, d))
        ss = np.zeros((n_meas, d))
        acc = np.zeros(n_meas, np.int64)
        aux = np.zeros((n_meas, 4), np.int64)
        rc = lib().orc_on_run_aux(_ptr(self.s, _f64p), d, _ptr(self.dims, _i32p), self.n, self.beta,
                                  self.step, self.key[0], self.key[1], self.stream,
                                  _ptr(self.sweep, _u64p), _ptr(sched, _i64p), sched.shape[0],
                                  int(n_meas), _ptr(e, _f64p), _ptr(m, _f64p), _ptr(sc, _f64p),
                                  _ptr(ss, _f64p), _ptr(acc, _i64p), _ptr(aux, _i64p))
        if rc:
            raise ValueError(f"orc_on_run failed ({rc})")
        return {"E": e, "M": m, "stiff_c": sc, "stiff_s": ss, "accepted": acc, "aux": aux}


_EXTRA_DECLS.update({
    "orc_wl_run": ([_i8p, ctypes.c_int, _i32p, ctypes.c_int, _i32p, _f64p, _i64p, ctypes.c_int,
                    ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u64p,
                    ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int64], ctypes.c_int),
    "orc_muca_run": ([_i8p, ctypes.c_int, _i32p, ctypes.c_int, _i32p, _i32p, _f64p, _i64p,
                      ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                      _u64p, ctypes.c_int64, ctypes.c_int64, _i32p, _i64p], ctypes.c_int),
})


class FlatWalker:
    """One flat-histogram walker replayed on the CPU (q = 0: Ising, else Potts)."""

    def __init__(self, s, dims, q: int, idx: int, lo: int, hi: int, key, stream: int = 0):
        self.s = np.ascontiguousarray(s, dtype=np.int8).copy()
        self.dims = np.ascontiguousarray(dims, dtype=np.int32)
        self.q = int(q)
        self.idx = np.array([idx], np.int32)
        self.lo, self.hi = int(lo), int(hi)
        self.key = (int(key[0]), int(key[1]))
        self.stream = int(stream)
        self.moves = np.zeros(1, np.uint64)
        if self.q == 0:
            self.order = np.array([int(self.s.astype(np.int64).sum())], np.int32)
        else:
            self.order = np.bincount(self.s, minlength=self.q).astype(np.int32)

    def wl(self, ln_g: np.ndarray, hist: np.ndarray, ln_f: float, mode_1t: bool, n_levels: float,
           n_moves: int) -> None:
        lib().orc_wl_run(_ptr(self.s, _i8p), len(self.dims), _ptr(self.dims, _i32p), self.q,
                         _ptr(self.idx, _i32p), _ptr(ln_g, _f64p), _ptr(hist, _i64p), self.lo,
                         self.hi, self.key[0], self.key[1], self.stream, _ptr(self.moves, _u64p),
                         float(ln_f), int(mode_1t), float(n_levels), int(n_moves))

    def muca(self, ln_w: np.ndarray, hist: np.ndarray, n_meas: int, moves_per_meas: int):
        idx_out = np.zeros(n_meas, np.int32)
        ord_out = np.zeros(n_meas, np.int64)
        lib().orc_muca_run(_ptr(self.s, _i8p), len(self.dims), _ptr(self.dims, _i32p), self.q,
                           _ptr(self.idx, _i32p), _ptr(self.order, _i32p), _ptr(ln_w, _f64p),
                           _ptr(hist, _i64p), self.lo, self.hi, self.key[0], self.key[1],
                           self.stream, _ptr(self.moves, _u64p), int(n_meas), int(moves_per_meas),
                           _ptr(idx_out, _i32p), _ptr(ord_out, _i64p))
        return idx_out, ord_out


_EXTRA_DECLS.update({
    "orc_table_run": ([_i8p, ctypes.c_int64, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_double,
                       ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u64p,
                       _i64p, ctypes.c_int, ctypes.c_int64, _i64p, _i64p, _i64p, _i64p, _i64p],
                      ctypes.c_int),
})


class TableReplay:
    """One Ising (``q`` = 0) or Potts chain on a lattice given by its neighbour table,
    in the reference's own site orders (replay_table.c); mutates ``s``."""

    def __init__(self, s, nbr, q: int, beta: float, h: float, key, stream: int = 0,
                 sweep: int = 0):
        self.s = np.ascontiguousarray(s, dtype=np.int8).copy()
        self.nbr = np.ascontiguousarray(nbr, dtype=np.int32)
        self.N, self.z = self.nbr.shape
        self.q, self.beta, self.h = int(q), float(beta), float(h)
        self.key = (int(key[0]), int(key[1]))
        self.stream = int(stream)
        self.sweep = np.array([sweep], np.uint64)
        e = self._energy()
        self.e = np.array([e], np.int64)
        if self.q == 0:
            self.mc = np.array([int(self.s.astype(np.int64).sum())], np.int64)
        else:
            self.mc = np.bincount(self.s, minlength=self.q).astype(np.int64)

    def _energy(self) -> int:
        s = self.s.astype(np.int64)
        fwd = self.nbr[:, 0::2]
        if self.q == 0:
            return int(-(s[:, None] * s[fwd]).sum())
        return int(-(s[:, None] == s[fwd]).sum())

    def run(self, schedule, n_meas: int) -> dict[str, np.ndarray]:
        sched = np.ascontiguousarray(schedule, np.int64).reshape(-1, 2)
        nmc = 1 if self.q == 0 else self.q
        e_out = np.zeros(n_meas, np.int64)
        mc_out = np.zeros((n_meas, nmc), np.int64)
        aux = np.zeros((n_meas, 4), np.int64)
        rc = lib().orc_table_run(
            _ptr(self.s, _i8p), self.N, self.z, _ptr(self.nbr, _i32p), self.q, self.beta, self.h,
            self.key[0], self.key[1], self.stream, _ptr(self.sweep, _u64p), _ptr(sched, _i64p),
            sched.shape[0], int(n_meas), _ptr(self.e, _i64p), _ptr(self.mc, _i64p),
            _ptr(e_out, _i64p), _ptr(mc_out, _i64p), _ptr(aux, _i64p))
        if rc:
            raise ValueError(f"orc_table_run failed ({rc})")
        out = {"E": e_out, "aux": aux}
        if self.q == 0:
            out["M"] = mc_out[:, 0]
        else:
            out["counts"] = mc_out
        return out


_EXTRA_DECLS.update({
    "orc_seq_on_run": ([_f64p, ctypes.c_int64, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_double,
                        ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u64p,
                        _i64p, ctypes.c_int, ctypes.c_int64, _f64p, _f64p, _f64p, _f64p, _i64p],
                       ctypes.c_int),
})


class SeqOnReplay:
    """One O(n) chain (n = 2, 3) in FP64 on a lattice given by its neighbour table,
    in the reference's own site orders (replay_seq_on.c); mutates ``s`` ([N, n])."""

    def __init__(self, s, nbr, n: int, beta: float, step: float, key, stream: int = 0,
                 sweep: int = 0):
        self.s = np.ascontiguousarray(s, dtype=np.float64).reshape(-1, int(n)).copy()
        self.nbr = np.ascontiguousarray(nbr, dtype=np.int32)
        self.N, self.z = self.nbr.shape
        self.n, self.beta, self.step = int(n), float(beta), float(step)
        self.key = (int(key[0]), int(key[1]))
        self.stream = int(stream)
        self.sweep = np.array([sweep], np.uint64)

    def run(self, schedule, n_meas: int) -> dict[str, np.ndarray]:
        sched = np.ascontiguousarray(schedule, np.int64).reshape(-1, 2)
        A = self.z // 2
        e = np.zeros(n_meas)
        m = np.zeros((n_meas, self.n))
        sc = np.zeros((n_meas, A))
        ss = np.zeros((n_meas, A))
        aux = np.zeros((n_meas, 4), np.int64)
        rc = lib().orc_seq_on_run(
            _ptr(self.s, _f64p), self.N, self.z, _ptr(self.nbr, _i32p), self.n, self.beta,
            self.step, self.key[0], self.key[1], self.stream, _ptr(self.sweep, _u64p),
            _ptr(sched, _i64p), sched.shape[0], int(n_meas), _ptr(e, _f64p), _ptr(m, _f64p),
            _ptr(sc, _f64p), _ptr(ss, _f64p), _ptr(aux, _i64p))
        if rc:
            raise ValueError(f"orc_seq_on_run failed ({rc})")
        return {"E": e, "M": m, "stiff_c": sc, "stiff_s": ss, "aux": aux}
