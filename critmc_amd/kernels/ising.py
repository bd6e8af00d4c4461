"""Checkerboard Ising kernels: host side of ``cm_ising_*``.

Replaces critmc/kernels/ising.py's ``run`` (:173-237) for the Metropolis
(:68-87) and heat-bath (:90-107) sweeps, and with :func:`run_cluster` for
schedules with Wolff / Swendsen-Wang rows (:110-170).  Spins stay on the
device as int8 ``[R, N]`` (x fastest) or as the packed ``[R, 2, rows, G]`` u32
image.
"""

from __future__ import annotations

import ctypes
from collections.abc import Sequence

import numpy as np
import torch

from .._lib import check, lib
from ..device import device, host_ptr, ptr, stream
from .chains import Chains
from .common import HEATBATH, METROPOLIS, SWENDSEN_WANG, WOLFF

LAYOUT_INT8 = 0
LAYOUT_PACKED = 1
FORM_AUTO, FORM_PLAIN, FORM_BATCHED, FORM_WIDE = 0, 1, 2, 3
START_CODES = {"random": 0, "ordered": 1, "up": 1, "down": 2}


def _dims(shape: Sequence[int]) -> np.ndarray:
    return np.ascontiguousarray(shape, dtype=np.int32)


def packed_words(shape: Sequence[int]) -> int:
    """u32 words per replica in the packed layout."""
    L0 = int(shape[0])
    rows = int(np.prod(shape)) // L0
    return 2 * rows * ((L0 // 2 + 31) // 32)


def thresholds_table(beta: float, h: float, z: int) -> np.ndarray:
    """The library's class table for one chain (64 u32), for inspection and tests."""
    out = np.zeros(64, dtype=np.uint32)
    check(lib().cm_ising_tables(float(beta), float(h), int(z), host_ptr(out, ctypes.c_uint32)),
          "cm_ising_tables")
    return out


def init_spins(spins: torch.Tensor, shape, chains: Chains, start: str = "random") -> None:
    """Fill int8 spins ``[R, N]`` on the device (random draws use the chain's Philox key)."""
    if start not in START_CODES:
        raise ValueError(f"unknown start {start!r}")
    d = _dims(shape)
    rc = lib().cm_ising_init(ptr(spins), len(d), host_ptr(d, ctypes.c_int32), chains.R,
                             ptr(chains.keys_d), ptr(chains.sids_d), START_CODES[start], stream())
    check(rc, "cm_ising_init")


def measure(spins: torch.Tensor, shape, R: int) -> tuple[torch.Tensor, torch.Tensor]:
    """Exact E_b and M per replica (int32 device tensors), critmc/kernels/ising.py:31-44."""
    d = _dims(shape)
    e = torch.empty(R, dtype=torch.int32, device=spins.device)
    m = torch.empty(R, dtype=torch.int32, device=spins.device)
    rc = lib().cm_ising_measure(ptr(spins), len(d), host_ptr(d, ctypes.c_int32), R, ptr(e),
                                ptr(m), stream())
    check(rc, "cm_ising_measure")
    return e, m


def pack(spins: torch.Tensor, shape, R: int) -> torch.Tensor:
    d = _dims(shape)
    out = torch.empty(R * packed_words(shape), dtype=torch.int32, device=spins.device)
    rc = lib().cm_ising_convert(ptr(spins), ptr(out), len(d), host_ptr(d, ctypes.c_int32), R, 0,
                                stream())
    check(rc, "cm_ising_convert")
    return out


def unpack(packed: torch.Tensor, shape, R: int) -> torch.Tensor:
    d = _dims(shape)
    out = torch.empty((R, int(np.prod(shape))), dtype=torch.int8, device=packed.device)
    rc = lib().cm_ising_convert(ptr(out), ptr(packed), len(d), host_ptr(d, ctypes.c_int32), R, 1,
                                stream())
    check(rc, "cm_ising_convert")
    return out


class Workspace:
    """Device scratch owned by the caller (class tables, one 256-byte row per chain).

    The tables are built and uploaded by ``cm_ising_prepare`` when the betas,
    field, lattice dimension, update codes or stream differ from the previous
    call that used this workspace; otherwise ``cm_ising_run`` reuses them with
    no host work.
    """

    def __init__(self, R: int, dev: torch.device | None = None):
        nbytes = int(lib().cm_ising_workspace_bytes(R))
        self.nbytes = nbytes
        self.buf = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=dev or device())
        self._key = None
        self._info = np.zeros(2, dtype=np.int32)

    def prepare(self, dims: np.ndarray, betas: np.ndarray, h: float, sched: np.ndarray):
        """The ``prepared`` record for these tables, uploading them first if needed."""
        st = stream()
        key = (len(dims), betas.tobytes(), float(h), tuple(sorted(set(sched[:, 0].tolist()))),
               st.value)
        if key != self._key:
            self._key = None
            rc = lib().cm_ising_prepare(len(dims), host_ptr(dims, ctypes.c_int32), len(betas),
                                        host_ptr(betas, ctypes.c_double), float(h),
                                        host_ptr(sched, ctypes.c_int64), sched.shape[0],
                                        ptr(self.buf), self.nbytes,
                                        host_ptr(self._info, ctypes.c_int32), st)
            check(rc, "cm_ising_prepare")
            self._key = key
        return host_ptr(self._info, ctypes.c_int32)


def run(spins: torch.Tensor, shape, betas, h: float, chains: Chains, schedule: np.ndarray,
        n_meas: int, e_run: torch.Tensor, m_run: torch.Tensor, *, layout: int = LAYOUT_INT8,
        form: int = FORM_AUTO, block: int = 0, grid: int = 0, ws: Workspace | None = None,
        out: tuple[torch.Tensor, torch.Tensor, torch.Tensor] | None = None,
        n_meas_rep: torch.Tensor | None = None):
    """Apply ``schedule`` ``n_meas`` times to every replica.

    Returns device int32 ``(E_b, M, accepted)`` series of shape ``[R, n_meas]``;
    ``e_run``/``m_run`` and the chains' sweep counters are advanced in place.
    ``n_meas_rep`` (int64 device ``[R]``, batched and wide forms) caps each replica's
    count: replica r performs ``min(n_meas, n_meas_rep[r])`` measurements.
    """
    R = chains.R
    betas = np.ascontiguousarray(np.broadcast_to(np.asarray(betas, dtype=np.float64), (R,)))
    sched = np.ascontiguousarray(schedule, dtype=np.int64).reshape(-1, 2)
    for code in sched[:, 0]:
        if code not in (METROPOLIS, HEATBATH):
            raise ValueError(f"update code {int(code)} has no checkerboard Ising kernel")
    d = _dims(shape)
    dev = spins.device
    if out is None:
        out = tuple(torch.empty((R, n_meas), dtype=torch.int32, device=dev) for _ in range(3))
    e_out, m_out, acc_out = out
    ws = ws or Workspace(R, dev)
    prepared = ws.prepare(d, betas, h, sched) if n_meas > 0 else None
    rc = lib().cm_ising_run(
        ptr(spins), layout, len(d), host_ptr(d, ctypes.c_int32), R,
        host_ptr(betas, ctypes.c_double), float(h), ptr(chains.keys_d), ptr(chains.sids_d),
        ptr(chains.sweeps_d), host_ptr(chains.sweeps_h, ctypes.c_uint64),
        host_ptr(sched, ctypes.c_int64), sched.shape[0], int(n_meas), ptr(e_run), ptr(m_run),
        ptr(e_out), ptr(m_out), ptr(acc_out), ptr(ws.buf), ws.nbytes, form, block, grid,
        ptr(n_meas_rep), prepared, stream(),
    )
    check(rc, "cm_ising_run")
    if n_meas_rep is None:
        chains.advance(int(sched[:, 1].sum()) * int(n_meas))
    return e_out, m_out, acc_out


class BoundRun:
    """One ``cm_ising_run`` call with every argument marshalled once.

    For loops that repeat the same call (benchmarks, fixed-shape slices):
    calling the object enqueues the run on the stream that was current at
    binding time and advances the chains' host counters, with no other host
    work.  The tables are prepared at binding time; the betas, field and
    schedule are fixed for the object's life.  Batched and wide forms only
    (the plain form reads the host counters per launch).
    """

    def __init__(self, spins: torch.Tensor, shape, betas, h: float, chains: Chains,
                 schedule: np.ndarray, n_meas: int, e_run: torch.Tensor, m_run: torch.Tensor,
                 out: tuple[torch.Tensor, torch.Tensor, torch.Tensor], *,
                 layout: int = LAYOUT_INT8, form: int = FORM_BATCHED, block: int = 0,
                 grid: int = 0, ws: Workspace | None = None,
                 n_meas_rep: torch.Tensor | None = None):
        if form not in (FORM_BATCHED, FORM_WIDE):
            raise ValueError("BoundRun needs the batched or the wide form")
        R = chains.R
        self._betas = np.ascontiguousarray(np.broadcast_to(np.asarray(betas, np.float64), (R,)))
        self._sched = np.ascontiguousarray(schedule, dtype=np.int64).reshape(-1, 2)
        self._dims = _dims(shape)
        if n_meas < 1:
            raise ValueError("n_meas must be positive")
        for t in out:
            if t.shape != (R, n_meas) or t.dtype != torch.int32:
                raise ValueError("out tensors must be int32 [R, n_meas]")
        self.ws = ws or Workspace(R, spins.device)
        self._keep = (spins, chains, e_run, m_run, out, n_meas_rep)
        self._chains = chains
        self._per_call = int(self._sched[:, 1].sum()) * int(n_meas) if n_meas_rep is None else 0
        prepared = self.ws.prepare(self._dims, self._betas, h, self._sched)
        self._fn = lib().cm_ising_run
        self._args = (
            ptr(spins), layout, len(self._dims), host_ptr(self._dims, ctypes.c_int32), R,
            host_ptr(self._betas, ctypes.c_double), float(h), ptr(chains.keys_d),
            ptr(chains.sids_d), ptr(chains.sweeps_d),
            host_ptr(chains.sweeps_h, ctypes.c_uint64), host_ptr(self._sched, ctypes.c_int64),
            self._sched.shape[0], int(n_meas), ptr(e_run), ptr(m_run), ptr(out[0]), ptr(out[1]),
            ptr(out[2]), ptr(self.ws.buf), self.ws.nbytes, form, block, grid, ptr(n_meas_rep),
            prepared, stream())

    def __call__(self) -> None:
        check(self._fn(*self._args), "cm_ising_run")
        if self._per_call:
            self._chains.advance(self._per_call)


class ClusterWorkspace:
    """Device scratch of :func:`run_cluster` (tables, parent/size/queue/mark arrays)."""

    def __init__(self, shape, R: int, dev: torch.device | None = None):
        d = _dims(shape)
        self.nbytes = int(lib().cm_ising_cluster_workspace_bytes(len(d), host_ptr(d, ctypes.c_int32), R))
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=dev or device())


def run_cluster(spins: torch.Tensor, shape, betas, h: float, chains: Chains, schedule,
                n_meas: int, e_run: torch.Tensor, m_run: torch.Tensor, *,
                ws: ClusterWorkspace | None = None):
    """Apply a schedule that may contain Wolff (3) and Swendsen-Wang (4) rows.

    Returns device ``(E_b, M)`` int32 ``[R, n_meas]`` and ``aux`` int64
    ``[R, n_meas, 4]`` (accepted, Wolff sizes, sum |C|^2, clusters).  Every
    repetition of every row advances the chains' sweep counters by one.
    """
    R = chains.R
    betas = np.ascontiguousarray(np.broadcast_to(np.asarray(betas, dtype=np.float64), (R,)))
    sched = np.ascontiguousarray(schedule, dtype=np.int64).reshape(-1, 2)
    for code in sched[:, 0]:
        if code not in (METROPOLIS, HEATBATH, WOLFF, SWENDSEN_WANG):
            raise ValueError(f"update code {int(code)} has no Ising GPU kernel")
    if h != 0.0 and np.isin(sched[:, 0], (WOLFF, SWENDSEN_WANG)).any():
        raise ValueError("cluster updates require zero field h")
    d = _dims(shape)
    dev = spins.device
    e_out = torch.empty((R, n_meas), dtype=torch.int32, device=dev)
    m_out = torch.empty((R, n_meas), dtype=torch.int32, device=dev)
    aux = torch.empty((R, n_meas, 4), dtype=torch.int64, device=dev)
    ws = ws or ClusterWorkspace(shape, R, dev)
    rc = lib().cm_ising_cluster_run(
        ptr(spins), len(d), host_ptr(d, ctypes.c_int32), R, host_ptr(betas, ctypes.c_double),
        float(h), ptr(chains.keys_d), ptr(chains.sids_d), ptr(chains.sweeps_d),
        host_ptr(chains.sweeps_h, ctypes.c_uint64), host_ptr(sched, ctypes.c_int64),
        sched.shape[0], int(n_meas), ptr(e_run), ptr(m_run), ptr(e_out), ptr(m_out), ptr(aux),
        ptr(ws.buf), ws.nbytes, stream())
    check(rc, "cm_ising_cluster_run")
    return e_out, m_out, aux
