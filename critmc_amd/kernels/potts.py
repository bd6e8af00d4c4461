"""Checkerboard Potts kernels: host side of ``cm_potts_*``.

Replaces critmc/kernels/potts.py's ``run`` (:170-235) for the Metropolis
(:44-71) and heat-bath (:74-103) sweeps; states are int8 ``[R, N]`` on the device.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from .._lib import check, lib
from ..device import device, host_ptr, ptr, stream
from .chains import Chains
from .common import HEATBATH, METROPOLIS, SWENDSEN_WANG, WOLFF

START_CODES = {"random": 0, "ordered": 1, "up": 1}


def _dims(shape) -> np.ndarray:
    return np.ascontiguousarray(shape, dtype=np.int32)


class Workspace:
    """Device scratch owned by the caller: the chains' tables (thresholds and
    heat-bath weights), uploaded by ``cm_potts_prepare`` when the betas, the
    lattice dimension or the stream differ from the previous call that used
    this workspace, and otherwise reused with no host work."""

    def __init__(self, R: int, dev: torch.device | None = None):
        self.nbytes = int(lib().cm_potts_workspace_bytes(R))
        self.buf = torch.empty(max(self.nbytes, 256), dtype=torch.uint8, device=dev or device())
        self._key = None

    def prepare(self, dims: np.ndarray, betas: np.ndarray) -> None:
        st = stream()
        key = (len(dims), betas.tobytes(), st.value)
        if key != self._key:
            self._key = None
            check(lib().cm_potts_prepare(len(dims), host_ptr(dims, ctypes.c_int32), len(betas),
                                         host_ptr(betas, ctypes.c_double), ptr(self.buf), self.nbytes,
                                         st), "cm_potts_prepare")
            self._key = key


def init_states(spins: torch.Tensor, shape, q: int, chains: Chains, start: str = "random") -> None:
    if start not in START_CODES:
        raise ValueError(f"unknown start {start!r}")
    d = _dims(shape)
    check(lib().cm_potts_init(ptr(spins), len(d), host_ptr(d, ctypes.c_int32), chains.R, int(q),
                              ptr(chains.keys_d), ptr(chains.sids_d), START_CODES[start], stream()),
          "cm_potts_init")


def measure(spins: torch.Tensor, shape, q: int, R: int) -> tuple[torch.Tensor, torch.Tensor]:
    """E = -#satisfied bonds and occupation counts per replica (potts.py:27-41)."""
    d = _dims(shape)
    e = torch.empty(R, dtype=torch.int32, device=spins.device)
    counts = torch.empty((R, q), dtype=torch.int32, device=spins.device)
    check(lib().cm_potts_measure(ptr(spins), len(d), host_ptr(d, ctypes.c_int32), R, int(q),
                                 ptr(e), ptr(counts), stream()), "cm_potts_measure")
    return e, counts


def run(spins: torch.Tensor, shape, q: int, betas, chains: Chains, schedule, n_meas: int,
        e_run: torch.Tensor, counts_run: torch.Tensor, *, block: int = 0, grid: int = 0,
        ws: Workspace | None = None, n_meas_rep: torch.Tensor | None = None):
    """Returns device int32 ``(E [R, n_meas], accepted [R, n_meas], counts [R, n_meas, q])``."""
    R = chains.R
    betas = np.ascontiguousarray(np.broadcast_to(np.asarray(betas, dtype=np.float64), (R,)))
    sched = np.ascontiguousarray(schedule, dtype=np.int64).reshape(-1, 2)
    for code in sched[:, 0]:
        if code not in (METROPOLIS, HEATBATH):
            raise ValueError(f"update code {int(code)} has no checkerboard Potts kernel")
    d = _dims(shape)
    dev = spins.device
    e_out = torch.empty((R, n_meas), dtype=torch.int32, device=dev)
    acc = torch.empty((R, n_meas), dtype=torch.int32, device=dev)
    counts = torch.empty((R, n_meas, q), dtype=torch.int32, device=dev)
    if ws is None:  # one-shot scratch: the run uploads the tables itself
        ws = Workspace(R, dev)
        tabs = host_ptr(betas, ctypes.c_double)
    else:  # a caller's workspace keeps its tables across calls with the same betas
        if n_meas > 0:
            ws.prepare(d, betas)
        tabs = None
    rc = lib().cm_potts_run(
        ptr(spins), len(d), host_ptr(d, ctypes.c_int32), R, int(q),
        tabs, ptr(chains.keys_d), ptr(chains.sids_d),
        ptr(chains.sweeps_d), host_ptr(sched, ctypes.c_int64), sched.shape[0], int(n_meas),
        ptr(e_run), ptr(counts_run), ptr(e_out), ptr(acc), ptr(counts), ptr(ws.buf), ws.nbytes,
        block, grid, ptr(n_meas_rep), stream())
    check(rc, "cm_potts_run")
    if n_meas_rep is None:
        chains.advance(int(sched[:, 1].sum()) * int(n_meas))
    return e_out, acc, counts


class ClusterWorkspace:
    """Device scratch of :func:`run_cluster`."""

    def __init__(self, shape, R: int, q: int, dev: torch.device | None = None):
        d = _dims(shape)
        self.nbytes = int(lib().cm_potts_cluster_workspace_bytes(len(d), host_ptr(d, ctypes.c_int32),
                                                                 R, int(q)))
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=dev or device())


def run_cluster(spins: torch.Tensor, shape, q: int, betas, chains: Chains, schedule, n_meas: int,
                e_run: torch.Tensor, counts_run: torch.Tensor, *,
                ws: ClusterWorkspace | None = None):
    """Apply a schedule that may contain Wolff (3) and Swendsen-Wang (4) rows
    (critmc/kernels/potts.py:106-167).  Returns device ``(E [R, n_meas],
    counts [R, n_meas, q])`` int32 and ``aux`` int64 ``[R, n_meas, 4]``."""
    R = chains.R
    betas = np.ascontiguousarray(np.broadcast_to(np.asarray(betas, dtype=np.float64), (R,)))
    sched = np.ascontiguousarray(schedule, dtype=np.int64).reshape(-1, 2)
    for code in sched[:, 0]:
        if code not in (METROPOLIS, HEATBATH, WOLFF, SWENDSEN_WANG):
            raise ValueError(f"update code {int(code)} has no Potts GPU kernel")
    d = _dims(shape)
    dev = spins.device
    e_out = torch.empty((R, n_meas), dtype=torch.int32, device=dev)
    counts = torch.empty((R, n_meas, q), dtype=torch.int32, device=dev)
    aux = torch.empty((R, n_meas, 4), dtype=torch.int64, device=dev)
    ws = ws or ClusterWorkspace(shape, R, q, dev)
    rc = lib().cm_potts_cluster_run(
        ptr(spins), len(d), host_ptr(d, ctypes.c_int32), R, int(q), host_ptr(betas, ctypes.c_double),
        ptr(chains.keys_d), ptr(chains.sids_d), ptr(chains.sweeps_d), host_ptr(sched, ctypes.c_int64),
        sched.shape[0], int(n_meas), ptr(e_run), ptr(counts_run), ptr(e_out), ptr(counts), ptr(aux),
        ptr(ws.buf), ws.nbytes, stream())
    check(rc, "cm_potts_cluster_run")
    chains.advance(int(sched[:, 1].sum()) * int(n_meas))
    return e_out, counts, aux
