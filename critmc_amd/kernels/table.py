"""Chains on any lattice: host side of ``cm_ising_table_run`` / ``cm_potts_table_run``
and ``cm_on_table_run``.

For lattices the checkerboard kernels cannot sweep (odd extents, the
triangular lattice, or a schedule with ``metropolis-random``), the schedule
runs in the reference's own site orders (critmc/kernels/ising.py:68-107,
critmc/kernels/potts.py:44-103): typewriter Metropolis and heat bath and
random-site Metropolis, one lane per chain, with the Wolff and Swendsen-Wang
kernels reading neighbours from the lattice's table (critmc/lattice.py:129-138).
"""

from __future__ import annotations

import ctypes
import weakref

import numpy as np
import torch

from .._lib import check, lib
from ..device import device, host_ptr, ptr, stream
from ..lattice import Lattice
from .chains import Chains

Z_MAX = 8
_TABLES: "weakref.WeakKeyDictionary[Lattice, dict]" = weakref.WeakKeyDictionary()


def check_lattice(lattice: Lattice) -> None:
    """The table path needs a pair-form neighbour table (column 2a = +e_a,
    2a+1 = -e_a) with an even number of columns up to 8."""
    nbr = np.asarray(lattice.nbr)
    n, z = nbr.shape
    if z < 2 or z > Z_MAX or z % 2:
        raise ValueError(f"GPU chains need 2..{Z_MAX} neighbour columns in +/- pairs, got {z}")
    if n >= 1 << 28:
        raise ValueError("lattice too large")
    rows = np.arange(n)
    for a in range(z // 2):
        if not np.array_equal(nbr[nbr[:, 2 * a], 2 * a + 1], rows):
            raise ValueError("lattice neighbour table is not in forward/backward pair form")


def neighbours(lattice: Lattice, dev: torch.device | None = None) -> torch.Tensor:
    """The lattice's neighbour table as a device int32 ``[N, z]`` tensor (cached)."""
    dev = dev or device()
    per_lattice = _TABLES.setdefault(lattice, {})
    if dev not in per_lattice:
        check_lattice(lattice)
        per_lattice[dev] = torch.from_numpy(
            np.ascontiguousarray(lattice.nbr, dtype=np.int32)).to(dev)
    return per_lattice[dev]


class Workspace:
    """Device scratch: per-chain tables and the cluster arrays."""

    def __init__(self, n_sites: int, R: int, q: int = 0, dev: torch.device | None = None):
        self.nbytes = int(lib().cm_table_workspace_bytes(int(n_sites), int(R), int(q)))
        self.buf = torch.empty(max(self.nbytes, 256), dtype=torch.uint8, device=dev or device())


def _schedule(schedule) -> np.ndarray:
    return np.ascontiguousarray(schedule, dtype=np.int64).reshape(-1, 2)


def ising_run(spins: torch.Tensor, nbr: torch.Tensor, betas, h: float, chains: Chains, schedule,
              n_meas: int, e_run: torch.Tensor, m_run: torch.Tensor, *,
              ws: Workspace | None = None):
    """Apply ``schedule`` ``n_meas`` times to int8 spins ``[R, N]``.

    Returns device ``(E_b, M)`` int32 ``[R, n_meas]`` and ``aux`` int64
    ``[R, n_meas, 4]``; ``e_run``, ``m_run`` and the sweep counters advance.
    """
    R = chains.R
    N, z = nbr.shape
    betas = np.ascontiguousarray(np.broadcast_to(np.asarray(betas, np.float64), (R,)))
    sched = _schedule(schedule)
    dv = spins.device
    e_out = torch.empty((R, n_meas), dtype=torch.int32, device=dv)
    m_out = torch.empty((R, n_meas), dtype=torch.int32, device=dv)
    aux = torch.empty((R, n_meas, 4), dtype=torch.int64, device=dv)
    ws = ws or Workspace(N, R, 0, dv)
    rc = lib().cm_ising_table_run(
        ptr(spins), N, z, ptr(nbr), R, host_ptr(betas, ctypes.c_double), float(h),
        ptr(chains.keys_d), ptr(chains.sids_d), ptr(chains.sweeps_d),
        host_ptr(sched, ctypes.c_int64), sched.shape[0], int(n_meas), ptr(e_run), ptr(m_run),
        ptr(e_out), ptr(m_out), ptr(aux), ptr(ws.buf), ws.nbytes, stream())
    check(rc, "cm_ising_table_run")
    chains.advance(int(sched[:, 1].sum()) * int(n_meas))
    return e_out, m_out, aux


def potts_run(spins: torch.Tensor, nbr: torch.Tensor, q: int, betas, chains: Chains, schedule,
              n_meas: int, e_run: torch.Tensor, counts_run: torch.Tensor, *,
              ws: Workspace | None = None):
    """Potts states ``[R, N]`` (0..q-1): returns device ``E`` int32 ``[R, n_meas]``,
    ``counts`` int32 ``[R, n_meas, q]`` and ``aux`` int64 ``[R, n_meas, 4]``."""
    R = chains.R
    N, z = nbr.shape
    betas = np.ascontiguousarray(np.broadcast_to(np.asarray(betas, np.float64), (R,)))
    sched = _schedule(schedule)
    dv = spins.device
    e_out = torch.empty((R, n_meas), dtype=torch.int32, device=dv)
    c_out = torch.empty((R, n_meas, q), dtype=torch.int32, device=dv)
    aux = torch.empty((R, n_meas, 4), dtype=torch.int64, device=dv)
    ws = ws or Workspace(N, R, q, dv)
    rc = lib().cm_potts_table_run(
        ptr(spins), N, z, ptr(nbr), R, int(q), host_ptr(betas, ctypes.c_double),
        ptr(chains.keys_d), ptr(chains.sids_d), ptr(chains.sweeps_d),
        host_ptr(sched, ctypes.c_int64), sched.shape[0], int(n_meas), ptr(e_run),
        ptr(counts_run), ptr(e_out), ptr(c_out), ptr(aux), ptr(ws.buf), ws.nbytes, stream())
    check(rc, "cm_potts_table_run")
    chains.advance(int(sched[:, 1].sum()) * int(n_meas))
    return e_out, c_out, aux


def measure(spins: torch.Tensor, nbr: torch.Tensor, R: int, q: int = 0):
    """Exact energy and (Ising) magnetisation or (Potts) occupation counts per chain."""
    N, z = nbr.shape
    e = torch.empty(R, dtype=torch.int32, device=spins.device)
    mc = torch.empty((R, max(q, 1)), dtype=torch.int32, device=spins.device)
    check(lib().cm_table_measure(ptr(spins), N, z, ptr(nbr), R, int(q), ptr(e), ptr(mc), stream()),
          "cm_table_measure")
    return e, (mc[:, 0] if q == 0 else mc)


class OnWorkspace:
    """Device scratch of the O(n) neighbour-table kernel: per-chain β and step,
    union-find / queue arrays and projections for cluster rows."""

    def __init__(self, n_sites: int, R: int, dev: torch.device | None = None):
        self.nbytes = int(lib().cm_on_table_workspace_bytes(int(n_sites), int(R)))
        self.buf = torch.empty(max(self.nbytes, 256), dtype=torch.uint8, device=dev or device())


def on_run(spins: torch.Tensor, nbr: torch.Tensor, n: int, betas, steps, chains: Chains, schedule,
           n_meas: int, *, ws: OnWorkspace | None = None):
    """Apply ``schedule`` ``n_meas`` times to FP64 O(n) spins ``[R, N, n]`` (site order),
    one lane per chain, in the reference's own site orders (critmc/kernels/on.py:112-261,
    348-418).  Returns device FP64 ``E [R, n_meas]``, ``M [R, n_meas, n]``, stiffness sums
    ``c, s [R, n_meas, z/2]`` and ``aux`` int64 ``[R, n_meas, 4]``."""
    R = chains.R
    N, z = nbr.shape
    if spins.dtype != torch.float64 or tuple(spins.shape) != (R, N, n):
        raise ValueError(f"spins must be float64 [{R}, {N}, {n}]")
    betas = np.ascontiguousarray(np.broadcast_to(np.asarray(betas, np.float64), (R,)))
    steps = np.ascontiguousarray(np.broadcast_to(np.asarray(steps, np.float64), (R,)))
    sched = _schedule(schedule)
    dv = spins.device
    A = z // 2
    e_out = torch.empty((R, n_meas), dtype=torch.float64, device=dv)
    m_out = torch.empty((R, n_meas, n), dtype=torch.float64, device=dv)
    c_out = torch.empty((R, n_meas, A), dtype=torch.float64, device=dv)
    s_out = torch.empty((R, n_meas, A), dtype=torch.float64, device=dv)
    aux = torch.empty((R, n_meas, 4), dtype=torch.int64, device=dv)
    ws = ws or OnWorkspace(N, R, dv)
    rc = lib().cm_on_table_run(
        ptr(spins), N, z, ptr(nbr), R, int(n), host_ptr(betas, ctypes.c_double),
        host_ptr(steps, ctypes.c_double), ptr(chains.keys_d), ptr(chains.sids_d),
        ptr(chains.sweeps_d), host_ptr(sched, ctypes.c_int64), sched.shape[0], int(n_meas),
        ptr(e_out), ptr(m_out), ptr(c_out), ptr(s_out), ptr(aux), ptr(ws.buf), ws.nbytes, stream())
    check(rc, "cm_on_table_run")
    chains.advance(int(sched[:, 1].sum()) * int(n_meas))
    return e_out, m_out, c_out, s_out, aux
