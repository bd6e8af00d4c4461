"""Checkerboard O(n) kernels (n = 2 XY, n = 3 Heisenberg): host side of ``cm_on_*``.

Replaces critmc/kernels/on.py's ``run`` (:348-418) for Metropolis (:112-137)
and over-relaxation (:140-183) sweeps; spins are FP32 ``[R, N, n]`` on the
device, measurements (after an FP64 renormalisation) are FP64.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from .._lib import check, lib
from ..device import device, host_ptr, ptr, stream
from .chains import Chains
from .common import CLUSTER_CODES, HEATBATH, METROPOLIS, OVERRELAX, SWENDSEN_WANG, WOLFF


def _dims(shape) -> np.ndarray:
    return np.ascontiguousarray(shape, dtype=np.int32)


class Workspace:
    def __init__(self, shape, R: int, n: int, dev: torch.device | None = None):
        d = _dims(shape)
        self.nbytes = int(lib().cm_on_workspace_bytes(len(d), host_ptr(d, ctypes.c_int32), R, n))
        self.buf = torch.empty(max(self.nbytes, 256), dtype=torch.uint8, device=dev or device())


class ClusterWorkspace:
    """Device scratch for Wolff / Swendsen-Wang rows (parent, size, queue, mark, projections)."""

    def __init__(self, shape, R: int, dev: torch.device | None = None):
        d = _dims(shape)
        self.nbytes = int(lib().cm_on_cluster_workspace_bytes(len(d), host_ptr(d, ctypes.c_int32), R))
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=dev or device())


def to_split(sites: torch.Tensor, shape, n: int) -> torch.Tensor:
    """Site-ordered ``[R, N, n]`` spins -> the device's colour-split layout (same shape)."""
    d = _dims(shape)
    out = torch.empty_like(sites)
    check(lib().cm_on_convert(ptr(sites.contiguous()), ptr(out), len(d), host_ptr(d, ctypes.c_int32),
                              sites.shape[0], int(n), 0, stream()), "cm_on_convert")
    return out


def to_sites(split: torch.Tensor, shape, n: int) -> torch.Tensor:
    """Colour-split device spins -> site order ``[R, N, n]`` (x fastest)."""
    d = _dims(shape)
    out = torch.empty_like(split)
    check(lib().cm_on_convert(ptr(split.contiguous()), ptr(out), len(d), host_ptr(d, ctypes.c_int32),
                              split.shape[0], int(n), 1, stream()), "cm_on_convert")
    return out


def init_spins(spins: torch.Tensor, shape, n: int, chains: Chains) -> None:
    """Random unit vectors (XY angle uniform in [0, 2 pi), Heisenberg isotropic), written
    in the colour-split layout."""
    d = _dims(shape)
    check(lib().cm_on_init(ptr(spins), len(d), host_ptr(d, ctypes.c_int32), chains.R, int(n),
                           ptr(chains.keys_d), ptr(chains.sids_d), stream()), "cm_on_init")


def _per_chain_f32(x, R: int, dv: torch.device) -> torch.Tensor:
    """A per-chain FP32 device vector: a device tensor as given, or host values
    copied from pinned memory without waiting for the stream."""
    if isinstance(x, torch.Tensor) and x.device == dv:
        if x.dtype != torch.float32 or x.numel() != R or not x.is_contiguous():
            raise ValueError("per-chain device values must be contiguous float32 of length R")
        return x
    h = torch.from_numpy(np.broadcast_to(np.asarray(x, np.float32), (R,)).copy())
    return h.pin_memory().to(dv, non_blocking=True)


def run(spins: torch.Tensor, shape, n: int, betas, steps, chains: Chains, schedule, n_meas: int,
        *, ws: Workspace | None = None, n_meas_rep: torch.Tensor | None = None,
        cluster_ws: ClusterWorkspace | None = None, with_aux: bool = False):
    """``betas`` and ``steps``: per-chain values (host, or float32 device tensors of
    length R).  Returns device tensors (accepted int32 [R, n_meas], E f64 [R, n_meas],
    M f64 [R, n_meas, n], stiff_c / stiff_s f64 [R, n_meas, d]), plus ``aux``
    int64 [R, n_meas, 4] (accepted, Wolff sizes, sum |C|^2, clusters) when
    ``with_aux``."""
    R = chains.R
    sched = np.ascontiguousarray(schedule, dtype=np.int64).reshape(-1, 2)
    for code in sched[:, 0]:
        if code not in (METROPOLIS, HEATBATH, OVERRELAX, WOLFF, SWENDSEN_WANG):
            raise ValueError(f"update code {int(code)} has no O(n) GPU kernel")
    if np.isin(sched[:, 0], CLUSTER_CODES).any() and cluster_ws is None:
        cluster_ws = ClusterWorkspace(shape, R, spins.device)
    dv = spins.device
    d = _dims(shape)
    b = _per_chain_f32(betas, R, dv)
    st = _per_chain_f32(steps, R, dv)
    acc = torch.empty((R, n_meas), dtype=torch.int32, device=dv)
    e = torch.empty((R, n_meas), dtype=torch.float64, device=dv)
    m = torch.empty((R, n_meas, n), dtype=torch.float64, device=dv)
    sc = torch.empty((R, n_meas, len(d)), dtype=torch.float64, device=dv)
    ss = torch.empty((R, n_meas, len(d)), dtype=torch.float64, device=dv)
    ws = ws or Workspace(shape, R, n, dv)
    aux = torch.empty((R, n_meas, 4), dtype=torch.int64, device=dv) if with_aux else None
    rc = lib().cm_on_run(ptr(spins), len(d), host_ptr(d, ctypes.c_int32), R, int(n), ptr(b),
                         ptr(st), ptr(chains.keys_d), ptr(chains.sids_d), ptr(chains.sweeps_d),
                         host_ptr(sched, ctypes.c_int64), sched.shape[0], int(n_meas), ptr(acc),
                         ptr(e), ptr(m), ptr(sc), ptr(ss), ptr(ws.buf), ws.nbytes,
                         ptr(n_meas_rep), ptr(aux), ptr(cluster_ws.buf) if cluster_ws else None,
                         cluster_ws.nbytes if cluster_ws else 0, stream())
    check(rc, "cm_on_run")
    if n_meas_rep is None:
        chains.advance(int(sched[:, 1].sum()) * int(n_meas))
    if with_aux:
        return acc, e, m, sc, ss, aux
    return acc, e, m, sc, ss
