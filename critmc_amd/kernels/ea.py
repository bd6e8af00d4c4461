"""Batched Edwards-Anderson parallel tempering on the device (``cm_ea_*``).

Replaces critmc/kernels/glass.py's ``pt_run`` (:80-163) for many disorder
samples at once, of possibly different linear sizes.  All samples of a batch
share the dimension (square or cubic) and the number of temperatures ``m``.
A batch whose couplings are all +-1 runs on int8 couplings with exact int32
energies; any other couplings (Gaussian disorder) run in FP64.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from .._lib import check, lib
from ..device import device, from_numpy_u32, host_ptr, ptr, stream
from .common import HEATBATH, METROPOLIS

N_OBS = 7
OBS_E, OBS_QL, OBS_Q, OBS_Q2, OBS_Q4, OBS_QK2, OBS_COUNT = range(N_OBS)


class EaBatch:
    """Device state of S samples x 2 replica sets x m temperatures.

    ``couplings[s]``: sample s's forward-bond couplings, axis-major
    (``bonds[a*N + i] = (i, i + e_a)``, critmc/lattice.py:136-138); +-1 values
    select the integer path (``gaussian`` False), others the FP64 one.
    ``betas``: ``[S, m]`` or ``[m]``, hot -> cold.  ``keys``: ``[S, 2]`` Philox keys.
    ``spins``: optional host ``[2, m, N_s]`` int8 start per sample (default:
    device draw from the sample's key).
    """

    def __init__(self, Ls, ndim: int, couplings, betas, keys, spins=None,
                 dev: torch.device | None = None):
        self.dev = dev or device()
        self.L = np.ascontiguousarray(Ls, dtype=np.int32)
        self.S = len(self.L)
        self.d = int(ndim)
        if self.d not in (2, 3):
            raise ValueError("EA tempering supports square and cubic lattices")
        if np.any(self.L < 2) or np.any(self.L % 2):
            raise ValueError("checkerboard sweeps need even extents >= 2")
        self.N = self.L.astype(np.int64) ** self.d
        betas = np.asarray(betas, dtype=np.float64)
        if betas.ndim == 1:
            betas = np.tile(betas, (self.S, 1))
        self.betas = np.ascontiguousarray(betas)
        self.m = self.betas.shape[1]
        if self.m < 2:
            raise ValueError("need at least two temperatures")
        self.spin_off = np.concatenate([[0], np.cumsum(2 * self.m * self.N)[:-1]]).astype(np.int64)
        self.j_off = np.concatenate([[0], np.cumsum(self.d * self.N)[:-1]]).astype(np.int64)
        couplings = [np.asarray(c, dtype=np.float64) for c in couplings]
        if len(couplings) != self.S:
            raise ValueError("need the couplings of every sample")
        for s, c in enumerate(couplings):
            if c.shape != (self.d * self.N[s],):
                raise ValueError("need one coupling per bond")
            if not np.all(np.isfinite(c)):
                raise ValueError("couplings must be finite")
        self.gaussian = not all(np.all(np.abs(c) == 1.0) for c in couplings)
        self.jtype = 1 if self.gaussian else 0
        J = np.empty(int((self.d * self.N).sum()), dtype=np.float64 if self.gaussian else np.int8)
        for s, c in enumerate(couplings):
            J[self.j_off[s]:self.j_off[s] + c.size] = c
        dv = self.dev
        self.J = torch.from_numpy(J).to(dv)
        self.L_d = torch.from_numpy(self.L).to(dv)
        self.spin_off_d = torch.from_numpy(self.spin_off).to(dv)
        self.j_off_d = torch.from_numpy(self.j_off).to(dv)
        self.keys = np.asarray(keys, dtype=np.uint64).reshape(self.S, 2)
        self.keys_d = from_numpy_u32(self.keys, dv)
        self.betas_d = torch.from_numpy(self.betas).to(dv)
        total = int((2 * self.m * self.N).sum())
        self.spins = torch.empty(total, dtype=torch.int8, device=dv)
        if spins is None:
            check(lib().cm_ea_init(ptr(self.spins), ptr(self.L_d), ptr(self.spin_off_d), self.S,
                                   self.m, self.d, ptr(self.keys_d), stream()), "cm_ea_init")
        else:
            host = np.concatenate([np.asarray(sp, dtype=np.int8).reshape(-1) for sp in spins])
            if host.size != total:
                raise ValueError("spins do not match the samples' sizes")
            self.spins.copy_(torch.from_numpy(host))
        S2m = self.S * 2 * self.m
        i32 = dict(dtype=torch.int32, device=dv)
        self.sweeps = torch.zeros(S2m, dtype=torch.int64, device=dv)
        self.energies = torch.empty(S2m, dtype=torch.float64 if self.gaussian else torch.int32,
                                    device=dv)
        ar = torch.arange(self.m, **i32).repeat(self.S * 2)
        self.perm = ar.clone()
        self.kofc = ar.clone()
        self.labels = torch.zeros(S2m, **i32)
        self.trips = torch.zeros(self.S * 2, **i32)
        self.label_hist = torch.zeros(S2m * 2, **i32)
        self.swap_acc = torch.zeros(self.S * 2 * (self.m - 1), **i32)
        self.swap_try = torch.zeros(self.S * 2 * (self.m - 1), **i32)
        nbytes = int(lib().cm_ea_workspace_bytes(self.S, self.m, int(self.N.sum())))
        self.ws_bytes = nbytes
        self.ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=dv)
        self.step = 0
        self.recompute_energies()

    @property
    def multispin(self) -> bool:
        """Whether ``run`` takes this batch multi-spin coded, which fixes its uniform
        mapping (one bit per configuration, planes shared by a set)."""
        return bool(lib().cm_ea_multispin(self.jtype, self.m, self.d, int(self.L.max())))

    def recompute_energies(self) -> None:
        check(lib().cm_ea_energy(ptr(self.spins), ptr(self.J), self.jtype, ptr(self.L_d),
                                 ptr(self.spin_off_d), ptr(self.j_off_d), self.S, self.m, self.d,
                                 ptr(self.energies), stream()), "cm_ea_energy")

    def configurations(self, s: int) -> np.ndarray:
        """Host copy of sample s's spins ``[2, m, N]``."""
        o, n = int(self.spin_off[s]), int(self.N[s])
        return self.spins[o:o + 2 * self.m * n].cpu().numpy().reshape(2, self.m, n)

    def run(self, n_steps: int, n_bins: int, *, update: str = "metropolis",
            sweeps_per_step: int = 1, record: bool = False, block: int = 0):
        """Advance every sample ``n_steps`` PT steps; fresh accumulators and counters per call."""
        code = {"metropolis": METROPOLIS, "heatbath": HEATBATH}.get(update)
        if code is None:
            raise ValueError(f"update {update!r} has no checkerboard EA kernel")
        dv = self.dev
        acc = torch.zeros(self.S * n_bins * self.m * N_OBS, dtype=torch.float64, device=dv)
        for t in (self.trips, self.label_hist, self.swap_acc, self.swap_try):
            t.zero_()
        e_series = q_series = None
        if record:
            e_series = torch.empty(self.S * 2 * self.m * n_steps, device=dv,
                                   dtype=torch.float64 if self.gaussian else torch.int32)
            q_series = torch.empty(self.S * self.m * n_steps, dtype=torch.int32, device=dv)
        rc = lib().cm_ea_pt_run(
            ptr(self.spins), ptr(self.J), self.jtype, ptr(self.L_d), ptr(self.spin_off_d),
            ptr(self.j_off_d), self.S, self.m, self.d, host_ptr(self.L, ctypes.c_int32),
            host_ptr(self.betas, ctypes.c_double), ptr(self.betas_d), ptr(self.keys_d),
            ptr(self.sweeps), ptr(self.energies), ptr(self.perm), ptr(self.kofc), ptr(self.labels),
            ptr(self.trips), ptr(self.label_hist), ptr(self.swap_acc), ptr(self.swap_try),
            ptr(acc), int(n_bins), ptr(e_series), ptr(q_series), code, int(sweeps_per_step),
            int(self.step), int(n_steps), ptr(self.ws), self.ws_bytes, block, stream())
        check(rc, "cm_ea_pt_run")
        self.step += int(n_steps)
        out = {
            "acc": acc.reshape(self.S, n_bins, self.m, N_OBS),
            "swap_acc": self.swap_acc.reshape(self.S, 2, self.m - 1),
            "swap_try": self.swap_try.reshape(self.S, 2, self.m - 1),
            "label_hist": self.label_hist.reshape(self.S, 2, self.m, 2),
            "trips": self.trips.reshape(self.S, 2),
        }
        if record:
            out["E"] = e_series.reshape(self.S, 2, self.m, n_steps)
            out["q"] = q_series.reshape(self.S, self.m, n_steps)
        return out
