"""Flat-histogram Monte Carlo on the device: Wang-Landau and multicanonical walkers.

Interface of critmc/flathist.py:36-545.  Wang-Landau (Wang & Landau 2001) walks
in energy space with acceptance ``min(1, g(E)/g(E'))`` and ``ln g(E) += ln f``
at every move; ``ln f`` is halved whenever the histogram is flat, and the 1/t
refinement (Belardinelli & Pereyra 2007) takes over once ``ln f`` falls below
``N_levels / t``.  A converged ``ln g`` weights a multicanonical production run
(Berg & Neuhaus 1992) whose samples reweight to any temperature.

The walks run on the GPU as many independent walkers, one lane each
(``cm_wl_run`` / ``cm_muca_run``): :class:`WangLandauBatch` and
:class:`MulticanonicalBatch` advance a batch of walkers (different seeds and
windows of one model) together, checking each walker's flatness on the host
between blocks of moves.  :class:`WangLandau` and :class:`Multicanonical` are
the reference's one-walker drivers on top of them.  A single walker is a
sequential chain, so the GPU pays off with many walkers.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch
from scipy.optimize import brentq
from scipy.special import logsumexp

from ._lib import check, lib
from .device import device, from_numpy_u32, host_ptr, ptr, stream
from .models import Ising, Model, Potts
from .rng import chain_key, make_rng

__all__ = ["EnergyLevels", "WangLandau", "WangLandauBatch", "WLResult", "Multicanonical",
           "MulticanonicalBatch", "MucaResult"]


@dataclass(frozen=True)
class EnergyLevels:
    """Level index <-> (bond) energy of a discrete model (critmc/flathist.py:36-93):
    Ising ``idx = (E_b + B) / 4``, Potts ``idx = E + B``."""

    kind: str
    n_sites: int
    n_bonds: int
    q: int = 2

    @classmethod
    def of(cls, model: Model) -> EnergyLevels:
        lat = model.lattice
        if not lat.hypercubic or lat.dim > 3:
            raise ValueError("GPU flat-histogram walkers support chain, square and cubic lattices")
        if isinstance(model, Ising):
            if model.h != 0.0:
                raise ValueError("flat-histogram sampling is implemented for h = 0")
            if lat.z % 2:
                raise ValueError("Ising energy levels require even coordination")
            return cls("ising", lat.n_sites, lat.n_bonds, 2)
        if isinstance(model, Potts):
            return cls("potts", lat.n_sites, lat.n_bonds, model.q)
        raise TypeError("flat-histogram methods support Ising and Potts models")

    @property
    def size(self) -> int:
        return self.n_bonds // 2 + 1 if self.kind == "ising" else self.n_bonds + 1

    @property
    def energies(self) -> np.ndarray:
        k = np.arange(self.size)
        return 4.0 * k - self.n_bonds if self.kind == "ising" else k - float(self.n_bonds)

    def _level(self, energy: float) -> float:
        return (energy + self.n_bonds) / 4 if self.kind == "ising" else energy + self.n_bonds

    def index(self, energy: float) -> int:
        return int(round(self._level(float(energy))))

    def window(self, e_min: float | None, e_max: float | None) -> tuple[int, int]:
        """Levels covering per-site energies ``[e_min, e_max]`` (``None`` = open end)."""
        n = self.n_sites
        lo = 0 if e_min is None else max(0, int(np.ceil(self._level(e_min * n))))
        hi = self.size - 1 if e_max is None else min(self.size - 1, int(np.floor(self._level(e_max * n))))
        if lo >= hi:
            raise ValueError("empty energy window")
        return lo, hi

    @property
    def log_total(self) -> float:
        """ln of the number of configurations, the normalisation of g."""
        return self.n_sites * np.log(self.q)


def _walk_into_window(model: Model, state: np.ndarray, levels: EnergyLevels, lo: int, hi: int,
                      rng: np.random.Generator, max_steps: int = 10 ** 7) -> int:
    """Greedy single-site walk until the level is inside ``[lo, hi]`` (critmc/flathist.py:100-129);
    host-side and drawn from the walker's numpy generator, before its Philox key."""
    nbr = model.lattice.nbr
    idx = levels.index(model.energy(state))
    n = model.n_sites
    for _ in range(max_steps):
        if lo <= idx <= hi:
            return idx
        i = int(rng.integers(n))
        if levels.kind == "ising":
            new = idx + int(state[i] * state[nbr[i]].sum()) // 2
            new_val = -state[i]
        else:
            new_val = int(rng.integers(levels.q - 1))
            new_val += new_val >= state[i]
            neigh = state[nbr[i]]
            new = idx + int(np.sum(neigh == state[i])) - int(np.sum(neigh == new_val))
        if (idx < lo and new >= idx) or (idx > hi and new <= idx):
            state[i] = new_val
            idx = new
    raise RuntimeError("could not reach the energy window")


@dataclass
class WLResult:
    """Outcome of one Wang-Landau walker (critmc/flathist.py:132-171)."""

    levels: EnergyLevels
    ln_g: np.ndarray
    visited: np.ndarray
    window: tuple[int, int]
    trials: int
    ln_f: float
    wall_time: float
    stages: list[dict] = field(default_factory=list)
    checkpoints: list[tuple[int, float, np.ndarray]] = field(default_factory=list)

    @property
    def energies(self) -> np.ndarray:
        return self.levels.energies

    def normalized(self, ln_g: np.ndarray | None = None) -> np.ndarray:
        """ln g normalised to the number of states (full window) or to 0 at the lowest level."""
        ln_g = self.ln_g if ln_g is None else ln_g
        out = np.full_like(ln_g, -np.inf)
        v = self.visited
        full = self.window == (0, self.levels.size - 1)
        shift = self.levels.log_total - logsumexp(ln_g[v]) if full else -ln_g[v][0]
        out[v] = ln_g[v] + shift
        return out

    def canonical(self, beta) -> dict[str, np.ndarray]:
        """ln Z, energy and specific heat per site from ln g (restricted to the window)."""
        beta = np.atleast_1d(np.asarray(beta, dtype=np.float64))
        v = self.visited
        e = self.energies[v]
        a = self.normalized()[v][None, :] - beta[:, None] * e[None, :]
        lz = logsumexp(a, axis=1)
        p = np.exp(a - lz[:, None])
        mean = p @ e
        var = p @ (e * e) - mean ** 2
        n = self.levels.n_sites
        return {"lnZ": lz, "u": mean / n, "c": beta ** 2 * var / n}


class _Walkers:
    """Device state of R walkers of one model: spins, level, window, Philox stream."""

    def __init__(self, model: Model, levels: EnergyLevels, states, idx, windows, keys):
        self.model, self.levels = model, levels
        self.dev = device()
        dv = self.dev
        self.R = len(states)
        self.q = 0 if levels.kind == "ising" else levels.q
        self.dims = np.ascontiguousarray(model.lattice.shape, dtype=np.int32)
        self.spins = torch.from_numpy(np.stack(states).astype(np.int8)).to(dv)
        self.idx = torch.tensor(np.asarray(idx, np.int32), device=dv)
        w = np.asarray(windows, np.int32).reshape(self.R, 2)
        self.lo = torch.from_numpy(w[:, 0].copy()).to(dv)
        self.hi = torch.from_numpy(w[:, 1].copy()).to(dv)
        self.keys_d = from_numpy_u32(np.asarray(keys, np.uint64).reshape(self.R, 2), dv)
        self.sids_d = torch.zeros(self.R, dtype=torch.int32, device=dv)
        self.moves = torch.zeros(self.R, dtype=torch.int64, device=dv)

    def wl(self, ln_g, hist, ln_f, mode, n_levels_1t, active, n_moves: int) -> None:
        dv = self.dev
        f_d = torch.from_numpy(np.asarray(ln_f, np.float64)).to(dv)
        m_d = torch.from_numpy(np.asarray(mode, np.int32)).to(dv)
        n_d = torch.from_numpy(np.asarray(n_levels_1t, np.float64)).to(dv)
        a_d = torch.from_numpy(np.asarray(active, np.int32)).to(dv)
        rc = lib().cm_wl_run(ptr(self.spins), len(self.dims), host_ptr(self.dims, ctypes.c_int32),
                             self.R, self.q, ptr(self.idx), ptr(ln_g), ptr(hist), self.levels.size,
                             ptr(self.lo), ptr(self.hi), ptr(self.keys_d), ptr(self.sids_d),
                             ptr(self.moves), ptr(f_d), ptr(m_d), ptr(n_d), ptr(a_d), int(n_moves),
                             stream())
        check(rc, "cm_wl_run")

    def muca(self, order, ln_w, hist, n_meas: int, moves_per_meas: int):
        dv = self.dev
        idx_out = torch.empty((self.R, n_meas), dtype=torch.int32, device=dv)
        ord_out = torch.empty((self.R, n_meas), dtype=torch.int64, device=dv)
        rc = lib().cm_muca_run(ptr(self.spins), len(self.dims), host_ptr(self.dims, ctypes.c_int32),
                               self.R, self.q, ptr(self.idx), ptr(order), ptr(ln_w), ptr(hist),
                               self.levels.size, ptr(self.lo), ptr(self.hi), ptr(self.keys_d),
                               ptr(self.sids_d), ptr(self.moves), int(n_meas), int(moves_per_meas),
                               ptr(idx_out), ptr(ord_out), stream())
        check(rc, "cm_muca_run")
        return idx_out, ord_out


def _prepare(model, levels, window, seed):
    """Initial state and level as the reference draws them, then the Philox key."""
    rng = make_rng(seed)
    state = model.initial_state("random", rng)
    idx = _walk_into_window(model, state, levels, *window, rng)
    return state, idx, chain_key(rng)


class WangLandauBatch:
    """Many Wang-Landau walkers of one zero-field Ising or Potts model, one lane each.

    ``seeds``: one seed per walker (each walker's stream follows from its seed
    as in :class:`WangLandau`); ``e_windows``: ``None``, one ``(e_min, e_max)``
    for all walkers, or one per walker.  ``flatness`` and ``criterion`` as in
    critmc/flathist.py:174-210.
    """

    def __init__(self, model: Model, seeds, *, e_windows=None, flatness: float = 0.8,
                 criterion: str = "flat"):
        if criterion not in ("flat", "visit"):
            raise ValueError("criterion must be 'flat' or 'visit'")
        self.model = model
        self.levels = EnergyLevels.of(model)
        seeds = list(seeds)
        R = len(seeds)
        if e_windows is None or (len(e_windows) == 2 and not isinstance(e_windows[0], (tuple, list))):
            e_windows = [e_windows] * R
        self.windows = [self.levels.window(*w) if w else (0, self.levels.size - 1) for w in e_windows]
        self.flatness = float(flatness)
        self.criterion = criterion
        prep = [_prepare(model, self.levels, w, s) for w, s in zip(self.windows, seeds)]
        self.walkers = _Walkers(model, self.levels, [p[0] for p in prep], [p[1] for p in prep],
                                self.windows, [p[2] for p in prep])

    def run(self, ln_f_final: float = 1e-6, *, one_over_t: bool = True, check_sweeps: int = 100,
            max_trials: float = 1e12, checkpoints=None) -> list[WLResult]:
        """Iterate every walker until ``ln f < ln_f_final`` (critmc/flathist.py:247-318)."""
        W = self.walkers
        R, S = W.R, self.levels.size
        dv = W.dev
        ln_g = torch.zeros((R, S), dtype=torch.float64, device=dv)
        hist = torch.zeros((R, S), dtype=torch.int64, device=dv)
        seen = np.zeros((R, S), dtype=bool)
        cps = list(np.asarray(sorted(checkpoints), dtype=np.int64)) if checkpoints is not None else []
        saved: list[list] = [[] for _ in range(R)]
        stages: list[list] = [[] for _ in range(R)]
        check_moves = max(1, int(check_sweeps) * self.model.n_sites)
        ln_f = np.ones(R)
        mode = np.zeros(R, np.int32)
        n_levels = np.zeros(R)
        t = 0
        active = np.ones(R, np.int32)
        done_t = np.zeros(R, np.int64)
        t0 = time.perf_counter()
        while active.any():
            block = check_moves
            if cps and cps[0] - t < block:
                block = int(cps[0] - t)
            if block > 0:
                W.wl(ln_g, hist, ln_f, mode, n_levels, active, block)
                t += block
            h_all = hist.cpu().numpy()
            seen |= h_all > 0
            on = active.astype(bool)
            ln_f[on & (mode == 1)] = n_levels[on & (mode == 1)] / t
            while cps and cps[0] <= t:
                g_all = ln_g.cpu().numpy()
                for r in np.flatnonzero(on):
                    saved[r].append((int(cps[0]), float(ln_f[r]), g_all[r].copy()))
                cps.pop(0)
            for r in np.flatnonzero(on & (mode == 0)):
                h = h_all[r][seen[r]]
                ok = h.min() > 0 if self.criterion == "visit" else h.min() >= self.flatness * h.mean()
                if ok:
                    stages[r].append({"ln_f": float(ln_f[r]), "trials": t})
                    ln_f[r] /= 2.0
                    hist[r].zero_()
                    if one_over_t and ln_f[r] < seen[r].sum() / t:
                        mode[r] = 1
                        n_levels[r] = float(seen[r].sum())
                        ln_f[r] = n_levels[r] / t
            for r in np.flatnonzero(on):
                if ln_f[r] < ln_f_final or t >= max_trials:
                    active[r] = 0
                    done_t[r] = t
        wall = time.perf_counter() - t0
        g_all = ln_g.cpu().numpy()
        out = []
        for r in range(R):
            for c in cps:
                saved[r].append((int(c), float(ln_f[r]), g_all[r].copy()))
            out.append(WLResult(self.levels, g_all[r].copy(), seen[r].copy(), self.windows[r],
                                int(done_t[r]), float(ln_f[r]), wall, stages[r], saved[r]))
        return out


class WangLandau:
    """The reference's one-walker Wang-Landau sampler (critmc/flathist.py:174-318)."""

    def __init__(self, model: Model, *, e_window=None, flatness: float = 0.8,
                 criterion: str = "flat", seed=None):
        self._batch = WangLandauBatch(model, [seed], e_windows=[e_window], flatness=flatness,
                                      criterion=criterion)
        self.model = model
        self.levels = self._batch.levels
        self.window = self._batch.windows[0]

    def run(self, ln_f_final: float = 1e-6, **kw) -> WLResult:
        return self._batch.run(ln_f_final, **kw)[0]


@dataclass
class MucaResult:
    """Multicanonical series plus the weights that produced them (critmc/flathist.py:321-452)."""

    levels: EnergyLevels
    ln_g: np.ndarray
    level_series: np.ndarray
    order_series: np.ndarray  # Ising: M; Potts: sum_k n_k^2
    window: tuple[int, int]
    hist: np.ndarray

    @property
    def energy(self) -> np.ndarray:
        return self.levels.energies[self.level_series]

    def order_parameter_m2(self) -> np.ndarray:
        n = self.levels.n_sites
        if self.levels.kind == "ising":
            return (self.order_series / n) ** 2
        q = self.levels.q
        return np.clip((q * self.order_series / n ** 2 - 1.0) / (q - 1.0), 0.0, None)

    def log_weights(self, beta: float) -> np.ndarray:
        """ln[exp(-beta E) g(E)] of each sample (unnormalised)."""
        return -beta * self.energy + self.ln_g[self.level_series]

    def average(self, beta: float, obs: np.ndarray) -> float:
        lw = self.log_weights(beta)
        w = np.exp(lw - lw.max())
        return float(np.sum(w * obs) / np.sum(w))

    def canonical(self, beta) -> dict[str, np.ndarray]:
        beta = np.atleast_1d(np.asarray(beta, dtype=np.float64))
        n = self.levels.n_sites
        e = self.energy
        m2 = self.order_parameter_m2()
        out = {k: np.empty_like(beta) for k in ("u", "c", "m2", "m4")}
        for i, b in enumerate(beta):
            lw = self.log_weights(b)
            w = np.exp(lw - lw.max())
            w /= w.sum()
            mean = w @ e
            out["u"][i] = mean / n
            out["c"][i] = b * b * (w @ (e * e) - mean ** 2) / n
            out["m2"][i] = w @ m2
            out["m4"][i] = w @ (m2 * m2)
        return out

    def energy_distribution(self, beta: float, smooth: float = 0.0) -> tuple[np.ndarray, np.ndarray]:
        """Canonical P_beta(e) over levels from the histogram of visits, optionally
        smoothed by a Gaussian of width ``smooth`` (energy per site)."""
        hist = self.hist.astype(np.float64)
        v = hist > 0
        if smooth > 0:
            sigma = smooth * self.levels.n_sites / (4.0 if self.levels.kind == "ising" else 1.0)
            half = int(np.ceil(3 * sigma))
            if half >= 1:
                kern = np.exp(-0.5 * (np.arange(-half, half + 1) / sigma) ** 2)
                num = np.convolve(np.where(v, hist, 0.0), kern, mode="same")
                den = np.convolve(v.astype(float), kern, mode="same")
                hist = np.where(v, num / np.maximum(den, 1e-300), 0.0)
        a = np.full(hist.shape, -np.inf)
        a[v] = np.log(hist[v]) + self.ln_g[v] - beta * self.levels.energies[v]
        p = np.exp(a - logsumexp(a[v]))
        return self.levels.energies / self.levels.n_sites, p

    def double_peak(self, beta: float, split: float | None = None,
                    smooth: float = 0.0) -> dict[str, float]:
        """Peak positions / heights of a bimodal P_beta(e) separated at ``split``."""
        e, p = self.energy_distribution(beta, smooth)
        keep = p > 0
        e, p = e[keep], p[keep]
        if split is None:
            split = float(np.sum(e * p))
        low, high = e < split, e >= split
        i_lo = np.argmax(np.where(low, p, -1))
        i_hi = np.argmax(np.where(high, p, -1))
        between = (e > e[i_lo]) & (e < e[i_hi])
        return {"e_ordered": float(e[i_lo]), "e_disordered": float(e[i_hi]),
                "p_ordered": float(p[i_lo]), "p_disordered": float(p[i_hi]),
                "p_min": float(p[between].min()) if between.any() else float("nan"),
                "weight_ordered": float(p[low].sum())}

    def phase_means(self, beta: float, split: float) -> dict[str, float]:
        """Mean energy per site of the samples below / above ``split`` at ``beta``."""
        e = self.energy / self.levels.n_sites
        lw = self.log_weights(beta)
        w = np.exp(lw - lw.max())
        low = e < split
        w_lo, w_hi = w[low].sum(), w[~low].sum()
        return {"e_ordered": float(w[low] @ e[low] / w_lo),
                "e_disordered": float(w[~low] @ e[~low] / w_hi),
                "weight_ordered": float(w_lo / (w_lo + w_hi))}

    def equal_height_beta(self, bracket, split: float | None = None, smooth: float = 0.0) -> float:
        def gap(b):
            dp = self.double_peak(b, split, smooth)
            return np.log(dp["p_ordered"]) - np.log(dp["p_disordered"])

        return float(brentq(gap, *bracket, xtol=1e-10))

    def equal_weight_beta(self, bracket, split: float) -> float:
        return float(brentq(lambda b: self.phase_means(b, split)["weight_ordered"] - 0.5,
                            *bracket, xtol=1e-10))


def _muca_weights(levels: EnergyLevels, ln_g, window) -> tuple[np.ndarray, np.ndarray]:
    """ln g with never-visited levels forbidden, and the weights -ln g (critmc/flathist.py:477-490)."""
    if isinstance(ln_g, WLResult):
        ln_g = np.where(ln_g.visited, ln_g.ln_g, -np.inf)
    ln_g = np.asarray(ln_g, dtype=np.float64).copy()
    if len(ln_g) != levels.size:
        raise ValueError("ln_g does not match the model's energy levels")
    lo, hi = window
    inside = ln_g[lo:hi + 1]
    inside[~np.isfinite(inside)] = np.inf
    return ln_g, np.where(np.isfinite(ln_g), -ln_g, -np.inf)


class MulticanonicalBatch:
    """Multicanonical production of many walkers (one ``ln_g`` each, or one shared)."""

    def __init__(self, model: Model, ln_g, seeds, *, e_windows=None):
        self.model = model
        self.levels = EnergyLevels.of(model)
        seeds = list(seeds)
        R = len(seeds)
        if e_windows is None or (len(e_windows) == 2 and not isinstance(e_windows[0], (tuple, list))):
            e_windows = [e_windows] * R
        self.windows = [self.levels.window(*w) if w else (0, self.levels.size - 1) for w in e_windows]
        gs = ln_g if isinstance(ln_g, (list, tuple)) else [ln_g] * R
        pairs = [_muca_weights(self.levels, g, w) for g, w in zip(gs, self.windows)]
        self.ln_g = [p[0] for p in pairs]
        prep = [_prepare(model, self.levels, w, s) for w, s in zip(self.windows, seeds)]
        self.walkers = _Walkers(model, self.levels, [p[0] for p in prep], [p[1] for p in prep],
                                self.windows, [p[2] for p in prep])
        dv = self.walkers.dev
        self.ln_w = torch.from_numpy(np.stack([p[1] for p in pairs])).to(dv)
        if self.levels.kind == "ising":
            order = np.array([[int(p[0].astype(np.int64).sum())] for p in prep], np.int32)
        else:
            order = np.stack([np.bincount(p[0], minlength=self.levels.q) for p in prep]).astype(np.int32)
        self.order = torch.from_numpy(order).to(dv)
        self.hist = torch.zeros((R, self.levels.size), dtype=torch.int64, device=dv)

    def run(self, n_meas: int, sweeps_per_meas: int = 1, n_therm: int = 0) -> list[MucaResult]:
        W = self.walkers
        moves = int(sweeps_per_meas) * self.model.n_sites
        if n_therm:
            W.muca(self.order, self.ln_w, torch.zeros_like(self.hist), n_therm, moves)
        idx_out, ord_out = W.muca(self.order, self.ln_w, self.hist, n_meas, moves)
        idx_h, ord_h, hist = idx_out.cpu().numpy(), ord_out.cpu().numpy(), self.hist.cpu().numpy()
        return [MucaResult(self.levels, self.ln_g[r], idx_h[r].astype(np.int64), ord_h[r],
                           self.windows[r], hist[r].copy()) for r in range(W.R)]


class Multicanonical:
    """The reference's one-walker multicanonical driver (critmc/flathist.py:455-545)."""

    def __init__(self, model: Model, ln_g, *, e_window=None, seed=None):
        self._batch = MulticanonicalBatch(model, ln_g, [seed], e_windows=[e_window])
        self.model = model
        self.levels = self._batch.levels
        self.window = self._batch.windows[0]
        self.ln_g = self._batch.ln_g[0]

    def run(self, n_meas: int, sweeps_per_meas: int = 1, n_therm: int = 0) -> MucaResult:
        return self._batch.run(n_meas, sweeps_per_meas, n_therm)[0]
