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


def _normalised(log_w: np.ndarray) -> np.ndarray:
    """Weights proportional to exp(log_w), summing to one."""
    w = np.exp(log_w - np.max(log_w))
    return w / w.sum()


@dataclass(frozen=True)
class EnergyLevels:
    """Level index <-> (bond) energy of a discrete zero-field model (critmc/flathist.py:36-93).

    Ising bond energies run from -B to B in steps of 4 (even coordination),
    Potts energies (minus the satisfied bonds) from -B to 0 in steps of 1, so
    level k has energy ``k * step - B`` with B the number of bonds.
    """

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
    def _step(self) -> int:
        return 4 if self.kind == "ising" else 1

    @property
    def size(self) -> int:
        span = 2 * self.n_bonds if self.kind == "ising" else self.n_bonds
        return span // self._step + 1

    @property
    def energies(self) -> np.ndarray:
        return self._step * np.arange(self.size) - float(self.n_bonds)

    def _position(self, energy: float) -> float:
        """Fractional level of an energy."""
        return (energy + self.n_bonds) / self._step

    def index(self, energy: float) -> int:
        return int(round(self._position(float(energy))))

    def window(self, e_min: float | None, e_max: float | None) -> tuple[int, int]:
        """Levels covering per-site energies ``[e_min, e_max]`` (``None`` = open end)."""
        n = self.n_sites
        lo = 0 if e_min is None else int(np.ceil(self._position(e_min * n)))
        hi = self.size - 1 if e_max is None else int(np.floor(self._position(e_max * n)))
        lo, hi = max(lo, 0), min(hi, self.size - 1)
        if lo >= hi:
            raise ValueError("empty energy window")
        return lo, hi

    @property
    def log_total(self) -> float:
        """ln of the number of configurations, q^N: the normalisation of g."""
        return self.n_sites * np.log(self.q)


def _walk_into_window(model: Model, state: np.ndarray, levels: EnergyLevels, lo: int, hi: int,
                      rng: np.random.Generator, max_steps: int = 10 ** 7) -> int:
    """Bring ``state`` (in place) into levels ``[lo, hi]`` and return its level.

    Greedy single-site moves (critmc/flathist.py:100-129): a random site and,
    for Potts, a random other state are drawn from the walker's numpy
    generator, and the move is made unless it leads away from the window.
    Runs on the host, before the walker's Philox key is drawn.
    """
    nbr = model.lattice.nbr
    ising = levels.kind == "ising"
    idx = levels.index(model.energy(state))
    for _ in range(max_steps):
        if lo <= idx <= hi:
            return idx
        site = int(rng.integers(model.n_sites))
        old, neigh = state[site], state[nbr[site]]
        if ising:
            new = -old
            change = int(old * neigh.sum()) // 2  # bond energy +2 s_i h_i, in levels of 4
        else:
            new = int(rng.integers(levels.q - 1))
            new += new >= old
            change = int(np.count_nonzero(neigh == old)) - int(np.count_nonzero(neigh == new))
        towards = 1 if idx < lo else -1
        if towards * change >= 0:
            state[site] = new
            idx += change
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
        """ln g on the visited levels (-inf elsewhere), fixed so that g sums to the
        number of configurations when the walk covered every level, otherwise so
        that the lowest visited level has ln g = 0."""
        values = np.asarray(self.ln_g if ln_g is None else ln_g)
        seen = self.visited
        if self.window == (0, self.levels.size - 1):
            offset = self.levels.log_total - logsumexp(values[seen])
        else:
            offset = -values[seen][0]
        return np.where(seen, values + offset, -np.inf)

    def canonical(self, beta) -> dict[str, np.ndarray]:
        """ln Z, energy and specific heat per site from ln g over the visited levels."""
        beta = np.atleast_1d(np.asarray(beta, dtype=np.float64))
        seen = self.visited
        e = self.energies[seen]
        log_p = self.normalized()[seen] - np.outer(beta, e)
        ln_z = logsumexp(log_p, axis=1)
        p = np.exp(log_p - ln_z[:, None])
        first, second = p @ e, p @ (e * e)
        n = self.levels.n_sites
        return {"lnZ": ln_z, "u": first / n, "c": beta * beta * (second - first * first) / n}


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


def _smoothed_visits(hist: np.ndarray, seen: np.ndarray, sigma: float) -> np.ndarray:
    """Visit counts averaged over visited neighbouring levels with a Gaussian of
    ``sigma`` levels (cut at 3 sigma); unvisited levels stay 0."""
    reach = int(np.ceil(3 * sigma))
    if reach < 1:
        return hist
    kernel = np.exp(-0.5 * (np.arange(-reach, reach + 1) / sigma) ** 2)
    mass = np.convolve(seen.astype(np.float64), kernel, mode="same")
    total = np.convolve(np.where(seen, hist, 0.0), kernel, mode="same")
    return np.where(seen, total / np.maximum(mass, 1e-300), 0.0)


@dataclass
class MucaResult:
    """Multicanonical series and the ln g that weighted them (critmc/flathist.py:321-452).

    A sample at level E was drawn with weight 1/g(E), so canonical averages at
    beta weight it by exp(-beta E) g(E).
    """

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
        """Squared order parameter per site of every sample: (M/N)^2 for Ising,
        (q sum_k rho_k^2 - 1) / (q - 1), clipped at 0, for Potts."""
        n, q = self.levels.n_sites, self.levels.q
        if self.levels.kind == "ising":
            return (self.order_series / n) ** 2
        return np.maximum((q * self.order_series / (n * n) - 1.0) / (q - 1.0), 0.0)

    def log_weights(self, beta: float) -> np.ndarray:
        """ln[exp(-beta E) g(E)] of each sample (unnormalised)."""
        return self.ln_g[self.level_series] - beta * self.energy

    def average(self, beta: float, obs: np.ndarray) -> float:
        return float(np.dot(_normalised(self.log_weights(beta)), obs))

    def canonical(self, beta) -> dict[str, np.ndarray]:
        """u, c per site and <m^2>, <m^4> at each beta."""
        beta = np.atleast_1d(np.asarray(beta, dtype=np.float64))
        w = np.stack([_normalised(self.log_weights(b)) for b in beta])
        e, m2 = self.energy, self.order_parameter_m2()
        n = self.levels.n_sites
        first, second = w @ e, w @ (e * e)
        return {"u": first / n, "c": beta * beta * (second - first * first) / n,
                "m2": w @ m2, "m4": w @ (m2 * m2)}

    def energy_distribution(self, beta: float, smooth: float = 0.0) -> tuple[np.ndarray, np.ndarray]:
        """Per-site level energies and the canonical P_beta over levels, from the visit
        histogram; ``smooth`` > 0 first averages the visits over a Gaussian of that
        width in energy per site, which steadies peak positions and valley depths."""
        hist = self.hist.astype(np.float64)
        seen = hist > 0
        if smooth > 0:
            per_level = self.levels.n_sites / (4.0 if self.levels.kind == "ising" else 1.0)
            hist = _smoothed_visits(hist, seen, smooth * per_level)
        log_p = np.full(hist.shape, -np.inf)
        log_p[seen] = np.log(hist[seen]) + self.ln_g[seen] - beta * self.levels.energies[seen]
        p = np.exp(log_p - logsumexp(log_p[seen]))
        return self.levels.energies / self.levels.n_sites, p

    def double_peak(self, beta: float, split: float | None = None,
                    smooth: float = 0.0) -> dict[str, float]:
        """The largest P_beta(e) below and at or above ``split`` (default: the mean
        energy), their positions, the lowest P between them and the weight below."""
        e, p = self.energy_distribution(beta, smooth)
        e, p = e[p > 0], p[p > 0]
        cut = float(np.dot(e, p)) if split is None else split
        below = e < cut
        lo = int(np.argmax(np.where(below, p, -1.0)))
        hi = int(np.argmax(np.where(below, -1.0, p)))
        valley = p[(e > e[lo]) & (e < e[hi])]
        return {"e_ordered": float(e[lo]), "e_disordered": float(e[hi]),
                "p_ordered": float(p[lo]), "p_disordered": float(p[hi]),
                "p_min": float(valley.min()) if valley.size else float("nan"),
                "weight_ordered": float(p[below].sum())}

    def phase_means(self, beta: float, split: float) -> dict[str, float]:
        """Mean energy per site of the samples below and at or above ``split`` at beta,
        and the weight of those below."""
        e = self.energy / self.levels.n_sites
        w = _normalised(self.log_weights(beta))
        below = e < split
        w_below, w_above = w[below].sum(), w[~below].sum()
        return {"e_ordered": float(np.dot(w[below], e[below]) / w_below),
                "e_disordered": float(np.dot(w[~below], e[~below]) / w_above),
                "weight_ordered": float(w_below / (w_below + w_above))}

    def equal_height_beta(self, bracket, split: float | None = None, smooth: float = 0.0) -> float:
        """Beta in ``bracket`` where both peaks of P_beta(e) are equally high."""
        def log_ratio(b):
            peaks = self.double_peak(b, split, smooth)
            return np.log(peaks["p_ordered"] / peaks["p_disordered"])

        return float(brentq(log_ratio, *bracket, xtol=1e-10))

    def equal_weight_beta(self, bracket, split: float) -> float:
        """Beta in ``bracket`` where both phases carry probability 1/2."""
        return float(brentq(lambda b: self.phase_means(b, split)["weight_ordered"] - 0.5,
                            *bracket, xtol=1e-10))


def _muca_weights(levels: EnergyLevels, ln_g, window) -> tuple[np.ndarray, np.ndarray]:
    """The ln g a walker uses and its weights -ln g (critmc/flathist.py:477-490).

    From a :class:`WLResult`, levels its walk never visited get ln g = -inf.
    A non-finite ln g inside the window becomes +inf (weight -inf: forbidden);
    outside the window, -inf entries keep weight -inf as well.
    """
    if isinstance(ln_g, WLResult):
        ln_g = np.where(ln_g.visited, ln_g.ln_g, -np.inf)
    table = np.array(ln_g, dtype=np.float64)
    if table.shape != (levels.size,):
        raise ValueError("ln_g does not match the model's energy levels")
    lo, hi = window
    inside = np.zeros(levels.size, dtype=bool)
    inside[lo:hi + 1] = True
    table[inside & ~np.isfinite(table)] = np.inf
    return table, np.where(np.isfinite(table), -table, -np.inf)


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
