"""Canonical Markov-chain driver on the GPU.

The interface is the reference's (critmc/sim.py:48-297):

>>> from critmc_amd import Simulation, Ising, square
>>> sim = Simulation(Ising(square(64)), beta=0.4406868, updates="metropolis", seed=42)
>>> sim.thermalize(10_000)
>>> ts = sim.run(100_000)
>>> ts["e"].mean(), ts["absm"].mean()

The chain's state lives on the GPU between calls; ``sim.state`` copies it
back.  On bipartite chain / square / cubic lattices sweeps are checkerboard
(colour 0 then colour 1); on other lattices, and for schedules with
``metropolis-random``, the chain runs in the reference's own site orders
(typewriter sweeps, random sites).  Every draw is a counter-based Philox
number keyed by the chain, so a run split into chunks of any size reproduces
the same series (critmc tests/test_sim_chunks.py).

Schedules are ``+``-separated update names with optional ``k*`` repetition
counts.  A ``wolff`` term without a count is retuned during thermalisation
towards N / <|C|> clusters per measurement, and the O(n) Metropolis step
towards 50 % acceptance; production runs use the tuned values unchanged.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np
import torch

from .io import load_npz, save_npz
from .kernels.chains import Chains
from .kernels.common import (
    AUX_ACCEPTED,
    AUX_WOLFF_SIZE,
    HEATBATH,
    METROPOLIS,
    METROPOLIS_RANDOM,
    OVERRELAX,
    SWENDSEN_WANG,
    UPDATE_CODES,
    WOLFF,
)
from .models import Model, derive
from .rng import chain_key, make_rng

__all__ = ["TimeSeries", "Simulation", "parse_updates"]

SINGLE_SITE = (METROPOLIS, METROPOLIS_RANDOM)
FULL_SWEEPS = (METROPOLIS, METROPOLIS_RANDOM, HEATBATH, OVERRELAX, SWENDSEN_WANG)


@dataclass
class TimeSeries:
    """Measurement streams of one chain and its metadata.

    ``ts[name]`` is a raw stream (``"E"``, ``"M"``, ``"counts"`` ...) or a
    per-site observable derived from them on first use (``"e"``, ``"absm"``,
    ``"m2"`` ...; see :func:`critmc_amd.models.derive`).
    """

    raw: dict[str, np.ndarray]
    meta: dict[str, Any]
    _observables: dict[str, np.ndarray] | None = field(default=None, repr=False)

    def _derived(self) -> dict[str, np.ndarray]:
        if self._observables is None:
            self._observables = derive(self.raw, self.meta)
        return self._observables

    def __len__(self) -> int:
        return int(len(self.raw["E"]))

    def __getitem__(self, key: str) -> np.ndarray:
        stream = self.raw.get(key)
        if stream is not None:
            return stream
        derived = self._derived()
        if key not in derived:
            raise KeyError(f"no observable {key!r}; have {self.keys()}")
        return derived[key]

    def keys(self) -> list[str]:
        return sorted(self.raw.keys() | self._derived().keys())

    @property
    def beta(self) -> float:
        return float(self.meta["beta"])

    @property
    def n_sites(self) -> int:
        return int(self.meta["N"])

    @property
    def L(self) -> int:
        return int(self.meta["L"])

    def slice(self, start: int | None = None, stop: int | None = None) -> TimeSeries:
        """Measurements ``start:stop`` with a copy of the metadata."""
        window = slice(start, stop)
        return TimeSeries({name: s[window] for name, s in self.raw.items()}, dict(self.meta))

    @staticmethod
    def concatenate(parts: list[TimeSeries]) -> TimeSeries:
        """Parts of one chain joined in order; ``n_meas`` counts them all."""
        if not parts:
            raise ValueError("nothing to concatenate")
        first = parts[0]
        raw = {name: np.concatenate([p.raw[name] for p in parts]) for name in first.raw}
        return TimeSeries(raw, {**first.meta, "n_meas": sum(map(len, parts))})

    def save(self, path: str | Path) -> Path:
        return save_npz(path, self.raw, self.meta)

    @classmethod
    def load(cls, path: str | Path) -> TimeSeries:
        return cls(*load_npz(path))


def _schedule_term(token: str) -> tuple[str, int | None]:
    """``"4*overrelax"`` -> ("overrelax", 4); ``"wolff"`` -> ("wolff", None)."""
    head, star, name = token.partition("*")
    if not star:
        return token, None
    if not (head.isdecimal() and int(head) >= 1):
        raise ValueError(f"repetition count must be a positive integer in {token!r}")
    return name, int(head)


def parse_updates(updates: str) -> list[tuple[str, int | None]]:
    """``"heatbath+4*overrelax+wolff"`` -> ``[(name, count or None), ...]``.

    Spaces and case are ignored; the errors are those of critmc/sim.py:112-131.
    """
    cleaned = updates.replace(" ", "").lower()
    rows = [_schedule_term(token) for token in cleaned.split("+") if token]
    if not rows:
        raise ValueError("empty update schedule")
    unknown = [name for name, _ in rows if name not in UPDATE_CODES]
    if unknown:
        raise ValueError(f"unknown updates {unknown}; choose from {sorted(UPDATE_CODES)}")
    return rows


class Simulation:
    """One Markov chain at inverse temperature ``beta`` (critmc/sim.py:134-297).

    ``seed`` is an int, ``SeedSequence`` or ``Generator`` as in the reference.
    The initial state is drawn from it on the host (``start`` "random" or
    "ordered", or an explicit ``state``, which is copied), then the chain's
    Philox key is drawn from the same generator.
    """

    def __init__(self, model: Model, beta: float, updates: str = "metropolis", *,
                 seed: Any = None, start: str = "random", state: np.ndarray | None = None):
        rows = parse_updates(updates)
        model.check_updates([name for name, _ in rows])
        self.model = model
        self.beta = float(beta)
        self.updates = updates
        self.seed = seed if isinstance(seed, (int, np.integer)) else None
        self.rng = make_rng(seed)
        self.schedule = np.array([[UPDATE_CODES[name], count or 1] for name, count in rows],
                                 dtype=np.int64)
        # rows whose Wolff cluster count is retuned while thermalising
        self._auto_wolff = [i for i, (name, count) in enumerate(rows)
                            if name == "wolff" and count is None]
        self._mean_cluster = float("nan")
        self._wolff_history: list[int] = []
        if state is None:
            host, self.start = model.initial_state(start, self.rng), start
        else:
            host, self.start = np.array(state, copy=True), "given"
        self.chains = Chains([chain_key(self.rng)])
        self.running = model.begin(host)
        self._ds = model.to_device(host, self.running)
        self.ws = model.workspace()
        self.n_therm = 0
        self.n_meas = 0

    @property
    def state(self) -> np.ndarray:
        """Host copy of the current configuration."""
        return self.model.from_device(self._ds)

    def _count(self, *codes: int) -> int:
        """Updates of these codes per measurement."""
        return int(self.schedule[np.isin(self.schedule[:, 0], codes), 1].sum())

    def _advance(self, n_meas: int) -> dict[str, np.ndarray]:
        return self.model.advance(self._ds, self.running, self.beta, self.chains, self.schedule,
                                  n_meas, self.ws)

    # -- thermalisation ---------------------------------------------------------
    def _retune_wolff(self, wolff_sizes: np.ndarray) -> None:
        """Mean cluster size from the later (better equilibrated) half of a chunk;
        automatic Wolff rows move towards N / <|C|> clusters per measurement, by
        at most a factor 4 per chunk since a hot start has tiny clusters."""
        per_meas = self._count(WOLFF)
        late = wolff_sizes[len(wolff_sizes) // 2:]
        self._mean_cluster = float(late.sum() / (per_meas * len(late)))
        if not self._auto_wolff or not self._mean_cluster > 0:
            return
        goal = max(1, int(round(self.model.n_sites / self._mean_cluster)))
        for row in self._auto_wolff:
            now = int(self.schedule[row, 1])
            self.schedule[row, 1] = min(max(goal, max(1, now // 4)), 4 * now)
        self._wolff_history.append(int(self.schedule[self._auto_wolff[0], 1]))

    def _retune_step(self, accepted: np.ndarray, per_meas: int) -> None:
        """O(n) Metropolis step scaled by the acceptance over 1/2, within [1/2, 2] per
        chunk and [1e-3, 10] overall."""
        rate = accepted.sum() / (per_meas * self.model.n_sites * len(accepted))
        scaled = self.running["step"] * float(np.clip(rate / 0.5, 0.5, 2.0))
        self.running["step"] = float(np.clip(scaled, 1e-3, 10.0))

    def thermalize(self, n_meas: int, *, tune: bool = True, n_chunks: int = 10) -> dict[str, Any]:
        """Discard ``n_meas`` measurements in up to ``n_chunks`` chunks, retuning the
        automatic Wolff count and the O(n) step after each (critmc/sim.py:193-231)."""
        if n_meas < 0:
            raise ValueError("n_meas must not be negative")
        edges = np.linspace(0, n_meas, max(1, min(n_chunks, n_meas)) + 1).astype(int)
        single = self._count(*SINGLE_SITE)
        for size in np.diff(edges):
            if size == 0:
                continue
            aux = self._advance(int(size))["aux"]
            if not tune:
                continue
            if self._count(WOLFF):
                self._retune_wolff(aux[:, AUX_WOLFF_SIZE])
            if single and "step" in self.running:
                self._retune_step(aux[:, AUX_ACCEPTED], single)
        self.n_therm += int(n_meas)
        return self.tuning()

    def tuning(self) -> dict[str, Any]:
        """Current schedule and tuned parameters."""
        info: dict[str, Any] = {"schedule": self.schedule.tolist()}
        wolff = self._count(WOLFF)
        if wolff:
            info.update(wolff_per_meas=wolff, mean_cluster=self._mean_cluster)
            if self._wolff_history:
                info["wolff_history"] = list(self._wolff_history)
        if "step" in self.running:
            info["step"] = self.running["step"]
        return info

    # -- production -------------------------------------------------------------
    def run(self, n_meas: int, *, chunk: int = 1 << 16) -> TimeSeries:
        """Record ``n_meas`` measurements, advanced ``chunk`` at a time."""
        if n_meas < 1:
            raise ValueError("n_meas must be positive")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        parts = [self._advance(min(chunk, n_meas - done)) for done in range(0, n_meas, chunk)]
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        raw = parts[0] if len(parts) == 1 else {
            name: np.concatenate([p[name] for p in parts]) for name in parts[0]}
        self.n_meas += n_meas
        return TimeSeries(raw, self._run_meta(raw["aux"], n_meas, wall))

    def _run_meta(self, aux: np.ndarray, n_meas: int, wall: float) -> dict[str, Any]:
        """Run metadata with the reference's keys (critmc/sim.py:279-296).  Wolff
        clusters count as sweeps by their share of the lattice (:271-277)."""
        n = self.model.n_sites
        wolff = self._count(WOLFF)
        single = self._count(*SINGLE_SITE)
        mean_cluster = aux[:, AUX_WOLFF_SIZE].sum() / (wolff * n_meas) if wolff else float("nan")
        sweeps = self._count(*FULL_SWEEPS) + (wolff * mean_cluster / n if wolff else 0.0)
        acceptance = aux[:, AUX_ACCEPTED].sum() / (single * n * n_meas) if single else None
        return {
            **self.model.meta(),
            "beta": self.beta,
            "updates": self.updates,
            "schedule": self.schedule.tolist(),
            "seed": self.seed,
            "start": self.start,
            "n_therm": self.n_therm,
            "n_meas": n_meas,
            "sweeps_per_meas": sweeps,
            "mean_cluster": float(mean_cluster),
            "acceptance": None if acceptance is None else float(acceptance),
            "step": self.running.get("step"),
            "wall_time": wall,
            "site_updates_per_s": sweeps * n * n_meas / wall if wall > 0 else None,
            "backend": "hip",
            "sweep_order": self.model.sweep_order(self.schedule),
        }
