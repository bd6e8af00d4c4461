"""Batches of chains with a resumable on-disk cache (critmc/runner.py:34-223).

Same interface and cache semantics as the reference: a :class:`Task` is a
picklable description of one chain whose stream is ``(seed, *keys)``; with
``out_dir`` every result is written to ``<out_dir>/<filename>`` through a
``.part.npz`` + ``os.replace``, existing files are reused only when they record
the same task (:class:`StaleCacheError` otherwise), and interrupted partial
files are removed.

What changes is where the chains run: one GPU runs them all.  Tasks that share
model, schedule and run lengths (and differ in beta, seed or keys) advance
together in one batched kernel launch per phase; each task's series is
bit-identical to running it alone through :func:`run_task`, because every
chain keeps its own Philox key and sweep counter.  ``workers`` is accepted for
compatibility and only sizes :func:`parallel_map`'s process pool.
"""

from __future__ import annotations

import os
import sys
import time
from collections.abc import Callable, Iterable, Sequence
from concurrent.futures import ProcessPoolExecutor, as_completed
from dataclasses import asdict, dataclass, field
from pathlib import Path
from typing import Any

import numpy as np
import torch

from .io import to_jsonable
from .kernels import ising as k_ising
from .kernels.chains import Chains
from .kernels.common import N_AUX, UPDATE_CODES
from .lattice import checkerboard_ok
from .models import Ising, build_model
from .rng import chain_key, make_rng
from .sim import Simulation, TimeSeries, parse_updates

__all__ = ["Task", "StaleCacheError", "run_task", "run_tasks", "parallel_map", "default_workers"]


@dataclass(frozen=True)
class Task:
    """One canonical Markov chain; ``model`` is a spec such as
    ``{"model": "ising", "lattice": "square", "L": 32}``."""

    model: dict[str, Any]
    beta: float
    updates: str
    n_therm: int
    n_meas: int
    seed: int
    keys: tuple = ()
    start: str = "random"
    name: str = ""
    extra: dict[str, Any] = field(default_factory=dict)

    @property
    def filename(self) -> str:
        if self.name:
            return f"{self.name}.npz"
        tag = "_".join(str(k) for k in self.keys) or "run"
        return f"{self.model['model']}_L{self.model['L']}_b{self.beta:.8f}_{tag}.npz"


class StaleCacheError(ValueError):
    """A cached result file was produced by a different task than the one requested."""


def _task_record(task: Task) -> dict[str, Any]:
    """The task as its JSON metadata record reads back from a result file."""
    return to_jsonable(asdict(task))


def _record_mismatch(found: Any, wanted: dict[str, Any]) -> str:
    """Why a file's task record is not the requested one, field by field."""
    if not isinstance(found, dict):
        return "it does not record the task that produced it"
    return "; ".join(f"{name}: file has {found.get(name)!r}, requested {value!r}"
                     for name, value in wanted.items() if found.get(name) != value)


def _load_cached(task: Task, path: Path) -> TimeSeries:
    """A cached result, after checking that ``task`` produced it (critmc/runner.py:59-87)."""
    ts = TimeSeries.load(path)
    wanted = _task_record(task)
    found = ts.meta.get("task")
    if found != wanted:
        raise StaleCacheError(f"{path} was not produced by the requested task "
                              f"({_record_mismatch(found, wanted)}). "
                              "Use a different out_dir or delete the file.")
    return ts


def run_task(task: Task) -> TimeSeries:
    """Run one task on the GPU through :class:`Simulation` (critmc/runner.py:90-98); its
    stream is ``(seed, *keys)`` and its metadata records the task and the tuning."""
    sim = Simulation(build_model(task.model), task.beta, task.updates,
                     seed=make_rng(task.seed, *task.keys), start=task.start)
    sim.thermalize(task.n_therm)
    ts = sim.run(task.n_meas)
    ts.meta["task"] = asdict(task)
    ts.meta["tuning"] = sim.tuning()
    return ts


def _batchable(task: Task) -> bool:
    """Ising tasks with checkerboard rows only, on a lattice those kernels sweep."""
    if task.model.get("model") != "ising":
        return False
    try:
        rows = parse_updates(task.updates)
        lattice = build_model(task.model).lattice
    except ValueError:
        return False
    return checkerboard_ok(lattice) and all(name in ("metropolis", "heatbath") for name, _ in rows)


def _group_key(task: Task):
    return (tuple(sorted(to_jsonable(task.model).items())), task.updates, task.n_therm,
            task.n_meas, task.start)


def _run_ising_group(tasks: list[Task]) -> list[TimeSeries]:
    """Ising tasks of one group in a single batched launch per phase.

    Each chain is set up exactly as :class:`Simulation` does (host initial
    state from the task's generator, then the Philox key from the same
    generator), so its series equals :func:`run_task`'s bit for bit.
    """
    model = build_model(tasks[0].model)
    assert isinstance(model, Ising)
    lat = model.lattice
    rows = parse_updates(tasks[0].updates)
    for name, _ in rows:
        model.check_updates([name])
    sched = np.array([[UPDATE_CODES[n], 1 if k is None else k] for n, k in rows], dtype=np.int64)
    states, keys = [], []
    for t in tasks:
        rng = make_rng(t.seed, *t.keys)
        states.append(model.initial_state(t.start, rng))
        keys.append(chain_key(rng))
    dev = torch.device("cuda", torch.cuda.current_device())
    R = len(tasks)
    s = torch.from_numpy(np.stack(states).astype(np.int8)).to(dev)
    chains = Chains(keys, np.zeros(R, np.uint32), None, dev)
    betas = np.array([t.beta for t in tasks], dtype=np.float64)
    e, m = k_ising.measure(s, lat.shape, R)
    ws = k_ising.Workspace(R, dev)
    n_therm, n_meas = tasks[0].n_therm, tasks[0].n_meas
    if n_therm:
        k_ising.run(s, lat.shape, betas, model.h, chains, sched, n_therm, e, m, ws=ws)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e_out, m_out, acc = k_ising.run(s, lat.shape, betas, model.h, chains, sched, n_meas, e, m, ws=ws)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    e_out = e_out.cpu().numpy().astype(np.int64)
    m_out = m_out.cpu().numpy().astype(np.int64)
    acc = acc.cpu().numpy().astype(np.int64)
    n = model.n_sites
    sweeps = int(sched[:, 1].sum())
    out = []
    for r, t in enumerate(tasks):
        aux = np.zeros((n_meas, N_AUX), dtype=np.int64)
        aux[:, 0] = acc[r]
        eb, mm = e_out[r], m_out[r]
        raw = {"E": eb - model.h * mm if model.h != 0.0 else eb, "Eb": eb, "M": mm, "aux": aux}
        meta = {**model.meta(), "beta": t.beta, "updates": t.updates, "schedule": sched.tolist(),
                "seed": None, "start": t.start, "n_therm": n_therm, "n_meas": n_meas,
                "sweeps_per_meas": sweeps, "mean_cluster": float("nan"),
                "acceptance": float(acc[r].sum() / (sweeps * n * n_meas)) if sweeps else None,
                "step": None, "wall_time": wall / R, "site_updates_per_s": sweeps * n * n_meas * R / wall,
                "backend": "hip", "sweep_order": "checkerboard", "batched_with": R,
                "task": asdict(t), "tuning": {"schedule": sched.tolist()}}
        out.append(TimeSeries(raw, meta))
    return out


def _save(ts: TimeSeries, out_dir: str, task: Task) -> str:
    path = Path(out_dir) / task.filename
    tmp = path.with_suffix(".part.npz")
    ts.save(tmp)
    os.replace(tmp, path)
    return str(path)


def default_workers() -> int:
    return max(1, (os.cpu_count() or 2) - 1)


def run_tasks(tasks: Sequence[Task], *, workers: int | None = None, out_dir=None,
              progress: bool = True, batch: bool = True) -> list[TimeSeries]:
    """Run tasks on the GPU and return their time series in task order.

    With ``out_dir`` results are cached on disk and reused only if produced by
    the same task; otherwise :class:`StaleCacheError` is raised before anything
    runs (critmc/runner.py:154-199).
    """
    out = str(out_dir) if out_dir is not None else None
    if out is not None:
        Path(out).mkdir(parents=True, exist_ok=True)
        for stale in Path(out).glob("*.part.npz"):
            stale.unlink()
    results: dict[int, TimeSeries] = {}
    pending = []
    for i, t in enumerate(tasks):
        path = Path(out) / t.filename if out is not None else None
        if path is not None and path.exists():
            results[i] = _load_cached(t, path)
        else:
            pending.append(i)
    groups: dict[Any, list[int]] = {}
    singles = []
    for i in pending:
        if batch and _batchable(tasks[i]):
            groups.setdefault(_group_key(tasks[i]), []).append(i)
        else:
            singles.append(i)
    bar = _Progress(len(pending), progress)
    for idx in groups.values():
        for i, ts in zip(idx, _run_ising_group([tasks[i] for i in idx])):
            results[i] = ts
            if out is not None:
                _save(ts, out, tasks[i])
        bar.add(len(idx))
    for i in singles:
        results[i] = run_task(tasks[i])
        if out is not None:
            _save(results[i], out, tasks[i])
        bar.add()
    return [results[i] for i in range(len(tasks))]


class _Progress:
    """A stderr line rewritten as items finish: count, elapsed time, linear estimate of
    the time left; a newline after the last item."""

    def __init__(self, total: int, enabled: bool = True):
        self.total, self.enabled = total, enabled
        self.done = 0
        self.t0 = time.perf_counter()

    def add(self, n: int = 1) -> None:
        self.done += n
        if not self.enabled:
            return
        elapsed = time.perf_counter() - self.t0
        left = elapsed / max(self.done, 1) * (self.total - self.done)
        end = "\n" if self.done == self.total else ""
        sys.stderr.write(f"\r  {self.done}/{self.total} done, {elapsed:7.1f}s elapsed, "
                         f"~{left:7.1f}s left{end}")
        sys.stderr.flush()


def parallel_map(fn: Callable, items: Iterable, *, workers: int | None = None,
                 progress: bool = True) -> list:
    """``[fn(x) for x in items]`` over a process pool, in input order, for picklable
    top-level functions (critmc/runner.py:202-223); one worker runs in-process."""
    items = list(items)
    n_workers = default_workers() if workers is None else max(1, int(workers))
    bar = _Progress(len(items), progress)
    out: list = [None] * len(items)
    if n_workers == 1 or len(items) <= 1:
        for k, item in enumerate(items):
            out[k] = fn(item)
            bar.add()
        return out
    with ProcessPoolExecutor(max_workers=min(n_workers, len(items))) as pool:
        slot = {pool.submit(fn, item): k for k, item in enumerate(items)}
        for fut in as_completed(slot):
            out[slot[fut]] = fut.result()
            bar.add()
    return out
