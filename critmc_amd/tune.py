"""Launch-shape autotuner for the checkerboard kernels.

For a lattice size and replica count it times the candidate launch shapes on
the GPU and keeps the fastest:

* Ising: the kernel form (batched, one LDS-resident workgroup per replica, or
  wide, one launch per half sweep over all replicas), the workgroup size of
  the batched form (64, 128 or 256 threads per replica) and the HBM layout
  (int8 or 32-spin packed);
* Potts: the form (one workgroup per replica with the states in LDS, or wide)
  and the workgroup size.

A few sweeps of each candidate run on scratch replicas.  Results are cached
per device name, in memory and optionally in a JSON file, so the measurement
runs once per shape.  Tuning never changes results: every candidate draws the
same uniforms for the same site and sweep, so all of them run the same chains
bit for bit.
"""

from __future__ import annotations

import json
import time
from pathlib import Path

import numpy as np
import torch

from .kernels import ising as k_ising
from .kernels import potts as k_potts
from .kernels.chains import Chains

_CACHE: dict[str, dict] = {}

POTTS_LDS_SITES = 96 * 1024  # states per replica that fit the LDS-resident form


def _device_name() -> str:
    return torch.cuda.get_device_name(torch.cuda.current_device())


def _key(kind: str, L: int, R: int) -> str:
    return f"{_device_name()}|{kind}|L{L}|R{R}"


def _cached(key: str, cache_file) -> dict | None:
    if key in _CACHE:
        return _CACHE[key]
    if cache_file is not None and Path(cache_file).exists():
        disk = json.loads(Path(cache_file).read_text())
        if key in disk:
            _CACHE[key] = disk[key]
            return disk[key]
    return None


def _store(key: str, best: dict, cache_file) -> dict:
    _CACHE[key] = best
    if cache_file is not None:
        p = Path(cache_file)
        disk = json.loads(p.read_text()) if p.exists() else {}
        disk[key] = best
        p.write_text(json.dumps(disk, indent=1))
    return best


def ising_candidates(L: int, *, forms=("batched", "wide"), blocks=(64, 128, 256),
                     layouts=("int8", "packed")) -> list[dict]:
    """Launch shapes that apply to an L x L lattice."""
    W = L * ((L // 2 + 31) // 32)  # words per colour
    out = []
    for layout in layouts:
        for form in forms:
            if form == "wide":
                if L % 64 == 0:
                    out.append({"form": "wide", "block": 0, "layout": layout})
                continue
            if not (L <= 64 or L % 64 == 0):
                continue
            for block in blocks:
                if block > max(64, ((W + 63) // 64) * 64) and block != blocks[0]:
                    continue  # more threads than words per colour: idle lanes only
                out.append({"form": "batched", "block": block, "layout": layout})
    return out


def _ising_form(c: dict) -> int:
    return k_ising.FORM_WIDE if c["form"] == "wide" else k_ising.FORM_BATCHED


def _ising_layout(c: dict) -> int:
    return k_ising.LAYOUT_PACKED if c["layout"] == "packed" else k_ising.LAYOUT_INT8


def run_ising_candidate(c: dict, s0: torch.Tensor, L: int, betas, sweeps: int, seed: int = 0):
    """Run ``sweeps`` Metropolis sweeps of candidate ``c`` from int8 states ``s0``;
    returns (final int8 states, (E_b, M, accepted) series)."""
    R = s0.shape[0]
    dev = s0.device
    shape = (L, L)
    keys = (np.arange(2 * R, dtype=np.uint64) + seed).reshape(R, 2)
    ch = Chains(keys, np.arange(R), None, dev)
    s = s0.clone()
    e, m = k_ising.measure(s, shape, R)
    lay = _ising_layout(c)
    st = k_ising.pack(s, shape, R) if lay == k_ising.LAYOUT_PACKED else s
    series = k_ising.run(st, shape, betas, 0.0, ch, np.array([[0, 1]]), sweeps, e, m, layout=lay,
                         form=_ising_form(c), block=c["block"])
    final = k_ising.unpack(st, shape, R) if lay == k_ising.LAYOUT_PACKED else st
    return final, series


def tune_ising(L: int, R: int, *, sweeps: int = 20, forms=("batched", "wide"),
               blocks=(64, 128, 256), layouts=("int8", "packed"),
               cache_file: str | Path | None = None) -> dict:
    """Best ``{"form", "block", "layout", "rate"}`` for R replicas of an L x L Ising batch
    (checkerboard Metropolis near the critical point), plus every candidate's rate."""
    key = _key("ising2d", L, R)
    hit = _cached(key, cache_file)
    if hit is not None:
        return hit
    dev = torch.device("cuda", torch.cuda.current_device())
    shape = (L, L)
    keys = np.arange(2 * R, dtype=np.uint64).reshape(R, 2)
    betas = np.full(R, 0.44)
    results = []
    for c in ising_candidates(L, forms=forms, blocks=blocks, layouts=layouts):
        ch = Chains(keys, np.arange(R), None, dev)
        s = torch.empty((R, L * L), dtype=torch.int8, device=dev)
        k_ising.init_spins(s, shape, ch)
        e, m = k_ising.measure(s, shape, R)
        lay = _ising_layout(c)
        st = k_ising.pack(s, shape, R) if lay == k_ising.LAYOUT_PACKED else s
        ws = k_ising.Workspace(R, dev)
        kw = dict(layout=lay, form=_ising_form(c), block=c["block"], ws=ws)
        k_ising.run(st, shape, betas, 0.0, ch, np.array([[0, 1]]), 2, e, m, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k_ising.run(st, shape, betas, 0.0, ch, np.array([[0, 1]]), sweeps, e, m, **kw)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        results.append({**c, "rate": R * L * L * sweeps / dt / 1e9})
    best = max(results, key=lambda r: r["rate"])
    return _store(key, {**best, "candidates": results}, cache_file)


def potts_candidates(L: int, *, blocks=(64, 128, 256)) -> list[dict]:
    out = [{"form": "wide", "block": 0}]
    if L * L <= POTTS_LDS_SITES:
        out += [{"form": "per-replica", "block": b} for b in blocks]
    return out


def run_potts_candidate(c: dict, s0: torch.Tensor, L: int, q: int, betas, code: int, sweeps: int,
                        seed: int = 0):
    """Run ``sweeps`` sweeps (update ``code``) of candidate ``c`` from states ``s0``;
    returns (final states, (E, accepted, counts) series)."""
    R = s0.shape[0]
    shape = (L, L)
    keys = (np.arange(2 * R, dtype=np.uint64) + seed).reshape(R, 2)
    ch = Chains(keys, np.arange(R), None, s0.device)
    s = s0.clone()
    e, cnt = k_potts.measure(s, shape, q, R)
    grid = -1 if c["form"] == "wide" else R
    series = k_potts.run(s, shape, q, betas, ch, np.array([[code, 1]]), sweeps, e, cnt, grid=grid,
                         block=c["block"])
    return s, series


def tune_potts(L: int, R: int, q: int, *, code: int = 2, sweeps: int = 10, blocks=(64, 128, 256),
               cache_file: str | Path | None = None) -> dict:
    """Best ``{"form", "block", "rate"}`` for R replicas of an L x L q-state Potts batch
    (``code`` 2 = heat bath, 0 = Metropolis), plus every candidate's rate."""
    key = _key(f"potts{q}c{code}", L, R)
    hit = _cached(key, cache_file)
    if hit is not None:
        return hit
    dev = torch.device("cuda", torch.cuda.current_device())
    shape = (L, L)
    keys = np.arange(2 * R, dtype=np.uint64).reshape(R, 2)
    betas = np.full(R, np.log(1.0 + np.sqrt(q)))
    results = []
    for c in potts_candidates(L, blocks=blocks):
        ch = Chains(keys, np.arange(R), None, dev)
        s = torch.empty((R, L * L), dtype=torch.int8, device=dev)
        k_potts.init_states(s, shape, q, ch, "random")
        e, cnt = k_potts.measure(s, shape, q, R)
        ws = k_potts.Workspace(R, dev)
        kw = dict(grid=-1 if c["form"] == "wide" else R, block=c["block"], ws=ws)
        k_potts.run(s, shape, q, betas, ch, np.array([[code, 1]]), 2, e, cnt, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k_potts.run(s, shape, q, betas, ch, np.array([[code, 1]]), sweeps, e, cnt, **kw)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        results.append({**c, "rate": R * L * L * sweeps / dt / 1e9})
    best = max(results, key=lambda r: r["rate"])
    return _store(key, {**best, "candidates": results}, cache_file)
