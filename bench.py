"""Benchmark: batched 2D Ising, 4096 replicas of a 256x256 torus, checkerboard Metropolis.

Workload (one GPU): 4096 independent replicas, int8 spins (or the 32-spin
packed layout with --layout packed), hot start drawn on the device from seed
1234, beta_r ~ U[0.40, 0.48] per replica, one Metropolis sweep per
measurement with integer E_b, M and accepted counts recorded per replica per
sweep.  A "step" is one sweep of the whole batch.  Inputs are resident in HBM
before timing; the timed region is K steps bracketed by synchronisation.

Printed (one JSON line): site updates/ns, the HBM-roofline record of the
sweep kernel (algorithmic 2 B per site update for int8, 0.25 B packed), the
plain one-chain-per-launch form of the same kernel on one replica for
comparison, and the CPU baseline (the oracle's OpenMP restatement of the
reference's typewriter Metropolis) on a bounded sample.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
L, R, SEED = 256, 4096, 1234


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10_000, help="timed sweeps of the whole batch")
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--layout", choices=("int8", "packed"), default="int8")
    ap.add_argument("--sweeps-per-launch", type=int, default=0,
                    help="sweeps per kernel launch (0 = all timed steps in one launch)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--replicas", type=int, default=R)
    ap.add_argument("--plain-sweeps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=str(ROOT / "profiles" / "ising2d_traffic.json"))
    return ap.parse_args()


def cpu_baseline(seconds: float) -> dict:
    """Oracle typewriter Metropolis (the reference's sweep order) on host cores."""
    import oracle

    threads = min(16, os.cpu_count() or 1)
    reps = threads
    rng = np.random.default_rng(SEED)
    s = (2 * rng.integers(0, 2, size=(reps, L * L)) - 1).astype(np.int8)
    betas = rng.uniform(0.40, 0.48, size=reps)
    keys = rng.integers(0, 1 << 32, size=(reps, 2), dtype=np.uint64)
    t0 = time.perf_counter()
    oracle.ising_batch_typewriter(s, (L, L), betas, 0.0, keys, 2, threads)
    dt = time.perf_counter() - t0
    sweeps = max(2, int(2 * seconds / max(dt, 1e-3)))
    t0 = time.perf_counter()
    oracle.ising_batch_typewriter(s, (L, L), betas, 0.0, keys, sweeps, threads)
    dt = time.perf_counter() - t0
    return {"value": reps * L * L * sweeps / dt / 1e9, "unit": "site updates/ns", "cores": threads,
            "kind": "port",
            "sample": f"{reps} replicas x {L}x{L} x {sweeps} typewriter Metropolis sweeps, "
                      f"one replica per thread ({dt:.1f} s)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; replicas shard across GPUs with no "
                                   "traffic, so --gpus N > 1 is not measured here"}))
        sys.exit(2)

    import torch

    from critmc_amd.kernels import ising as K
    from critmc_amd.kernels.chains import Chains
    from critmc_amd.rng import seed_sequence

    dev = torch.device("cuda", 0)
    Rn = args.replicas
    shape = (L, L)
    layout = K.LAYOUT_PACKED if args.layout == "packed" else K.LAYOUT_INT8
    bytes_per_update = 0.25 if layout == K.LAYOUT_PACKED else 2.0

    # synthetic inputs: keys from seed 1234, hot start on the device, beta ~ U[0.40, 0.48]
    key = seed_sequence(SEED).generate_state(2, np.uint32)
    keys = np.tile(key.astype(np.uint64), (Rn, 1))
    chains = Chains(keys, np.arange(Rn, dtype=np.uint32), None, dev)
    betas = np.random.default_rng(SEED).uniform(0.40, 0.48, size=Rn)
    s = torch.empty((Rn, L * L), dtype=torch.int8, device=dev)
    K.init_spins(s, shape, chains, "random")
    e_run, m_run = K.measure(s, shape, Rn)
    state = K.pack(s, shape, Rn) if layout == K.LAYOUT_PACKED else s
    ws = K.Workspace(Rn, dev)
    sched = np.array([[0, 1]])

    spl = args.sweeps_per_launch or args.steps
    n_launch = -(-args.steps // spl)
    out = tuple(torch.empty((Rn, spl), dtype=torch.int32, device=dev) for _ in range(3))

    def sweep_block(n):
        o = out if n == spl else None
        return K.run(state, shape, betas, 0.0, chains, sched, n, e_run, m_run, layout=layout,
                     form=K.FORM_BATCHED, block=args.block, ws=ws, out=o)

    if args.warmup:
        sweep_block(min(args.warmup, spl))
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(n_launch + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    done = 0
    for i in range(n_launch):
        k = min(spl, args.steps - done)
        sweep_block(k)
        done += k
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [ev[i].elapsed_time(ev[i + 1]) for i in range(n_launch)]
    gpu_ms = sum(launch_ms)
    updates = Rn * L * L * args.steps
    value = updates / (gpu_ms * 1e6)

    # consistency: incremental bookkeeping equals a recount
    chk_s = K.unpack(state, shape, Rn) if layout == K.LAYOUT_PACKED else state
    e_chk, m_chk = K.measure(chk_s, shape, Rn)
    consistent = bool(torch.equal(e_chk, e_run) and torch.equal(m_chk, m_run))

    # roofline of the sweep kernel: algorithmic bytes per launch / mean launch time
    full = [t for t, i in zip(launch_ms, range(n_launch)) if (i + 1) * spl <= args.steps] or launch_ms
    mean_launch_s = float(np.mean(full)) / 1e3
    achieved = bytes_per_update * Rn * L * L * spl / mean_launch_s / 1e9
    traffic = None
    tpath = Path(args.traffic)
    if tpath.exists():
        t = json.loads(tpath.read_text())
        traffic = t.get(f"{args.layout}_spl{spl}_bytes_per_launch")

    # plain one-chain-per-launch form of the same kernel on one replica
    plain_rate = None
    if args.plain_sweeps > 0:
        c1 = Chains(keys[:1], [0], None, dev)
        s1 = s[:1].clone() if layout == K.LAYOUT_INT8 else K.unpack(state, shape, Rn)[:1].clone()
        e1, m1 = K.measure(s1, shape, 1)
        K.run(s1, shape, betas[:1], 0.0, c1, sched, 5, e1, m1, form=K.FORM_PLAIN)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        K.run(s1, shape, betas[:1], 0.0, c1, sched, args.plain_sweeps, e1, m1, form=K.FORM_PLAIN)
        torch.cuda.synchronize()
        plain_rate = L * L * args.plain_sweeps / ((time.perf_counter() - t1) * 1e9)

    cpu = None if args.no_cpu else cpu_baseline(args.cpu_seconds)

    line = {
        "metric": "site updates/ns, 2D Ising L=256 x4096 reps vs 1-chain GPU; p50/p99 job latency",
        "value": round(value, 3),
        "unit": "site updates/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": gpu_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8" if layout == K.LAYOUT_INT8 else "u32 (32-spin packed)",
        "data": "synthetic: device hot start from seed 1234, beta ~ U[0.40,0.48] per replica",
        "config": {"workload": "ising2d_batched_checkerboard_metropolis", "L": L, "replicas": Rn,
                   "layout": args.layout, "sweeps_per_launch": spl, "launches": n_launch,
                   "record": "int32 E_b, M, accepted per replica per sweep"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "plain_one_chain_rate": plain_rate,
        "vs_plain_one_chain": (value / plain_rate) if plain_rate else None,
        "wall_s": wall,
        "bookkeeping_consistent": consistent,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
