"""Continuous-batching request loop for lattice Monte Carlo jobs on one GPU.

Plays the role of critmc/runner.py:154-199 (``run_tasks``) for a stream of
independent chains: jobs ``(model, L, beta, sweeps, seed)`` arrive over time,
are admitted against an HBM byte budget, and advance together.  Jobs of the
same (model, L) share a resident slot pool; every tick advances every active
slot by up to ``slice_sweeps`` sweeps in one launch per pool (each slice on
the least loaded of a few HIP streams, so pools overlap on the device), with
per-slot caps so a job stops exactly at its sweep count.  A finished job frees its slot and budget at
once, and queued jobs take them on the next tick.

A job's chain depends only on its seed (Philox key, stream 0, its own sweep
counter), never on its slot or on the other jobs, so its result is identical
alone and in any mix.  Time series stay in device memory until the caller
asks for them.
"""

from __future__ import annotations

import heapq
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from ._lib import check, lib
from .device import device, ptr, stream
from .kernels import ising as k_ising
from .kernels import on as k_on
from .kernels import potts as k_potts
from .rng import chain_key, make_rng

MODELS = ("ising", "potts", "xy")
SCHEDULE = {"ising": [[0, 1]], "potts": [[2, 1]], "xy": [[0, 1]]}


@dataclass
class Job:
    """One chain to run: ``sweeps`` local sweeps, one measurement per sweep."""

    model: str
    L: int
    beta: float
    sweeps: int
    seed: int
    q: int = 10
    step: float = 1.0
    arrival: float = 0.0
    id: int = -1
    admitted_at: float = math.nan
    finished_at: float = math.nan
    done: int = 0
    series: dict = field(default_factory=dict, repr=False)

    @property
    def sites(self) -> int:
        return self.L * self.L

    @property
    def latency(self) -> float:
        return self.finished_at - self.arrival

    def state_bytes(self) -> int:
        return self.sites * (8 if self.model == "xy" else 1) + 64

    def series_bytes(self) -> int:
        per = {"ising": 8, "potts": 4 + 4 * self.q, "xy": 8 * 3}[self.model]
        return self.sweeps * per

    def footprint(self) -> int:
        return self.state_bytes() + self.series_bytes()

    def results(self) -> dict[str, np.ndarray]:
        return {k: v.cpu().numpy() for k, v in self.series.items()}


# Largest Potts replica, in states, that a serving pool runs in the per-replica
# form (one workgroup per replica, the lattice in LDS): each replica is one
# workgroup, so its sweep takes longer as L grows; measured on the 10^4-job
# stream, 256^2 pools did better with the launch per half sweep.
POTTS_RESIDENT_STATES = 128 * 128
# Pools of more jobs than this leave the form to the kernels' automatic choice,
# which is tuned for throughput over many replicas.
RESIDENT_MAX_JOBS = 256


def _ising_resident_ok(L: int) -> bool:
    """Whether an L x L Ising replica fits the batched form (its packed words in LDS)."""
    M = L // 2
    return (M <= 32 or M % 32 == 0) and 2 * L * (-(-M // 32)) * 4 <= 96 * 1024


class _Pool:
    """Resident state of all slots of one (model, L) group."""

    def __init__(self, model: str, L: int, q: int, dev, stream, resident: bool = True):
        self.model, self.L, self.q = model, L, q
        self.dev, self.stream = dev, stream
        # One launch per slice (each replica a workgroup holding its lattice on
        # chip for every sweep of the slice) instead of the launch per half
        # sweep the automatic choice makes for few replicas of a large lattice:
        # a serving stream is many small pools sharing the GPU, and the host's
        # kernel launches, not CUs per pool, set its pace.  Ising: wherever the
        # lattice fits; Potts: up to 128^2 (see POTTS_RESIDENT_STATES).
        self.resident = resident
        self.cap = 0
        self.jobs: list[Job | None] = []
        self.free: list[int] = []
        self.pending: list[Job] = []      # admitted, joining at the next slice boundary
        self.inflight = None              # caps of the slice on the device
        self.enqueue = None               # future of a slice being enqueued by a launch thread
        self.start_event = torch.cuda.Event(enable_timing=True)
        self.done_event = torch.cuda.Event(enable_timing=True)
        self.busy_s = 0.0                 # device time of this pool's slices
        self.sec_per_sweep = 0.0          # running estimate of one slice sweep's device time
        self.site_updates = 0
        self.ready = torch.cuda.Event()   # recorded after the pool's last queued work
        self.stream_idx = 0
        self.est_s = 0.0                  # estimated device time of the slice in flight
        with torch.cuda.stream(self.stream):
            self._grow(8)
            self.ready.record(self.stream)

    def use_stream(self, stream, idx: int):
        """Queue the pool's next work on ``stream``.  A pool has one slice in
        flight at most, so any stream will do; the new stream first waits for
        the pool's last work on the old one (on the device, not the host)."""
        if stream is not self.stream:
            stream.wait_event(self.ready)
            self.stream = stream
        self.stream_idx = idx

    # -- storage --------------------------------------------------------------
    def _alloc(self, cap: int):
        N, dv = self.L * self.L, self.dev
        st = {
            "keys": torch.zeros((cap, 2), dtype=torch.int32, device=dv),
            "sids": torch.zeros(cap, dtype=torch.int32, device=dv),
            "sweeps": torch.zeros(cap, dtype=torch.int64, device=dv),
            "e": torch.zeros(cap, dtype=torch.int32, device=dv),
        }
        if self.model == "xy":
            st["spins"] = torch.zeros((cap, N, 2), dtype=torch.float32, device=dv)
            st["spins"][:, :, 0] = 1.0
        else:
            st["spins"] = torch.zeros((cap, N), dtype=torch.int8, device=dv)
            if self.model == "ising":
                st["spins"].fill_(1)
                st["m"] = torch.zeros(cap, dtype=torch.int32, device=dv)
            else:
                st["counts"] = torch.zeros((cap, self.q), dtype=torch.int32, device=dv)
        return st

    def _grow(self, cap: int):
        new = self._alloc(cap)
        if self.cap:
            cur = torch.cuda.current_stream()
            for k, v in self.state.items():
                new[k][: self.cap].copy_(v)
                # v was allocated on the stream the pool used then: keep its
                # memory from reuse there until this copy has run
                v.record_stream(cur)
        self.state = new
        self.free.extend(range(self.cap, cap))
        self.jobs.extend([None] * (cap - self.cap))
        old = self.cap
        self.cap = cap
        self.betas = np.resize(getattr(self, "betas", np.zeros(0)), cap)
        self.steps = np.resize(getattr(self, "steps", np.zeros(0)), cap)
        # Host values reach the device from pinned buffers with copies that do
        # not wait for the stream (a blocking copy would wait for every slice
        # queued on the stream, including other pools').  A buffer is written
        # again only after the slice that followed its copy has completed.
        keys_h = torch.zeros((cap, 2), dtype=torch.int32).pin_memory()
        if old:
            keys_h[:old].copy_(self.keys_h)
        self.keys_h = keys_h
        self.caps_h = torch.zeros(cap, dtype=torch.int64).pin_memory()
        self.caps_d = torch.zeros(cap, dtype=torch.int64, device=self.dev)
        # (dst, src, bytes) of every job's slice of the time series, one copy launch per slice
        self.spans_h = torch.zeros((2 * cap, 3), dtype=torch.int64).pin_memory()
        self.spans_d = torch.zeros((2 * cap, 3), dtype=torch.int64, device=self.dev)
        if self.model == "xy":
            self.onpar_h = torch.zeros((2, cap), dtype=torch.float32).pin_memory()
            self.onpar_d = torch.zeros((2, cap), dtype=torch.float32, device=self.dev)
            self.onpar_h[0].copy_(torch.from_numpy(self.betas.astype(np.float32)))
            self.onpar_h[1].copy_(torch.from_numpy(self.steps.astype(np.float32)))
            self.onpar_dirty = True
        self.chains = None
        # scratch kept for the pool's life (the Ising tables are rebuilt only
        # when a slot's beta changes)
        shape = (self.L, self.L)
        if self.model == "ising":
            self.ws = k_ising.Workspace(cap, self.dev)
        elif self.model == "potts":
            self.ws = k_potts.Workspace(cap, self.dev)
        else:
            self.ws = k_on.Workspace(shape, cap, 2, self.dev)

    def add(self, job: Job):
        with torch.cuda.stream(self.stream):
            self._add(job)
            self.ready.record(self.stream)

    def _add(self, job: Job):
        if not self.free:
            self._grow(2 * self.cap)
        slot = min(self.free)
        self.free.remove(slot)
        self.jobs[slot] = job
        self.betas[slot] = job.beta
        self.steps[slot] = job.step
        if self.model == "xy":
            self.onpar_h[0, slot] = float(np.float32(job.beta))
            self.onpar_h[1, slot] = float(np.float32(job.step))
            self.onpar_dirty = True
        key = np.array(chain_key(make_rng(job.seed)), dtype=np.uint64).astype(np.uint32)
        self.keys_h[slot] = torch.from_numpy(key.view(np.int32))
        st = self.state
        st["keys"][slot].copy_(self.keys_h[slot], non_blocking=True)
        # fill kernels: assigning a Python number to a device element would
        # copy it from pageable host memory and wait for the stream
        st["sids"][slot: slot + 1].zero_()
        st["sweeps"][slot: slot + 1].zero_()
        one = _DeviceChains(st["keys"][slot: slot + 1], st["sids"][slot: slot + 1],
                            st["sweeps"][slot: slot + 1], 1)
        shape = (self.L, self.L)
        if self.model == "ising":
            s = st["spins"][slot: slot + 1]
            k_ising.init_spins(s, shape, one, "random")
            e, m = k_ising.measure(s, shape, 1)
            st["e"][slot] = e[0]
            st["m"][slot] = m[0]
        elif self.model == "potts":
            s = st["spins"][slot: slot + 1]
            k_potts.init_states(s, shape, self.q, one, "random")
            e, c = k_potts.measure(s, shape, self.q, 1)
            st["e"][slot] = e[0]
            st["counts"][slot] = c[0]
        else:
            k_on.init_spins(st["spins"][slot: slot + 1], shape, 2, one)
        job.series = self._series_buffers(job)
        self.chains = None

    def _series_buffers(self, job: Job) -> dict:
        n, dv = job.sweeps, self.dev
        if self.model == "ising":
            return {"E": torch.empty(n, dtype=torch.int32, device=dv),
                    "M": torch.empty(n, dtype=torch.int32, device=dv)}
        if self.model == "potts":
            return {"E": torch.empty(n, dtype=torch.int32, device=dv),
                    "counts": torch.empty((n, self.q), dtype=torch.int32, device=dv)}
        return {"E": torch.empty(n, dtype=torch.float64, device=dv),
                "M": torch.empty((n, 2), dtype=torch.float64, device=dv)}

    def remove(self, slot: int):
        self.jobs[slot] = None
        self.free.append(slot)

    @property
    def active(self) -> int:
        return sum(j is not None for j in self.jobs)

    # -- one tick ------------------------------------------------------------
    def launch(self, slice_sweeps: int):
        """Queue one slice for every active slot on this pool's stream."""
        caps = np.zeros(self.cap, dtype=np.int64)
        for s, j in enumerate(self.jobs):
            if j is not None:
                caps[s] = min(slice_sweeps, j.sweeps - j.done)
        n = int(caps.max())
        if n == 0:
            return None
        st = self.state
        with torch.cuda.stream(self.stream):
            self.caps_h.numpy()[:] = caps
            caps_d = self.caps_d
            caps_d.copy_(self.caps_h, non_blocking=True)
            if self.model == "xy" and self.onpar_dirty:
                self.onpar_d.copy_(self.onpar_h, non_blocking=True)
                self.onpar_dirty = False
            self.start_event.record(self.stream)
            ch = _DeviceChains(st["keys"], st["sids"], st["sweeps"], self.cap)
            sched = np.array(SCHEDULE[self.model])
            shape = (self.L, self.L)
            if self.model == "ising":
                small = self.active <= RESIDENT_MAX_JOBS
                form = (k_ising.FORM_BATCHED if self.resident and small and _ising_resident_ok(self.L)
                        else k_ising.FORM_AUTO)
                out = k_ising.run(st["spins"], shape, self.betas, 0.0, ch, sched, n, st["e"], st["m"],
                                  form=form, ws=self.ws, n_meas_rep=caps_d)
                rec = {"E": out[0], "M": out[1]}
            elif self.model == "potts":
                grid = (self.cap if self.resident and self.active <= RESIDENT_MAX_JOBS
                        and self.L * self.L <= POTTS_RESIDENT_STATES else 0)
                e, _, counts = k_potts.run(st["spins"], shape, self.q, self.betas, ch, sched, n,
                                           st["e"], st["counts"], grid=grid, ws=self.ws,
                                           n_meas_rep=caps_d)
                rec = {"E": e, "counts": counts}
            else:
                _, e, m, _, _ = k_on.run(st["spins"], shape, 2, self.onpar_d[0], self.onpar_d[1], ch,
                                         sched, n, ws=self.ws, n_meas_rep=caps_d)
                rec = {"E": e, "M": m}
            spans = self.spans_h.numpy()
            ns = 0
            for s, j in enumerate(self.jobs):
                if j is not None and caps[s]:
                    for k, v in rec.items():
                        dst = j.series[k]
                        per = dst[0].numel() * dst.element_size()  # bytes of one measurement
                        spans[ns] = (dst.data_ptr() + j.done * per,
                                     v.data_ptr() + s * v.stride(0) * v.element_size(),
                                     int(caps[s]) * per)
                        ns += 1
            if ns:
                self.spans_d[:ns].copy_(self.spans_h[:ns], non_blocking=True)
                check(lib().cm_copy_spans(ptr(self.spans_d), ns, stream()), "cm_copy_spans")
            self.done_event.record(self.stream)
            self.ready.record(self.stream)
        self.inflight = caps
        return caps


class _DeviceChains:
    """Chains view over a pool's device arrays (no host mirror needed)."""

    def __init__(self, keys, sids, sweeps, R):
        self.keys_d, self.sids_d, self.sweeps_d, self.R = keys, sids, sweeps, R
        self.sweeps_h = np.zeros(R, dtype=np.uint64)

    def advance(self, n):  # the device counters are advanced by the kernels
        pass


class Server:
    """Admits jobs against ``budget_bytes`` and advances them in ticks of ``slice_sweeps``."""

    def __init__(self, budget_bytes: float = 200e9, slice_sweeps: int = 500, n_streams: int = 4,
                 max_active: int = 1 << 16, slice_seconds: float = 0.0, min_slice: int = 16,
                 resident: bool = True, launch_threads: int = 0, balance_streams: bool = True):
        """``slice_sweeps`` caps the sweeps of one slice.  With ``slice_seconds`` > 0 a
        pool's slice is instead sized from its measured device time per sweep to take
        about that long (at least ``min_slice`` sweeps), so a large pool does not hold
        its stream for long while smaller pools sharing it wait.  Slicing never
        changes results: a job's chain depends only on its seed.  ``resident`` runs
        each Ising and Potts pool whose lattice fits on chip in its one-launch
        per-slice form; False leaves the form to the kernels' automatic choice
        (the same chains either way).  ``launch_threads`` > 0 enqueues the pools'
        slices from that many host threads (the library releases the interpreter
        while it launches), so a pool whose slice is thousands of launches does
        not hold up the others; 0 enqueues them from the calling thread.
        ``balance_streams`` queues each slice on the stream with the least
        estimated device time still queued (a pool's last slices' time per
        sweep times the slice's sweeps), so a long job's many slices do not
        each wait behind the pools that happen to share a fixed stream; False
        keeps every pool on the stream it was given when it was created.
        Streams beyond the process's hardware queues share them (HIP gives 4
        unless the environment sets ``GPU_MAX_HW_QUEUES`` before HIP starts)."""
        self.dev = device()
        self.budget = float(budget_bytes)
        self.used = 0.0
        self.slice = int(slice_sweeps)
        self.slice_seconds = float(slice_seconds)
        self.min_slice = max(1, int(min_slice))
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(n_streams)]
        self.pools: dict[tuple, _Pool] = {}
        self.queue: list[tuple[float, int, Job]] = []
        self.waiting: list[Job] = []
        self.finished: list[Job] = []
        self.max_active = max_active
        self.resident = bool(resident)
        self.balance_streams = bool(balance_streams)
        self._stream_est = [0.0] * len(self.streams)  # estimated device seconds queued per stream
        self.launch_threads = int(launch_threads)
        self._launcher = None
        if self.launch_threads > 0:
            from concurrent.futures import ThreadPoolExecutor

            self._launcher = ThreadPoolExecutor(max_workers=self.launch_threads,
                                                thread_name_prefix="critmc-serve")
        self.n_active = 0
        self._ids = 0
        self.site_updates = 0
        self.peak_used = 0.0
        self.ticks = 0

    def submit(self, job: Job) -> Job:
        if job.model not in MODELS:
            raise ValueError(f"unknown model {job.model!r}; choose from {MODELS}")
        if job.L < 2 or job.L % 2:
            raise ValueError("checkerboard sweeps need an even L >= 2")
        if job.sweeps < 1:
            raise ValueError("sweeps must be positive")
        if job.footprint() > self.budget:
            raise ValueError("job does not fit the memory budget")
        job.id = self._ids
        self._ids += 1
        heapq.heappush(self.queue, (job.arrival, job.id, job))
        return job

    def _pool(self, job: Job) -> _Pool:
        key = (job.model, job.L, job.q if job.model == "potts" else 0)
        if key not in self.pools:
            idx = len(self.pools) % len(self.streams)
            pool = _Pool(job.model, job.L, job.q, self.dev, self.streams[idx], resident=self.resident)
            pool.stream_idx = idx
            self.pools[key] = pool
        return self.pools[key]

    def _admit(self, now: float):
        while self.queue and self.queue[0][0] <= now:
            self.waiting.append(heapq.heappop(self.queue)[2])
        still = []
        for job in self.waiting:
            if self.n_active < self.max_active and self.used + job.footprint() <= self.budget:
                self.used += job.footprint()
                self.peak_used = max(self.peak_used, self.used)
                self._pool(job).pending.append(job)
                job.admitted_at = now
                self.n_active += 1
            else:
                still.append(job)
        self.waiting = still

    def _queued(self, pool: _Pool, est_s: float):
        """Book (est_s > 0) or release (est_s = 0) the pool's slice on its stream."""
        i = pool.stream_idx
        self._stream_est[i] = max(0.0, self._stream_est[i] - pool.est_s) + est_s
        pool.est_s = est_s

    def _retire(self, pool: _Pool, now: float):
        self._queued(pool, 0.0)
        caps, pool.inflight = pool.inflight, None
        dt = pool.start_event.elapsed_time(pool.done_event) / 1e3
        pool.busy_s += dt
        sps = dt / max(int(caps.max()), 1)
        pool.sec_per_sweep = sps if pool.sec_per_sweep == 0.0 else 0.5 * (pool.sec_per_sweep + sps)
        for slot, job in enumerate(pool.jobs):
            if job is None or not caps[slot]:
                continue
            job.done += int(caps[slot])
            self.site_updates += int(caps[slot]) * job.sites
            pool.site_updates += int(caps[slot]) * job.sites
            if job.done >= job.sweeps:
                job.finished_at = now
                pool.remove(slot)
                self.used -= job.footprint()
                self.n_active -= 1
                self.finished.append(job)

    def run(self, *, realtime: bool = True, t0: float | None = None,
            progress: float = 0.0) -> list[Job]:
        """Serve until every submitted job has finished; returns them in completion order.

        Pools run independently: whenever a pool's slice has completed on the
        device, its finished jobs are retired, its newly admitted jobs join,
        and its next slice is launched, while other pools keep running.  With
        ``realtime`` the arrival times are seconds after ``t0`` (default: now);
        otherwise every job is available at once.  ``progress`` > 0 prints a
        status line to stderr every that many seconds.
        """
        import sys

        t0 = time.perf_counter() if t0 is None else t0
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: math.inf)
        stamp = lambda: time.perf_counter() - t0  # noqa: E731
        last = time.perf_counter()
        torch.cuda.synchronize()
        while self.queue or self.waiting or self.n_active:
            self._admit(clock())
            busy = False
            for pool in self.pools.values():
                if pool.enqueue is not None:  # its slice is still being enqueued
                    if not pool.enqueue.done():
                        busy = True
                        continue
                    caps, pool.enqueue = pool.enqueue.result(), None
                    if caps is not None:
                        self.ticks += 1
                    else:
                        self._queued(pool, 0.0)
                if pool.inflight is not None:
                    if not pool.done_event.query():
                        busy = True
                        continue
                    self._retire(pool, stamp())
                if not (pool.pending or pool.active):
                    continue
                if self.balance_streams:
                    i = min(range(len(self.streams)), key=self._stream_est.__getitem__)
                    pool.use_stream(self.streams[i], i)
                for job in pool.pending:
                    pool.add(job)
                pool.pending = []
                if not pool.active:
                    continue
                n = self._slice_for(pool)
                self._queued(pool, pool.sec_per_sweep * n if pool.sec_per_sweep > 0.0 else 1e-3)
                if self._launcher is not None:
                    pool.enqueue = self._launcher.submit(pool.launch, n)
                    busy = True
                elif pool.launch(n) is not None:
                    self.ticks += 1
                    busy = True
                else:
                    self._queued(pool, 0.0)
            if progress > 0 and time.perf_counter() - last > progress:
                last = time.perf_counter()
                print(f"serve: t={last - t0:.1f}s finished={len(self.finished)} active={self.n_active} "
                      f"queued={len(self.queue) + len(self.waiting)} launches={self.ticks}",
                      file=sys.stderr, flush=True)
            if not busy:
                if self.queue and not self.waiting and realtime:
                    time.sleep(max(0.0, min(self.queue[0][0] - clock(), 0.002)))
            else:
                time.sleep(20e-6)
        torch.cuda.synchronize()
        return self.finished

    def _slice_for(self, pool: _Pool) -> int:
        if self.slice_seconds <= 0.0 or pool.sec_per_sweep <= 0.0:
            return self.slice
        return max(self.min_slice, min(self.slice, int(self.slice_seconds / pool.sec_per_sweep)))

    def pool_stats(self) -> dict[str, dict]:
        """Per (model, L) pool: device seconds of its slices (pools overlap on the
        device, so these add up to more than the wall time) and site updates."""
        return {f"{m}-L{L}" + (f"-q{q}" if m == "potts" else ""):
                {"busy_s": round(p.busy_s, 3), "site_updates": p.site_updates,
                 "site_updates_per_ns": round(p.site_updates / max(p.busy_s, 1e-9) / 1e9, 2)}
                for (m, L, q), p in sorted(self.pools.items())}


def poisson_jobs(n: int, rate: float, seed: int = 2024, L_choices=(16, 32, 64, 128, 256, 512),
                 sweeps_range=(1e3, 1e5), beta_spread: float = 0.10) -> list[Job]:
    """Synthetic stream: Poisson arrivals; model uniform over Ising / Potts q=10 / XY;
    L log-uniform over ``L_choices``; beta within +-``beta_spread`` of the model's
    critical or transition point; sweeps log-uniform; one seed per job."""
    rng = np.random.default_rng(seed)
    beta_c = {"ising": math.log(1 + math.sqrt(2)) / 2, "potts": math.log(1 + math.sqrt(10)),
              "xy": 1.0 / 0.89294}
    t = 0.0
    jobs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        model = MODELS[rng.integers(0, 3)]
        L = int(L_choices[rng.integers(0, len(L_choices))])
        beta = beta_c[model] * (1 + rng.uniform(-beta_spread, beta_spread))
        sweeps = int(round(math.exp(rng.uniform(math.log(sweeps_range[0]), math.log(sweeps_range[1])))))
        jobs.append(Job(model, L, float(beta), sweeps, seed=int(rng.integers(0, 2**31)), arrival=t))
    return jobs
