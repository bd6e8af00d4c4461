"""Continuous batching: each job's result is the same alone and in the mix.

Jobs arriving at different times, of different models and sizes, share
launches; their series must equal, bit for bit, the series of the same job
served alone.  Admission never exceeds the byte budget.
"""

import numpy as np
import pytest

from critmc_amd.serve import Job, Server, poisson_jobs

pytestmark = pytest.mark.gpu

PROBES = [Job("ising", 32, 0.44, 1300, seed=11), Job("potts", 16, 1.42, 700, seed=12),
          Job("xy", 16, 1.1, 900, seed=13), Job("ising", 16, 0.40, 333, seed=14)]


def _clone(j, arrival=0.0):
    return Job(j.model, j.L, j.beta, j.sweeps, j.seed, q=j.q, step=j.step, arrival=arrival)


def _alone(jobs, **kw):
    out = []
    for j in jobs:
        srv = Server(**kw)
        job = srv.submit(_clone(j))
        srv.run(realtime=False)
        out.append(job.results())
    return out


@pytest.fixture(scope="module")
def probes_alone():
    return _alone(PROBES, budget_bytes=1e9, slice_sweeps=256)


# resident: small pools in their one-launch-per-slice forms (the default) or in
# the kernels' automatic choice; launch_threads: slices enqueued from host
# threads (the serving bench uses 4) or from the serving loop itself;
# balance: each slice on the least loaded stream (the default), or every pool
# on the stream it was created with
@pytest.mark.parametrize("resident,threads,balance", [(True, 0, True), (False, 0, True),
                                                      (True, 4, True), (True, 0, False)])
def test_each_result_is_the_same_alone_and_in_the_mix(probes_alone, resident, threads, balance):
    alone = probes_alone
    mix = poisson_jobs(24, rate=2000.0, seed=5, L_choices=(16, 32, 64), sweeps_range=(200, 2000))
    srv = Server(budget_bytes=1e9, slice_sweeps=300, slice_seconds=0.002,  # adaptive slices too
                 resident=resident, launch_threads=threads, balance_streams=balance)
    probes = [srv.submit(_clone(j, arrival=0.003 * (k + 1))) for k, j in enumerate(PROBES)]
    for j in mix:
        srv.submit(j)
    done = srv.run(realtime=True)
    assert len(done) == len(mix) + len(PROBES)
    for res, job in zip(alone, probes):
        got = job.results()
        assert got.keys() == res.keys()
        for k in res:
            assert np.array_equal(got[k], res[k]), (job.model, k)
    assert srv.ticks > 1


# Large lattices: the pools that carry most of the serving work.  With the
# resident forms, Ising pools run one launch per slice (each replica on chip)
# at every L here and Potts pools at 128^2; Potts 256^2 / 512^2 take the wide
# form (one launch per half sweep over all replicas), and so does every Ising
# and Potts pool without the resident forms.  XY takes the O(n) LDS form at
# 128^2 and the half-sweep launches above it.  Sweep counts differ from the
# background jobs', so the per-slot caps stop each job at its own count
# inside shared slices.
LARGE_PROBES = [Job(m, L, b, sw, seed=100 + k, arrival=0.004 * (1 + k % 3))
                for k, (m, L, b, sw) in enumerate([
                    ("ising", 128, 0.44, 333), ("ising", 256, 0.43, 517), ("ising", 512, 0.45, 261),
                    ("potts", 128, 1.40, 311), ("potts", 256, 1.45, 205), ("potts", 512, 1.43, 133),
                    ("xy", 128, 1.10, 219), ("xy", 256, 1.12, 151), ("xy", 512, 1.09, 97)])]


@pytest.fixture(scope="module")
def large_alone():
    return _alone(LARGE_PROBES, budget_bytes=4e9, slice_sweeps=64)


@pytest.mark.parametrize("resident,threads", [(True, 0), (False, 4)])
def test_large_lattices_are_the_same_alone_and_joining_running_pools(large_alone, resident,
                                                                     threads):
    alone = large_alone
    srv = Server(budget_bytes=8e9, slice_sweeps=64, resident=resident, launch_threads=threads)
    background = []
    for k, j in enumerate(LARGE_PROBES):  # two running jobs per pool before any probe arrives
        for b in range(2):
            background.append(srv.submit(Job(j.model, j.L, j.beta * (1.01 + 0.01 * b),
                                             20 * j.sweeps + 41 * b, seed=500 + 2 * k + b)))
    probes = [srv.submit(_clone(j, arrival=j.arrival)) for j in LARGE_PROBES]
    done = srv.run(realtime=True)
    assert len(done) == len(background) + len(probes)
    for res, job, bg in zip(alone, probes, zip(background[::2], background[1::2])):
        # the pool was running when the probe joined it, and kept running after
        assert min(b.admitted_at for b in bg) < job.admitted_at < max(b.finished_at for b in bg), \
            (job.model, job.L)
        got = job.results()
        assert got.keys() == res.keys()
        for k in res:
            assert np.array_equal(got[k], res[k]), (job.model, job.L, k)


def test_admission_respects_the_budget():
    jobs = [Job("ising", 64, 0.44, 3000, seed=s) for s in range(6)]
    budget = 2.5 * jobs[0].footprint()
    srv = Server(budget_bytes=budget, slice_sweeps=1000)
    for j in jobs:
        srv.submit(j)
    done = srv.run(realtime=False)
    assert len(done) == 6
    assert srv.peak_used <= budget
    assert all(j.done == j.sweeps for j in done)
    # never more than two at a time: the third started only after one finished
    starts = sorted(j.admitted_at for j in done)
    assert starts[2] > 0


def test_latency_report():
    srv = Server(budget_bytes=2e9, slice_sweeps=500)
    for j in poisson_jobs(40, rate=500.0, seed=9, L_choices=(16, 32, 64, 128), sweeps_range=(300, 3000)):
        srv.submit(j)
    done = srv.run(realtime=True)
    lat = np.array([j.latency for j in done])
    assert np.all(lat > 0)
    assert np.percentile(lat, 99) >= np.percentile(lat, 50)
    assert srv.site_updates == sum(j.sweeps * j.sites for j in done)


def test_copy_spans_copies_every_span():
    """cm_copy_spans, the one launch per slice that moves each job's part of
    the pool's series into the job's buffers: spans of 0, 1, a few and many
    4-byte words, at unrelated offsets of two buffers."""
    import torch

    from critmc_amd._lib import check, lib
    from critmc_amd.device import ptr, stream

    dev = torch.device("cuda", 0)
    g = torch.Generator().manual_seed(3)
    src = torch.randint(-(1 << 31), (1 << 31) - 1, (9000,), dtype=torch.int32, generator=g)
    dst0 = torch.randint(-(1 << 31), (1 << 31) - 1, (9000,), dtype=torch.int32, generator=g)
    src_d, dst_d = src.to(dev), dst0.to(dev)
    spans = [(0, 0, 0), (5, 17, 1), (100, 3, 7), (300, 1000, 256), (1000, 2000, 257),
             (3000, 4000, 5000), (8999, 8999, 1), (20, 40, 33)]
    table = torch.tensor([[dst_d.data_ptr() + 4 * d, src_d.data_ptr() + 4 * s, 4 * n]
                          for d, s, n in spans], dtype=torch.int64).to(dev)
    check(lib().cm_copy_spans(ptr(table), len(spans), stream()), "cm_copy_spans")
    want = dst0.clone()
    for d, s, n in spans:
        want[d: d + n] = src[s: s + n]
    assert torch.equal(dst_d.cpu(), want)


def test_pools_that_grow_while_changing_streams_keep_every_result():
    """Twenty jobs of one pool arriving over time: the pool grows from 8 to 16
    to 32 slots (its state copied to new buffers) while its slices move between
    eight streams; every job's series equals the job run alone."""
    jobs = [Job(m, 16, b, 150 + 37 * k, seed=900 + k, arrival=0.0005 * k)
            for k in range(20) for m, b in [(("ising", "potts", "xy")[k % 3],
                                             (0.44, 1.42, 1.1)[k % 3])]]
    alone = _alone(jobs, budget_bytes=1e9, slice_sweeps=64)
    srv = Server(budget_bytes=1e9, slice_sweeps=64, n_streams=8)
    probes = [srv.submit(_clone(j, arrival=j.arrival)) for j in jobs]
    # three pools of about seven jobs each would not grow past 8 slots: twenty
    # more Ising jobs, all there at the start, take the Ising pool to 32
    extra = [srv.submit(Job("ising", 16, 0.43, 400 + 20 * k, seed=950 + k)) for k in range(20)]
    done = srv.run(realtime=True)
    assert len(done) == len(jobs) + len(extra)
    assert srv.pools[("ising", 16, 0)].cap >= 32
    for res, job in zip(alone, probes):
        got = job.results()
        for k in res:
            assert np.array_equal(got[k], res[k]), (job.model, job.seed, k)
