"""Measurement loops replayed as hipGraphs give the eager launches' results bit for
bit: Ising and Potts cluster schedules (with and without local rows, several SW
and Wolff repetitions), and the neighbour-table path, over several calls in a row
with eager work and synchronisation between them."""

import numpy as np
import pytest
import torch

from critmc_amd._lib import set_option
from critmc_amd import square, triangular
from critmc_amd.kernels import ising as KI
from critmc_amd.kernels import potts as KP
from critmc_amd.kernels import table as T
from critmc_amd.kernels.chains import Chains

pytestmark = pytest.mark.gpu


def _with_options(fn, **opts):
    old = {k: set_option(k, v) for k, v in opts.items()}
    try:
        return fn()
    finally:
        for k, v in old.items():
            set_option(k, v)


def _ising_calls(shape, sched, q=0):
    dev = torch.device("cuda", 0)
    R = 5
    rng = np.random.default_rng(31 + q)
    keys = rng.integers(0, 1 << 32, (R, 2), dtype=np.uint64)
    N = int(np.prod(shape))
    s0 = (2 * rng.integers(0, 2, (R, N)) - 1).astype(np.int8) if q == 0 else \
        rng.integers(0, q, (R, N)).astype(np.int8)
    betas = rng.uniform(0.3, 0.6, R) if q == 0 else rng.uniform(0.8, 1.6, R)

    def run():
        chains = Chains(keys, np.arange(R), None, dev)
        s = torch.from_numpy(s0).to(dev)
        out = []
        if q == 0:
            e, m = KI.measure(s, shape, R)
            ws = KI.ClusterWorkspace(shape, R, dev)
            for n in (40, 1, 23):
                res = KI.run_cluster(s, shape, betas, 0.0, chains, sched, n, e, m, ws=ws)
                torch.cuda.synchronize()
                KI.measure(s, shape, R)  # eager work between the calls
                out.append([t.cpu().numpy() for t in res])
        else:
            e, c = KP.measure(s, shape, q, R)
            ws = KP.ClusterWorkspace(shape, R, q, dev)
            for n in (40, 1, 23):
                res = KP.run_cluster(s, shape, q, betas, chains, sched, n, e, c, ws=ws)
                torch.cuda.synchronize()
                KP.measure(s, shape, q, R)
                out.append([t.cpu().numpy() for t in res])
        out.append([s.cpu().numpy(), chains.sweeps_d.cpu().numpy(), chains.sweeps_h.copy()])
        return out

    return run


def _table_calls(lat, sched, q):
    dev = torch.device("cuda", 0)
    R = 70
    rng = np.random.default_rng(7 + q)
    keys = rng.integers(0, 1 << 32, (R, 2), dtype=np.uint64)
    s0 = (2 * rng.integers(0, 2, (R, lat.n_sites)) - 1).astype(np.int8) if q == 0 else \
        rng.integers(0, q, (R, lat.n_sites)).astype(np.int8)
    betas = rng.uniform(0.2, 0.5, R) if q == 0 else rng.uniform(0.7, 1.3, R)

    def run():
        chains = Chains(keys, np.arange(R), None, dev)
        s = torch.from_numpy(s0).to(dev)
        nbr = T.neighbours(lat, dev)
        e, mc = T.measure(s, nbr, R, q)
        ws = T.Workspace(lat.n_sites, R, q, dev)
        out = []
        for n in (33, 2, 17):
            if q == 0:
                res = T.ising_run(s, nbr, betas, 0.0, chains, sched, n, e, mc, ws=ws)
            else:
                res = T.potts_run(s, nbr, q, betas, chains, sched, n, e, mc, ws=ws)
            torch.cuda.synchronize()
            T.measure(s, nbr, R, q)
            out.append([t.cpu().numpy() for t in res])
        out.append([s.cpu().numpy(), chains.sweeps_d.cpu().numpy()])
        return out

    return run


CASES = {
    "ising-sw": _ising_calls((16, 16), [[4, 2]]),
    "ising-wolff+metropolis": _ising_calls((32, 32), [[3, 3], [0, 1]]),
    "ising-heatbath+sw-3d": _ising_calls((8, 8, 8), [[2, 1], [4, 1]]),
    "potts-sw+heatbath": _ising_calls((16, 16), [[4, 1], [2, 1]], q=3),
    "potts-wolff": _ising_calls((24, 16), [[3, 4]], q=10),
    "table-ising": _table_calls(triangular(9, 7), [[1, 1], [4, 1], [3, 2], [2, 1]], 0),
    "table-potts": _table_calls(square(7, 5), [[0, 1], [3, 1], [4, 1]], 4),
}


@pytest.mark.parametrize("case", list(CASES))
def test_graph_replays_equal_eager_launches(case):
    """Multi-launch path (no LDS form): graph replays against eager launches."""
    run = CASES[case]
    print(f"graph vs eager: {case}", flush=True)
    eager = _with_options(run, graph_min_meas=0, cluster_lds_max_sites=0)
    graph = _with_options(run, graph_min_meas=1, cluster_lds_max_sites=0)
    for k, (a, b) in enumerate(zip(eager, graph)):
        for x, y in zip(a, b):
            assert np.array_equal(x, y), (case, k)


LDS_CASES = {
    "ising-sw": _ising_calls((16, 16), [[4, 2]]),
    "ising-wolff-sw": _ising_calls((64, 64), [[3, 5], [4, 1]]),
    "ising-wolff-3d": _ising_calls((8, 8, 8), [[3, 3]]),
    "ising-sw-chain": _ising_calls((40,), [[4, 1]]),
    "potts-sw": _ising_calls((16, 16), [[4, 1]], q=3),
    "potts-wolff+sw": _ising_calls((24, 16), [[3, 4], [4, 1]], q=10),
    "table-ising": _table_calls(triangular(9, 7), [[4, 1], [3, 2]], 0),
    "table-potts": _table_calls(square(7, 5), [[3, 1], [4, 1]], 4),
}


@pytest.mark.parametrize("case", list(LDS_CASES))
def test_one_launch_lds_form_equals_multi_launch_kernels(case):
    """Cluster-only schedules: every update of the call in one launch with the replica
    in LDS, against the multi-launch kernels."""
    run = LDS_CASES[case]
    print(f"LDS form vs multi-launch: {case}", flush=True)
    multi = _with_options(run, cluster_lds_max_sites=0)
    lds = _with_options(run, cluster_lds_max_sites=8192)
    for k, (a, b) in enumerate(zip(multi, lds)):
        for x, y in zip(a, b):
            assert np.array_equal(x, y), (case, k)
