"""CPU replay of the checkerboard Potts schedule (critmc/kernels/potts.py:44-103).

Pinned by: counts / energy bookkeeping against a recount, chunk invariance,
pulls against brute-force enumeration (q=3 on 2x4 and 4x4 would be too many
states; 2x4 and 2x6 are bipartite and enumerable), q=2 Potts = Ising at half
coupling, and a planted wrong beta that must fail.
"""

import numpy as np
import pytest

from oracle import exact
from oracle.stats import batch_mean_error


@pytest.mark.parametrize("shape,q", [((6, 4), 3), ((8,), 5), ((4, 4, 2), 10), ((2, 2), 2)])
@pytest.mark.parametrize("sched", [[[0, 1]], [[2, 1]], [[0, 1], [2, 2]]])
def test_potts_bookkeeping(oracle_lib, shape, q, sched):
    n = int(np.prod(shape))
    s0 = oracle_lib.potts_init(n, q, (3, 4), 0)
    r = oracle_lib.PottsReplay(s0, shape, q, 0.9, (5, 6), 2)
    out = r.run(sched, 30)
    e, c = oracle_lib.potts_energy(r.s, shape, q)
    assert e == out["E"][-1]
    assert np.array_equal(c, out["counts"][-1])
    assert np.array_equal(c, np.bincount(r.s, minlength=q))
    assert np.all(out["counts"].sum(axis=1) == n)


def test_potts_chunks(oracle_lib):
    s0 = oracle_lib.potts_init(64, 4, (1, 1))
    a = oracle_lib.PottsReplay(s0, (8, 8), 4, 1.0, (2, 2))
    x = a.run([[0, 1], [2, 1]], 40)
    b = oracle_lib.PottsReplay(s0, (8, 8), 4, 1.0, (2, 2))
    y1 = b.run([[0, 1], [2, 1]], 13)
    y2 = b.run([[0, 1], [2, 1]], 27)
    for k in x:
        assert np.array_equal(x[k], np.concatenate([y1[k], y2[k]]))


def _pull(oracle_lib, shape, q, beta, sched, n_meas, seed, beta_true=None):
    from critmc_amd import hypercubic

    lat = hypercubic(shape)
    n = lat.n_sites
    r = oracle_lib.PottsReplay(oracle_lib.potts_init(n, q, (seed, 0)), shape, q, beta, (seed, 1))
    r.run(sched, 2000)
    out = r.run(sched, n_meas)
    bt = beta if beta_true is None else beta_true
    zs = []
    e_ex = exact.potts_average(lat.bonds, n, q, bt, lambda e, c: e)
    m2 = lambda e, c: (q * np.sum((c / n) ** 2, axis=1) - 1) / (q - 1)  # noqa: E731
    m2_ex = exact.potts_average(lat.bonds, n, q, bt, m2)
    for series, ex in ((out["E"], e_ex), (m2(None, out["counts"]), m2_ex)):
        mean, err = batch_mean_error(series)
        zs.append((mean - ex) / err)
    return zs


@pytest.mark.parametrize("shape,q,beta,sched", [
    ((2, 4), 3, 0.8, [[0, 1]]),
    ((2, 4), 3, 0.8, [[2, 1]]),
    ((2, 6), 3, 1.1, [[0, 1], [2, 1]]),
    ((2, 4), 3, 1.0, [[4, 1]]),
    ((2, 4), 3, 1.0, [[3, 2]]),
    ((2, 6), 3, 1.2, [[2, 1], [4, 1], [3, 1]]),
    ((2, 2, 2), 4, 0.9, [[4, 1]]),
    ((2, 2), 10, 1.4, [[0, 1]]),
    ((2, 2), 10, 1.4, [[2, 1]]),
    ((6,), 4, 1.0, [[0, 1]]),
])
def test_potts_replay_matches_enumeration(oracle_lib, shape, q, beta, sched):
    zs = _pull(oracle_lib, shape, q, beta, sched, 150_000, seed=q + sum(shape))
    assert all(abs(z) < 4.5 for z in zs), zs


def test_potts_planted_wrong_beta_is_caught(oracle_lib):
    zs = _pull(oracle_lib, (2, 4), 3, 0.8 * 0.85, [[0, 1]], 150_000, seed=5, beta_true=0.8)
    assert max(abs(z) for z in zs) > 4.5, zs


@pytest.mark.parametrize("sched", [[[4, 1]], [[3, 3]]])
def test_potts_planted_wrong_beta_is_caught_by_cluster_updates(oracle_lib, sched):
    zs = _pull(oracle_lib, (2, 4), 3, 1.0 * 0.85, sched, 150_000, seed=6, beta_true=1.0)
    assert max(abs(z) for z in zs) > 4.5, zs


@pytest.mark.parametrize("shape,q", [((8, 8), 3), ((6, 4, 2), 10), ((12,), 2)])
def test_potts_cluster_bookkeeping(oracle_lib, shape, q):
    n = int(np.prod(shape))
    r = oracle_lib.PottsReplay(oracle_lib.potts_init(n, q, (8, 1), 0), shape, q, 1.1, (3, 9), 5)
    out = r.run([[0, 1], [3, 2], [4, 1]], 30)
    aux = out["aux"]
    assert np.array_equal(aux[:, 0], out["accepted"])
    assert np.all(aux[:, 1] >= 2) and np.all(aux[:, 1] <= 2 * n)
    assert np.all(aux[:, 3] >= 1) and np.all(aux[:, 2] >= n) and np.all(aux[:, 2] <= n * n)
    e, c = oracle_lib.potts_energy(r.s, shape, q)
    assert e == out["E"][-1] and np.array_equal(c, out["counts"][-1])
    assert int(r.sweep[0]) == 30 * 4
