"""The FP64 replay of O(n) chains in the reference's own site orders on lattices
given by their neighbour table (oracle/replay_seq_on.c), pinned against the
transfer operator of the ring (oracle/exact.py) on an odd ring of 7 sites, which
has no checkerboard: every update code and mixes, n = 2 and 3; plus
bookkeeping, chunk invariance and a planted error."""

import numpy as np
import pytest

import oracle
from critmc_amd import chain, triangular
from oracle import exact
from oracle.stats import batch_mean_error

SCHEDULES = [[[0, 1]], [[1, 1]], [[2, 1]], [[5, 1], [0, 1]], [[3, 2]], [[4, 1]],
             [[1, 1], [4, 1], [5, 1]]]


def _start(N, n, seed):
    s = np.random.default_rng(seed).standard_normal((N, n))
    return s / np.linalg.norm(s, axis=1, keepdims=True)


@pytest.mark.parametrize("n", [2, 3])
@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: "+".join(str(r[0]) for r in s))
def test_ring_of_7_matches_transfer_operator(n, sched):
    lat, beta = chain(7), 0.8
    rep = oracle.SeqOnReplay(_start(7, n, 1), lat.nbr, n, beta, 1.0, (3, 4))
    rep.run(sched, 1000)
    out = rep.run(sched, 30_000)
    mean, err = batch_mean_error(out["E"] / 7)
    assert abs(mean - exact.on_ring_energy(beta, 7, n)) < 4.5 * err
    # the last measurement describes the final, renormalised state
    s = rep.s
    np.testing.assert_allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-14)
    assert out["E"][-1] == pytest.approx(-np.sum(s[lat.bonds[:, 0]] * s[lat.bonds[:, 1]]), abs=1e-12)
    np.testing.assert_allclose(out["M"][-1], s.sum(axis=0), atol=1e-12)


def test_a_wrong_temperature_is_rejected():
    lat = chain(7)
    rep = oracle.SeqOnReplay(_start(7, 2, 2), lat.nbr, 2, 0.8, 1.0, (5, 6))
    out = rep.run([[2, 1]], 40_000)
    mean, err = batch_mean_error(out["E"] / 7)
    assert abs(mean - exact.on_ring_energy(0.9, 7, 2)) > 6 * err


@pytest.mark.parametrize("n", [2, 3])
def test_chunks_reproduce_one_run(n):
    lat = triangular(4, 3)
    sched = [[1, 2], [2, 1], [3, 1], [0, 1], [4, 1], [5, 1]]
    a = oracle.SeqOnReplay(_start(lat.n_sites, n, 7), lat.nbr, n, 0.7, 0.8, (2, 8), 3)
    b = oracle.SeqOnReplay(_start(lat.n_sites, n, 7), lat.nbr, n, 0.7, 0.8, (2, 8), 3)
    whole = a.run(sched, 20)
    parts = [b.run(sched, k) for k in (7, 1, 12)]
    for key in whole:
        np.testing.assert_array_equal(whole[key], np.concatenate([p[key] for p in parts]))
    np.testing.assert_array_equal(a.s, b.s)
    assert a.sweep[0] == b.sweep[0] == 20 * 7


def test_stiffness_sums_per_axis():
    """c_a, s_a over the forward bonds of axis a, in the reference's definition
    (critmc/kernels/on.py:58-75), on the triangular lattice's three axes."""
    lat = triangular(5, 4)
    rep = oracle.SeqOnReplay(_start(lat.n_sites, 2, 9), lat.nbr, 2, 0.5, 1.0, (1, 2))
    out = rep.run([[0, 1]], 1)
    s = rep.s
    for a in range(3):
        b = lat.bonds[lat.bond_axis == a]
        si, sj = s[b[:, 0]], s[b[:, 1]]
        assert out["stiff_c"][0, a] == pytest.approx(np.sum(si[:, 0] * sj[:, 0] + si[:, 1] * sj[:, 1]), abs=1e-12)
        assert out["stiff_s"][0, a] == pytest.approx(np.sum(si[:, 0] * sj[:, 1] - si[:, 1] * sj[:, 0]), abs=1e-12)
