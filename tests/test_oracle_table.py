"""The sequential schedule replay on lattices given by their neighbour table
(oracle/replay_table.c): the reference's own site orders on the non-bipartite
lattices its kernel tests use (critmc tests/test_kernels.py:21-52: square 3x3
and triangular 3x3), pinned against brute-force enumeration for every update
code, Ising and Potts, plus bookkeeping, chunk invariance and a planted error."""

import numpy as np
import pytest

import oracle
from critmc_amd import square, triangular
from oracle import exact
from oracle.stats import batch_mean_error

CODES = {"metropolis": 0, "metropolis-random": 1, "heatbath": 2, "wolff": 3, "sw": 4}
SCHEDULES = [[[0, 1]], [[1, 1]], [[2, 1]], [[3, 2]], [[4, 1]], [[1, 1], [3, 1], [2, 1]]]


def _start(lat, q, seed):
    rng = np.random.default_rng(seed)
    if q == 0:
        return (2 * rng.integers(0, 2, lat.n_sites) - 1).astype(np.int8)
    return rng.integers(0, q, lat.n_sites).astype(np.int8)


def _pull(series, exact_value):
    mean, err = batch_mean_error(series)
    return (mean - exact_value) / err


@pytest.mark.parametrize("lat", [square(3), triangular(3)], ids=["square3", "triangular3"])
@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: "+".join(str(r[0]) for r in s))
def test_ising_matches_enumeration(lat, sched):
    beta, n = 0.3, lat.n_sites
    rep = oracle.TableReplay(_start(lat, 0, 1), lat.nbr, 0, beta, 0.0, (5, 7))
    rep.run(sched, 2000)
    out = rep.run(sched, 150_000)
    e_exact = exact.ising_average(lat.bonds, n, beta, 0.0, lambda e, m: e)
    m2_exact = exact.ising_average(lat.bonds, n, beta, 0.0, lambda e, m: (m / n) ** 2)
    assert abs(_pull(out["E"].astype(float), e_exact)) < 4.5
    assert abs(_pull((out["M"] / n) ** 2, m2_exact)) < 4.5
    # running totals equal a recount of the final state
    s = rep.s.astype(np.int64)
    assert rep.e[0] == -(s[lat.bonds[:, 0]] * s[lat.bonds[:, 1]]).sum()
    assert rep.mc[0] == s.sum()
    if sched[0][0] in (0, 1, 2):
        assert out["aux"][:, 0].sum() > 0


def test_ising_in_a_field_matches_enumeration():
    lat, beta, h = triangular(3), 0.3, 0.35
    n = lat.n_sites
    for sched in ([[0, 1]], [[1, 1]], [[2, 1]]):
        rep = oracle.TableReplay(_start(lat, 0, 2), lat.nbr, 0, beta, h, (3, 4))
        out = rep.run(sched, 150_000)
        m_exact = exact.ising_average(lat.bonds, n, beta, h, lambda e, m: m)
        assert abs(_pull(out["M"].astype(float), m_exact)) < 4.5


@pytest.mark.parametrize("lat", [square(3), triangular(3)], ids=["square3", "triangular3"])
@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: "+".join(str(r[0]) for r in s))
def test_potts_q3_matches_enumeration(lat, sched):
    beta, q, n = 0.9, 3, lat.n_sites
    rep = oracle.TableReplay(_start(lat, q, 3), lat.nbr, q, beta, 0.0, (9, 1))
    rep.run(sched, 2000)
    out = rep.run(sched, 100_000)
    e_exact = exact.potts_average(lat.bonds, n, q, beta, lambda e, c: e)
    m2 = lambda c: (q * np.sum((c / n) ** 2, axis=1) - 1) / (q - 1)  # noqa: E731
    m2_exact = exact.potts_average(lat.bonds, n, q, beta, lambda e, c: m2(c))
    assert abs(_pull(out["E"].astype(float), e_exact)) < 4.5
    assert abs(_pull(m2(out["counts"].astype(float)), m2_exact)) < 4.5
    s = rep.s
    assert rep.e[0] == -(s[lat.bonds[:, 0]] == s[lat.bonds[:, 1]]).sum()
    np.testing.assert_array_equal(rep.mc, np.bincount(s, minlength=q))


def test_a_wrong_temperature_is_rejected():
    lat = triangular(3)
    n = lat.n_sites
    rep = oracle.TableReplay(_start(lat, 0, 4), lat.nbr, 0, 0.3, 0.0, (1, 1))
    out = rep.run([[1, 1]], 150_000)
    e_wrong = exact.ising_average(lat.bonds, n, 0.33, 0.0, lambda e, m: e)
    assert abs(_pull(out["E"].astype(float), e_wrong)) > 6


@pytest.mark.parametrize("q", [0, 4])
def test_chunks_reproduce_one_run(q):
    lat = triangular(4, 3)
    sched = [[1, 2], [2, 1], [3, 1], [0, 1], [4, 1]]
    a = oracle.TableReplay(_start(lat, q, 5), lat.nbr, q, 0.7, 0.1 if q == 0 else 0.0, (2, 8), 3)
    b = oracle.TableReplay(_start(lat, q, 5), lat.nbr, q, 0.7, 0.1 if q == 0 else 0.0, (2, 8), 3)
    whole = a.run(sched, 30)
    parts = [b.run(sched, k) for k in (7, 1, 22)]
    for key in whole:
        np.testing.assert_array_equal(whole[key], np.concatenate([p[key] for p in parts]))
    np.testing.assert_array_equal(a.s, b.s)
    assert a.sweep[0] == b.sweep[0] == 30 * 6
