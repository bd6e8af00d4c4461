"""FP64 replay of checkerboard O(n) sweeps (critmc/kernels/on.py:112-183).

Pinned by: O(n) rings against the transfer operator (critmc tests/
test_kernels.py:95-109, here with checkerboard order), over-relaxation
invariants (energy conserved, twice = identity, zero field = no change,
critmc tests/test_overrelax_xy.py), and a planted wrong beta.  The heat bath
(critmc/kernels/on.py:186-261) is pinned by its conditional law: on a 2-site
ring each update draws the relative angle afresh, so the samples are i.i.d.
von Mises (XY) or exponential in cos (Heisenberg) with kappa = 2 beta, checked
by Kolmogorov-Smirnov tests that also reject a kappa 15 % off.
"""

import numpy as np
import pytest

from oracle import exact
from oracle.stats import batch_mean_error


def _random_state(N, n, seed):
    s = np.random.default_rng(seed).standard_normal((N, n))
    return s / np.linalg.norm(s, axis=1, keepdims=True)


@pytest.mark.parametrize("n", [2, 3])
def test_overrelax_invariants(oracle_lib, n):
    from critmc_amd import square

    lat = square(8)
    s0 = _random_state(64, n, 1)
    rep = oracle_lib.OnReplay(s0, (8, 8), 1.0, 1.0, (1, 1))
    e0 = -np.sum(s0[lat.bonds[:, 0]] * s0[lat.bonds[:, 1]])
    rep.half_sweep(5, 0)
    e1 = -np.sum(rep.s[lat.bonds[:, 0]] * rep.s[lat.bonds[:, 1]])
    assert e1 == pytest.approx(e0, abs=1e-12)
    rep.half_sweep(5, 0)  # the reflection is an involution when the field is fixed
    assert np.allclose(rep.s, s0, atol=1e-12)
    z = np.zeros((4, n))
    z[:, 0] = 1.0
    z[1, 0] = z[3, 0] = -1.0  # chain of 4: both neighbours of sites 0, 2 cancel
    rep = oracle_lib.OnReplay(z, (4,), 1.0, 1.0, (1, 1))
    rep.half_sweep(5, 0)
    assert np.array_equal(rep.s, z)


def _ring(oracle_lib, n_comp, sched, beta, n_meas, seed, beta_true=None):
    rep = oracle_lib.OnReplay(_random_state(8, n_comp, seed), (8,), beta, 1.0, (seed, 3))
    for _ in range(10):  # step tuning as critmc/sim.py:226-229
        out = rep.run(sched, 500)
        nm = sum(c for code, c in sched if code == 0)
        if nm:
            rate = out["accepted"].sum() / (nm * 8 * 500)
            rep.step = float(np.clip(rep.step * np.clip(rate / 0.5, 0.5, 2.0), 1e-3, 10.0))
    out = rep.run(sched, n_meas)
    bt = beta if beta_true is None else beta_true
    mean, err = batch_mean_error(out["E"] / 8)
    return (mean - exact.on_ring_energy(bt, 8, n_comp)) / err


@pytest.mark.parametrize("n_comp", [2, 3])
@pytest.mark.parametrize("sched", [[[0, 1]], [[0, 1], [5, 3]], [[2, 1]], [[2, 1], [5, 2]],
                                   [[4, 1]], [[3, 2]], [[5, 2], [4, 1]]])
def test_ring_matches_transfer_operator(oracle_lib, n_comp, sched):
    z = _ring(oracle_lib, n_comp, sched, 1.1, 100_000, seed=5 + n_comp)
    assert abs(z) < 4.5, z


def test_ring_planted_wrong_beta_is_caught(oracle_lib):
    z = _ring(oracle_lib, 2, [[0, 1]], 1.1 * 0.85, 100_000, seed=3, beta_true=1.1)
    assert abs(z) > 4.5, z


def _pair_samples(oracle_lib, n_comp, beta, n, seed):
    """Relative angle (XY) or cosine (Heisenberg) of a 2-site ring after each heat-bath sweep."""
    rep = oracle_lib.OnReplay(_random_state(2, n_comp, seed), (2,), beta, 1.0, (seed, 11))
    out = np.empty(n)
    for t in range(n):
        rep.run([[2, 1]], 1)
        a, b = rep.s
        if n_comp == 2:
            out[t] = np.arctan2(a[0] * b[1] - a[1] * b[0], a @ b)
        else:
            out[t] = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
    return out


@pytest.mark.parametrize("beta", [0.7, 1.6, 1e-10])
def test_xy_heatbath_draws_von_mises(oracle_lib, beta):
    from scipy import stats

    kappa = 2.0 * beta
    x = _pair_samples(oracle_lib, 2, beta, 20_000, seed=int(beta * 100) + 1)
    assert stats.kstest(x, stats.vonmises(kappa).cdf).pvalue > 1e-3
    if beta > 0.1:  # the test has the power to see a wrong kappa
        assert stats.kstest(x, stats.vonmises(0.85 * kappa).cdf).pvalue < 1e-6
    else:
        assert stats.kstest(x, stats.uniform(-np.pi, 2 * np.pi).cdf).pvalue > 1e-3


@pytest.mark.parametrize("beta", [0.5, 1.3])
def test_heisenberg_heatbath_draws_exponential_cosine(oracle_lib, beta):
    from scipy import stats

    def cdf(k):
        return lambda u: np.expm1(k * (u + 1)) / np.expm1(2 * k)

    kappa = 2.0 * beta
    u = _pair_samples(oracle_lib, 3, beta, 20_000, seed=int(beta * 10) + 2)
    assert stats.kstest(u, cdf(kappa)).pvalue > 1e-3
    assert stats.kstest(u, cdf(0.85 * kappa)).pvalue < 1e-6


@pytest.mark.parametrize("sched", [[[4, 1]], [[3, 2]]])
def test_ring_cluster_planted_wrong_beta_is_caught(oracle_lib, sched):
    z = _ring(oracle_lib, 2, sched, 1.1 * 0.85, 100_000, seed=4, beta_true=1.1)
    assert abs(z) > 4.5, z


@pytest.mark.parametrize("n", [2, 3])
def test_cluster_reflections_keep_unit_length_and_count(oracle_lib, n):
    rep = oracle_lib.OnReplay(_random_state(64, n, 9), (8, 8), 0.9, 1.0, (3, 5))
    out = rep.run([[4, 1], [3, 3]], 20)
    aux = out["aux"]
    assert np.all(aux[:, 1] >= 3) and np.all(aux[:, 1] <= 3 * 64)
    assert np.all(aux[:, 3] >= 1) and np.all(aux[:, 2] >= 64) and np.all(aux[:, 2] <= 64 * 64)
    assert np.allclose(np.linalg.norm(rep.s, axis=1), 1.0, atol=1e-12)
    assert int(rep.sweep[0]) == 20 * 4


def test_sw_at_infinite_temperature_reflects_half_the_sites(oracle_lib):
    """beta = 0 activates no bond: every site is its own cluster and reflects
    with probability 1/2, which flips the sign of its projection on r."""
    s0 = _random_state(256, 2, 2)
    rep = oracle_lib.OnReplay(s0, (16, 16), 0.0, 1.0, (8, 8))
    out = rep.run([[4, 1]], 1)
    assert out["aux"][0, 3] == 256 and out["aux"][0, 2] == 256
    moved = np.linalg.norm(rep.s - s0, axis=1) > 1e-12
    assert 0.4 < moved.mean() < 0.6
