"""The CPU replay of the checkerboard Ising schedule, pinned without the GPU.

* thresholds: T = ceil(p 2^32) against multiprecision p for the reference
  tables (critmc/kernels/ising.py:47-65);
* the bit-plane uniform of a site against a direct transposition of the
  Philox words;
* incremental E_b / M against a recomputation, for 1D/2D/3D;
* chunk invariance (critmc tests/test_sim_chunks.py);
* pulls against brute-force enumeration (bipartite lattices: the reference's
  own 3x3 is not), and a planted mistake (wrong beta) that must fail.
"""

import mpmath as mp
import numpy as np
import pytest

from oracle import exact
from oracle.stats import batch_mean_error


@pytest.mark.parametrize("beta,h,z", [(0.4406868, 0.0, 4), (0.3, 0.4, 4), (1.7, -0.2, 6), (0.05, 0.0, 2)])
def test_thresholds_are_ceil_of_p(oracle_lib, beta, h, z):
    met, hb = oracle_lib.ising_thresholds(beta, h, z)
    mp.mp.dps = 40
    for a in range(-z, z + 1):
        for t in range(2):
            d_e = 2 * a + 2 * h * (2 * t - 1)
            p = mp.mpf(1) if d_e <= 0 else mp.exp(-mp.mpf(beta) * d_e)
            x = p * 2 ** 32
            assert 0 <= int(met[a + z, t]) - x < 1 + 1e-6
        p = 1 / (1 + mp.exp(-2 * mp.mpf(beta) * (a + h)))
        assert 0 <= int(hb[a + z]) - p * 2 ** 32 < 1 + 1e-6


def test_library_class_tables_use_oracle_thresholds(oracle_lib):
    from critmc_amd.kernels.ising import thresholds_table

    for beta, h, z in [(0.44, 0.0, 4), (0.3, 0.4, 4), (0.9, 0.1, 6)]:
        met, hb = oracle_lib.ising_thresholds(beta, h, z)
        tab = thresholds_table(beta, h, z)
        # Metropolis classes: count of aligned neighbours n, spin selector
        for k in range(tab[0]):
            n, sel = tab[1 + k] & 0xFF, tab[1 + k] >> 8
            a = 2 * n - z
            ts = [0, 1] if sel == 0 else [0] if sel == 1 else [1]
            for t in ts:
                assert int(met[a + z, t]) == int(tab[15 + k])
        listed = {(tab[1 + k] & 0xFF, tab[1 + k] >> 8) for k in range(tab[0])}
        for n in range(z + 1):
            for t in range(2):
                if int(met[2 * n, t]) < 2 ** 32:
                    assert (n, 0) in listed or (n, 1 + t) in listed
        hbt = tab[32:]
        for k in range(hbt[0]):
            n = hbt[1 + k]
            assert int(hb[2 * n]) == int(hbt[15 + k])
        assert hbt[0] == sum(int(hb[2 * n]) < 2 ** 32 for n in range(z + 1))


def test_site_uniform_is_bitplane_transpose(oracle_lib):
    key, stream, sweep, colour, group = (7, 9), 3, 12345, 1, 77
    words = []
    for q in range(8):
        ctr = [8 * group + q, sweep & 0xFFFFFFFF, (sweep >> 32) | (colour << 16) | (2 << 24), stream]
        words.extend(int(w) for w in oracle_lib.philox(ctr, key))
    for b in (0, 5, 31):
        x = 0
        for j in range(32):
            x = (x << 1) | ((words[j] >> b) & 1)
        got = oracle_lib.lib().orc_ising_site_uniform(key[0], key[1], stream, sweep, colour, group, b)
        assert got == x


@pytest.mark.parametrize("shape", [(8,), (6, 4), (64, 4), (128, 2), (4, 4, 6)])
@pytest.mark.parametrize("sched", [[[0, 1]], [[2, 1]], [[0, 2], [2, 1]]])
def test_incremental_bookkeeping_matches_recount(oracle_lib, shape, sched):
    n = int(np.prod(shape))
    s0 = oracle_lib.ising_init(n, (5, 6), 1)
    r = oracle_lib.IsingReplay(s0, shape, 0.5, 0.2, (11, 12), 4)
    out = r.run(sched, 25)
    assert oracle_lib.ising_energy(r.s, shape) == (int(out["Eb"][-1]), int(out["M"][-1]))
    assert np.all(out["accepted"] >= 0) and np.all(out["accepted"] <= n * sum(c for _, c in sched))
    assert int(r.sweep[0]) == 25 * sum(c for _, c in sched)


def test_chunks_reproduce_the_stream(oracle_lib):
    s0 = oracle_lib.ising_init(64 * 8, (1, 2), 0)
    full = oracle_lib.IsingReplay(s0, (64, 8), 0.44, 0.0, (3, 4))
    a = full.run([[0, 1], [2, 1]], 50)
    part = oracle_lib.IsingReplay(s0, (64, 8), 0.44, 0.0, (3, 4))
    b1 = part.run([[0, 1], [2, 1]], 17)
    b2 = part.run([[0, 1], [2, 1]], 33)
    for k in a:
        assert np.array_equal(a[k], np.concatenate([b1[k], b2[k]]))
    assert np.array_equal(full.s, part.s)


def _pull(oracle_lib, shape, beta, h, sched, n_meas, seed, beta_true=None):
    from critmc_amd import hypercubic

    lat = hypercubic(shape)
    n = lat.n_sites
    s0 = oracle_lib.ising_init(n, (seed, 99), 0)
    r = oracle_lib.IsingReplay(s0, shape, beta, h, (seed, 1))
    r.run(sched, 2000)
    out = r.run(sched, n_meas)
    bt = beta if beta_true is None else beta_true
    E = out["Eb"] - h * out["M"]
    zs = []
    for series, fn in ((E, lambda e, m: e - h * m), (out["M"].astype(float) ** 2 / n ** 2,
                                                      lambda e, m: (m / n) ** 2)):
        ex = exact.ising_average(lat.bonds, n, bt, h, fn)
        mean, err = batch_mean_error(series)
        zs.append((mean - ex) / err)
    return zs


@pytest.mark.parametrize("shape,beta,h,sched", [
    ((4, 4), 0.35, 0.0, [[0, 1]]),
    ((4, 4), 0.35, 0.0, [[2, 1]]),
    ((2, 6), 0.5, 0.0, [[0, 1]]),
    ((4, 4), 0.3, 0.4, [[0, 1]]),
    ((4, 4), 0.3, 0.4, [[2, 1]]),
    ((8,), 0.8, 0.0, [[0, 1], [2, 1]]),
    ((2, 2, 4), 0.3, 0.0, [[0, 1]]),
    ((4, 4), 0.44, 0.0, [[4, 1]]),
    ((4, 4), 0.44, 0.0, [[3, 2]]),
    ((2, 6), 0.6, 0.0, [[0, 1], [4, 1]]),
    ((2, 2, 4), 0.25, 0.0, [[3, 1], [2, 1]]),
])
def test_replay_matches_enumeration(oracle_lib, shape, beta, h, sched):
    zs = _pull(oracle_lib, shape, beta, h, sched, 200_000, seed=sum(shape))
    assert all(abs(z) < 4.5 for z in zs), zs


def test_planted_wrong_beta_is_caught(oracle_lib):
    zs = _pull(oracle_lib, (4, 4), 0.35 * 0.9, 0.0, [[0, 1]], 200_000, seed=8, beta_true=0.35)
    assert max(abs(z) for z in zs) > 4.5, zs


@pytest.mark.parametrize("sched", [[[4, 1]], [[3, 3]]])
def test_planted_wrong_beta_is_caught_by_cluster_updates(oracle_lib, sched):
    zs = _pull(oracle_lib, (4, 4), 0.44 * 0.9, 0.0, sched, 200_000, seed=8, beta_true=0.44)
    assert max(abs(z) for z in zs) > 4.5, zs


@pytest.mark.parametrize("shape", [(8, 8), (6, 4, 2), (16,)])
def test_cluster_diagnostics_and_bookkeeping(oracle_lib, shape):
    n = int(np.prod(shape))
    r = oracle_lib.IsingReplay(oracle_lib.ising_init(n, (2, 3), 0), shape, 0.4, 0.0, (7, 8))
    out = r.run([[0, 1], [3, 2], [4, 1]], 40)
    aux = out["aux"]
    assert np.array_equal(aux[:, 0], out["accepted"])
    assert np.all(aux[:, 1] >= 2) and np.all(aux[:, 1] <= 2 * n)  # two Wolff clusters
    assert np.all(aux[:, 3] >= 1) and np.all(aux[:, 3] <= n)      # SW cluster count
    assert np.all(aux[:, 2] >= n) and np.all(aux[:, 2] <= n * n)  # sum |C|^2 over a partition
    assert oracle_lib.ising_energy(r.s, shape) == (int(out["Eb"][-1]), int(out["M"][-1]))
    assert int(r.sweep[0]) == 40 * 4


def test_sw_at_infinite_temperature_flips_isolated_sites(oracle_lib):
    """beta = 0 activates no bond: n singleton clusters, each flipped with
    probability 1/2 by its own Philox bit."""
    shape, n = (8, 8), 64
    r = oracle_lib.IsingReplay(np.ones(n, np.int8), shape, 0.0, 0.0, (4, 4))
    out = r.run([[4, 1]], 400)
    assert np.all(out["aux"][:, 3] == n) and np.all(out["aux"][:, 2] == n)
    frac_down = (n - out["M"]) / (2 * n)
    assert abs(frac_down.mean() - 0.5) < 0.01
