"""Batched GPU flat-histogram walkers against the CPU replay and exact results.

Bit-exact: spins, level, ln g, visit histogram and move counter of every
walker after Wang-Landau blocks (halving and 1/t modes, windows, inactive
walkers), and the recorded level / order series of multicanonical runs, for
Ising and Potts.  Then the reference's drivers: WangLandau recovers the exact
4x4 Ising density of states and Multicanonical reweights to the exact energy.
"""

import numpy as np
import pytest
import torch

import oracle
from critmc_amd import Ising, Potts, hypercubic, square
from critmc_amd.flathist import (EnergyLevels, Multicanonical, MulticanonicalBatch, WangLandau,
                                 WangLandauBatch, _Walkers)
from oracle import exact

pytestmark = pytest.mark.gpu


def _setup(model, R, seed):
    lv = EnergyLevels.of(model)
    rng = np.random.default_rng(seed)
    n = model.n_sites
    states, idx, wins = [], [], []
    for r in range(R):
        if lv.kind == "ising":
            s = (2 * rng.integers(0, 2, n) - 1).astype(np.int8)
        else:
            s = rng.integers(0, lv.q, n).astype(np.int8)
        k = lv.index(model.energy(s))
        lo = max(0, k - int(rng.integers(0, lv.size // 2)))
        hi = min(lv.size - 1, k + int(rng.integers(1, lv.size // 2)))
        states.append(s)
        idx.append(k)
        wins.append((lo, hi))
    keys = rng.integers(0, 1 << 32, size=(R, 2), dtype=np.uint64)
    return lv, states, idx, wins, keys


@pytest.mark.parametrize("model", [Ising(square(6)), Potts(hypercubic((4, 4, 2)), 5),
                                   Ising(hypercubic((12,)))], ids=["ising2d", "potts3d", "chain"])
def test_wang_landau_blocks_are_bit_exact(model):
    R = 6
    lv, states, idx, wins, keys = _setup(model, R, 3)
    W = _Walkers(model, lv, states, idx, wins, keys)
    S = lv.size
    ln_g = torch.zeros((R, S), dtype=torch.float64, device=W.dev)
    hist = torch.zeros((R, S), dtype=torch.int64, device=W.dev)
    reps = [oracle.FlatWalker(states[r], model.lattice.shape, W.q, idx[r], *wins[r], keys[r])
            for r in range(R)]
    g_h = np.zeros((R, S))
    h_h = np.zeros((R, S), np.int64)
    ln_f = np.array([1.0, 0.5, 0.25, 1.0, 0.125, 1.0])
    mode = np.array([0, 0, 1, 0, 1, 0], np.int32)
    nl = np.array([0.0, 0.0, 7.0, 0.0, 5.0, 0.0])
    for block, active in ((500, [1] * R), (1700, [1, 0, 1, 1, 0, 1]), (300, [1] * R)):
        act = np.array(active, np.int32)
        W.wl(ln_g, hist, ln_f, mode, nl, act, block)
        for r in range(R):
            if act[r]:
                reps[r].wl(g_h[r], h_h[r], ln_f[r], bool(mode[r]), nl[r], block)
        assert np.array_equal(ln_g.cpu().numpy(), g_h)
        assert np.array_equal(hist.cpu().numpy(), h_h)
        assert np.array_equal(W.idx.cpu().numpy(), np.array([int(p.idx[0]) for p in reps]))
        assert np.array_equal(W.spins.cpu().numpy(), np.stack([p.s for p in reps]))
        assert np.array_equal(W.moves.cpu().numpy().view(np.uint64),
                              np.array([int(p.moves[0]) for p in reps], np.uint64))


@pytest.mark.parametrize("model", [Ising(square(6)), Potts(hypercubic((6, 4)), 4)],
                         ids=["ising", "potts"])
def test_multicanonical_is_bit_exact(model):
    R = 5
    lv, states, idx, wins, keys = _setup(model, R, 9)
    rng = np.random.default_rng(1)
    ln_g = rng.uniform(0, 20, size=lv.size)
    ln_g[rng.integers(0, lv.size, 3)] = np.inf  # forbidden levels
    ln_w = np.where(np.isfinite(ln_g), -ln_g, -np.inf)
    W = _Walkers(model, lv, states, idx, wins, keys)
    ln_w_d = torch.from_numpy(np.tile(ln_w, (R, 1))).to(W.dev)
    if W.q == 0:
        order = np.array([[int(s.astype(np.int64).sum())] for s in states], np.int32)
    else:
        order = np.stack([np.bincount(s, minlength=W.q) for s in states]).astype(np.int32)
    order_d = torch.from_numpy(order).to(W.dev)
    hist = torch.zeros((R, lv.size), dtype=torch.int64, device=W.dev)
    idx_out, ord_out = W.muca(order_d, ln_w_d, hist, 40, 37)
    for r in range(R):
        rep = oracle.FlatWalker(states[r], model.lattice.shape, W.q, idx[r], *wins[r], keys[r])
        h = np.zeros(lv.size, np.int64)
        i_ref, o_ref = rep.muca(ln_w.copy(), h, 40, 37)
        assert np.array_equal(idx_out[r].cpu().numpy(), i_ref)
        assert np.array_equal(ord_out[r].cpu().numpy(), o_ref)
        assert np.array_equal(hist[r].cpu().numpy(), h)
        assert np.array_equal(W.spins[r].cpu().numpy(), rep.s)
        assert np.array_equal(order_d[r].cpu().numpy(), rep.order)


def test_wang_landau_and_multicanonical_reproduce_exact_results():
    lat = square(4)
    model = Ising(lat)
    res = WangLandau(model, seed=4).run(1e-5, check_sweeps=50)
    e_b, _ = exact.ising_states(lat.bonds, 16)
    g = np.bincount(((e_b + 32) // 4).astype(int), minlength=17).astype(float)
    v = g > 0
    assert np.array_equal(res.visited, v)
    est = res.normalized()[v]
    assert np.max(np.abs(est - np.log(g[v]))) < 0.06
    can = res.canonical([0.3, 0.44])
    for k, b in enumerate((0.3, 0.44)):
        want = exact.ising_average(lat.bonds, 16, b, 0.0, lambda e, m: e) / 16
        assert can["u"][k] == pytest.approx(want, abs=0.01)
    mu = Multicanonical(model, res, seed=5).run(100_000, sweeps_per_meas=1, n_therm=100)
    for b in (0.3, 0.44):
        want = exact.ising_average(lat.bonds, 16, b, 0.0, lambda e, m: e) / 16
        assert mu.canonical(b)["u"][0] == pytest.approx(want, abs=0.02)


def test_many_walkers_reproduce_exact_energy():
    """256 independent walkers on the 8x8 torus: their canonical energies from ln g
    scatter around Kaufman's exact u(beta) by their own standard error."""
    model = Ising(square(8))
    batch = WangLandauBatch(model, seeds=range(256))
    runs = batch.run(1e-4, check_sweeps=20)
    v = runs[0].visited
    assert all(np.array_equal(r.visited, v) for r in runs)
    for beta in (0.3, 0.4406868, 0.5):
        u = np.array([r.canonical(beta)["u"][0] for r in runs])
        want = exact.kaufman(beta, 8)["u"]
        err = u.std(ddof=1) / np.sqrt(len(u))
        assert abs(u.mean() - want) < max(5 * err, 1e-4), (beta, u.mean(), want, err)
    b = MulticanonicalBatch(model, runs[0], seeds=range(64)).run(200)
    assert len(b) == 64 and all(len(r.level_series) == 200 for r in b)
