"""GPU O(n) chains in the reference's own site orders (FP64, one lane per chain,
neighbour table) against the FP64 replay (oracle/replay_seq_on.c).

The device runs the replay's operations in the same order; only the device's
libm (log, cos, sin, atan2, acos, exp) may differ from the host's in the last
bits.  So acceptance counts, Wolff sizes and cluster statistics agree exactly
(a decision could only flip if its uniform fell within such a rounding of the
threshold).  Over-relaxation amplifies a perturbation of a spin's field by about
2/|h| per sweep, and the XY heat bath's arccos has an unbounded derivative near
f = +-1, so those last-bit differences grow with the sweeps: spins agree to 1e-7
and FP64 sums to 1e-7 per site (measured: at most 1.1e-8 over 200 chains after
twelve sweeps of 108 sites mixing both, at most 1e-13 on the small lattices).  Then the reference's ring statistics
through the Simulation API on a ring of 7, which has no checkerboard, and
metropolis-random on a checkerboard lattice (FP32 colour-split state converted
for the call)."""

import numpy as np
import pytest
import torch

import oracle
from critmc_amd import XY, Heisenberg, Simulation, chain, cubic, square, triangular
from critmc_amd.kernels import table as T
from critmc_amd.kernels.chains import Chains
from oracle import exact
from oracle.stats import batch_mean_error

pytestmark = pytest.mark.gpu

LATTICES = {"chain7": chain(7), "square5x3": square(5, 3), "triangular5x4": triangular(5, 4),
            "cubic3": cubic(3), "square4": square(4)}
SCHEDULES = [[[0, 1]], [[1, 1]], [[2, 1]], [[5, 1], [0, 1]], [[3, 2]], [[4, 1]],
             [[1, 1], [4, 1], [5, 1], [2, 1]]]
TOL = 1e-7


def _units(shape, seed):
    s = np.random.default_rng(seed).standard_normal(shape)
    return s / np.linalg.norm(s, axis=-1, keepdims=True)


def _compare(lat, n, sched, R=3, n_meas=(4, 3)):
    dev = torch.device("cuda", 0)
    N = lat.n_sites
    rng = np.random.default_rng(11 + n + len(sched))
    s0 = _units((R, N, n), 5 + n)
    keys = rng.integers(0, 1 << 32, size=(R, 2), dtype=np.uint64)
    sids = np.arange(R, dtype=np.uint32) + 7
    betas = rng.uniform(0.4, 1.4, R)
    steps = rng.uniform(0.5, 2.0, R)
    chains = Chains(keys, sids, None, dev)
    s = torch.from_numpy(s0.copy()).to(dev)
    nbr = T.neighbours(lat, dev)
    ws = T.OnWorkspace(N, R, dev)
    outs = []
    for k in n_meas:  # a second call continues every chain
        outs.append([t.cpu().numpy() for t in T.on_run(s, nbr, n, betas, steps, chains, sched, k, ws=ws)])
    got = s.cpu().numpy()
    print(f"{lat!r} n={n} schedule={sched}", flush=True)
    worst = 0.0
    for r in range(R):
        rep = oracle.SeqOnReplay(s0[r], lat.nbr, n, betas[r], steps[r], keys[r], int(sids[r]))
        for k, nm in enumerate(n_meas):
            ref = rep.run(sched, nm)
            e, m, c, ss, aux = (o[r] for o in outs[k])
            assert np.array_equal(aux, ref["aux"]), (r, k, aux, ref["aux"])
            for name, a, b in (("E", e, ref["E"]), ("M", m, ref["M"]), ("c", c, ref["stiff_c"]),
                               ("s", ss, ref["stiff_s"])):
                np.testing.assert_allclose(a, b, rtol=0, atol=TOL * N, err_msg=f"{name} r={r} call={k}")
        worst = max(worst, float(np.abs(got[r] - rep.s).max()))
        np.testing.assert_allclose(got[r], rep.s, rtol=0, atol=TOL)
    print(f"  largest spin difference {worst:.1e}")
    assert np.array_equal(chains.sweeps_d.cpu().numpy().astype(np.uint64),
                          np.full(R, sum(n_meas) * int(np.sum(np.array(sched)[:, 1])), np.uint64))


@pytest.mark.parametrize("n", [2, 3])
@pytest.mark.parametrize("lat", list(LATTICES.values()), ids=list(LATTICES))
@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: "+".join(str(r[0]) for r in s))
def test_on_chains_match_fp64_replay(lat, n, sched):
    _compare(lat, n, sched)


def test_many_chains_on_a_larger_triangular_lattice():
    """200 chains (four workgroups of lanes) on a 12 x 9 triangular torus."""
    _compare(triangular(12, 9), 2, [[1, 1], [5, 1], [2, 1], [4, 1]], R=200, n_meas=(2, 1))


@pytest.mark.parametrize("n,updates", [(2, "heatbath"), (2, "metropolis-random"), (2, "overrelax+metropolis"),
                                       (2, "wolff"), (3, "sw"), (3, "heatbath")])
def test_ring_of_7_matches_transfer_operator(n, updates):
    lat, beta = chain(7), 0.8
    model = XY(lat) if n == 2 else Heisenberg(lat)
    sim = Simulation(model, beta, updates, seed=21)
    sim.thermalize(2000)
    ts = sim.run(60_000)
    assert ts.meta["sweep_order"] == "sequential"
    mean, err = batch_mean_error(ts["e"])
    assert abs(mean - exact.on_ring_energy(beta, 7, n)) < 4.5 * err, (mean, err)
    s = sim.state
    np.testing.assert_allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-12)


def test_metropolis_random_on_a_checkerboard_lattice():
    """XY on 6 x 4: metropolis-random runs sequentially on an FP64 site-order copy of
    the FP32 colour-split state; one call against the replay from the same start."""
    lat = square(6, 4)
    sim = Simulation(XY(lat), 0.9, "metropolis-random", seed=4)
    s0 = sim.state
    key = sim.chains.keys_h[0]
    ts = sim.run(5)
    assert ts.meta["sweep_order"] == "sequential"
    rep = oracle.SeqOnReplay(s0, lat.nbr, 2, 0.9, 1.0, key, int(sim.chains.sids_h[0]))
    ref = rep.run([[1, 1]], 5)
    assert np.array_equal(ts["aux"][:, 0], ref["aux"][:, 0])
    np.testing.assert_allclose(sim.state, rep.s, rtol=0, atol=2e-7)  # the FP32 store
    np.testing.assert_allclose(ts["E"], ref["E"], rtol=0, atol=TOL * lat.n_sites)
