"""GPU parity of the checkerboard Potts kernels against the CPU replay (bit-exact)."""

import numpy as np
import pytest
import torch

import oracle
from critmc_amd.kernels import potts as P
from critmc_amd.kernels.chains import Chains

pytestmark = pytest.mark.gpu


def _chains(R, seed, dev):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 1 << 32, size=(R, 2), dtype=np.uint64)
    sids = rng.integers(0, 1 << 20, size=R, dtype=np.uint64)
    sweeps = rng.integers(0, 1 << 36, size=R, dtype=np.uint64)
    return Chains(keys, sids, sweeps, dev), keys, sids, sweeps


@pytest.mark.parametrize("shape,q", [((4, 4), 3), ((32, 32), 10), ((16, 6), 2), ((10,), 4),
                                     ((4, 4, 6), 5), ((512, 512), 10), ((6, 4), 127),
                                     # rows of 8k sites with q <= 10 take the 8-byte-span quads:
                                     # two rows (+y = -y), one quad per row, q = 7 of 8 fields
                                     ((8, 2), 8), ((8, 8), 10), ((24, 10), 7)])
@pytest.mark.parametrize("sched", [[[0, 1]], [[2, 1]], [[0, 1], [2, 2]]])
@pytest.mark.parametrize("grid", [0, 2, -1], ids=["auto", "per-replica", "wide"])
def test_potts_run_is_bit_exact_against_replay(shape, q, sched, grid):
    dev = torch.device("cuda", 0)
    R = 1 if shape == (512, 512) else 3
    chains, keys, sids, sw0 = _chains(R, q + sum(shape), dev)
    n = int(np.prod(shape))
    s = torch.empty((R, n), dtype=torch.int8, device=dev)
    P.init_states(s, shape, q, chains, "random")
    s0 = s.cpu().numpy()
    for r in range(R):
        assert np.array_equal(s0[r], oracle.potts_init(n, q, keys[r], int(sids[r])))
    e, c = P.measure(s, shape, q, R)
    betas = np.random.default_rng(q).uniform(0.5, 1.6, size=R)
    n_meas = 2 if shape == (512, 512) else 6
    e_out, acc, counts = P.run(s, shape, q, betas, chains, np.array(sched), n_meas, e, c, grid=grid)
    got = s.cpu().numpy()
    e_out, acc, counts = (t.cpu().numpy() for t in (e_out, acc, counts))
    for r in range(R):
        rep = oracle.PottsReplay(s0[r], shape, q, betas[r], keys[r], int(sids[r]), int(sw0[r]))
        ref = rep.run(sched, n_meas)
        assert np.array_equal(got[r], rep.s), f"states differ (replica {r})"
        assert np.array_equal(e_out[r], ref["E"])
        assert np.array_equal(acc[r], ref["accepted"])
        assert np.array_equal(counts[r], ref["counts"])
    e2, c2 = P.measure(s, shape, q, R)
    assert torch.equal(e2, e) and torch.equal(c2, c)


def test_potts_simulation_api_matches_replay():
    from critmc_amd import Potts, Simulation, make_rng, square
    from critmc_amd.rng import chain_key

    sim = Simulation(Potts(square(8), 3), 0.9, "metropolis+heatbath", seed=21)
    ts = sim.run(40)
    rng = make_rng(21)
    s0 = Potts(square(8), 3).initial_state("random", rng)
    rep = oracle.PottsReplay(s0, (8, 8), 3, 0.9, chain_key(rng))
    ref = rep.run([[0, 1], [2, 1]], 40)
    assert np.array_equal(sim.state, rep.s)
    assert np.array_equal(ts["E"], ref["E"]) and np.array_equal(ts["counts"], ref["counts"])
    assert sim.running["energy"] == int(ref["E"][-1])
    assert np.array_equal(sim.running["counts"], np.bincount(sim.state, minlength=3))


@pytest.mark.parametrize("grid", [2, -1], ids=["per-replica", "wide"])
def test_caller_workspace_tables_follow_the_betas(grid):
    """A caller-owned workspace keeps its tables while the betas stay the same
    and is prepared again when they change; each call continues the chains
    exactly as the replay does with that call's betas."""
    dev = torch.device("cuda", 0)
    shape, q, R = (16, 16), 10, 2
    chains, keys, sids, sw0 = _chains(R, 77, dev)
    n = int(np.prod(shape))
    s = torch.empty((R, n), dtype=torch.int8, device=dev)
    P.init_states(s, shape, q, chains, "random")
    state = s.cpu().numpy()
    e, c = P.measure(s, shape, q, R)
    ws = P.Workspace(R, dev)
    sched = np.array([[2, 1], [0, 1]])
    rng = np.random.default_rng(5)
    beta_a, beta_b = rng.uniform(0.8, 1.6, size=R), rng.uniform(0.8, 1.6, size=R)
    sweep = sw0.astype(np.int64).copy()
    # calls alternate between two streams, each after the previous call's work
    # (the chains' state is shared); a call on another stream than the last
    # upload's prepares the tables again on its own stream
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for call, betas in enumerate((beta_a, beta_b, beta_b, beta_a, beta_a)):
        n_meas = 3
        st = streams[call % 2]
        st.wait_stream(streams[1 - call % 2])
        st.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(st):
            e_out, acc, counts = P.run(s, shape, q, betas, chains, sched, n_meas, e, c, grid=grid,
                                       ws=ws)
        torch.cuda.current_stream().wait_stream(st)
        got = s.cpu().numpy()
        e_out, acc, counts = (t.cpu().numpy() for t in (e_out, acc, counts))
        for r in range(R):
            rep = oracle.PottsReplay(state[r], shape, q, betas[r], keys[r], int(sids[r]), int(sweep[r]))
            ref = rep.run(sched.tolist(), n_meas)
            assert np.array_equal(got[r], rep.s), f"states differ (replica {r})"
            assert np.array_equal(e_out[r], ref["E"])
            assert np.array_equal(acc[r], ref["accepted"])
            assert np.array_equal(counts[r], ref["counts"])
        state = got
        sweep += n_meas * int(sched[:, 1].sum())
