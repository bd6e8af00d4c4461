"""GPU chains on lattices the checkerboard kernels cannot sweep, and random-site
Metropolis: the neighbour-table kernels against the CPU replay, bit for bit
(spins, E, M or counts, aux columns, continued calls), then the reference's own
statistical tests (critmc tests/test_kernels.py:21-52) through the Simulation
API on the square 3x3 and triangular 3x3 tori."""

import numpy as np
import pytest
import torch

import oracle
from critmc_amd import Ising, Potts, Simulation, chain, cubic, square, triangular
from critmc_amd.kernels import table as T
from critmc_amd.kernels.chains import Chains
from oracle import exact
from oracle.stats import batch_mean_error

pytestmark = pytest.mark.gpu

LATTICES = {"square3": square(3), "square5x3": square(5, 3), "triangular3": triangular(3),
            "triangular5x4": triangular(5, 4), "cubic3": cubic(3), "chain5": chain(5),
            "square4": square(4)}
SCHEDULES = [[[0, 1]], [[1, 1]], [[2, 1]], [[3, 2]], [[4, 1]], [[1, 2], [3, 1], [2, 1], [4, 1]]]


def _batch(lat, q, R, seed):
    rng = np.random.default_rng(seed)
    n = lat.n_sites
    s0 = (2 * rng.integers(0, 2, (R, n)) - 1).astype(np.int8) if q == 0 else \
        rng.integers(0, q, (R, n)).astype(np.int8)
    keys = rng.integers(0, 1 << 32, size=(R, 2), dtype=np.uint64)
    betas = rng.uniform(0.2, 0.6, R) if q == 0 else rng.uniform(0.6, 1.6, R)
    return s0, keys, betas


def _compare(lat, q, sched, R=4, h=0.0, n_meas=(6, 5)):
    dev = torch.device("cuda", 0)
    s0, keys, betas = _batch(lat, q, R, 17 + len(sched) + q)
    sids = np.arange(R, dtype=np.uint32) + 3
    chains = Chains(keys, sids, None, dev)
    s = torch.from_numpy(s0).to(dev)
    nbr = T.neighbours(lat, dev)
    e, mc = T.measure(s, nbr, R, q)
    ws = T.Workspace(lat.n_sites, R, q, dev)
    outs = []
    for n in n_meas:  # a second call continues every chain
        if q == 0:
            eo, mo, aux = T.ising_run(s, nbr, betas, h, chains, sched, n, e, mc, ws=ws)
        else:
            eo, mo, aux = T.potts_run(s, nbr, q, betas, chains, sched, n, e, mc, ws=ws)
        outs.append((eo.cpu().numpy(), mo.cpu().numpy(), aux.cpu().numpy()))
    got = s.cpu().numpy()
    print(f"{lat!r} q={q} schedule={sched} R={R}", flush=True)
    for r in range(R):
        rep = oracle.TableReplay(s0[r], lat.nbr, q, betas[r], h, keys[r], int(sids[r]))
        for k, n in enumerate(n_meas):
            ref = rep.run(sched, n)
            eo, mo, aux = outs[k]
            assert np.array_equal(eo[r], ref["E"]), (r, k, "E")
            want = ref["M"] if q == 0 else ref["counts"]
            assert np.array_equal(mo[r], want), (r, k, "M/counts")
            assert np.array_equal(aux[r], ref["aux"]), (r, k, "aux")
        assert np.array_equal(got[r], rep.s), (r, "spins")
    # running totals equal a recount
    e2, mc2 = T.measure(s, nbr, R, q)
    assert torch.equal(e2, e) and torch.equal(mc2, mc)
    assert np.array_equal(chains.sweeps_d.cpu().numpy().astype(np.uint64),
                          np.full(R, sum(n_meas) * int(np.sum(np.array(sched)[:, 1])), np.uint64))


@pytest.mark.parametrize("lat", list(LATTICES.values()), ids=list(LATTICES))
@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: "+".join(str(r[0]) for r in s))
def test_ising_chains_are_bit_exact_against_replay(lat, sched):
    _compare(lat, 0, sched)


@pytest.mark.parametrize("lat", [LATTICES["triangular5x4"], LATTICES["square5x3"]],
                         ids=["triangular5x4", "square5x3"])
def test_ising_in_a_field_is_bit_exact(lat):
    _compare(lat, 0, [[0, 1], [1, 1], [2, 1]], h=0.37)


@pytest.mark.parametrize("q", [3, 10])
@pytest.mark.parametrize("lat", [LATTICES["triangular3"], LATTICES["square5x3"], LATTICES["cubic3"],
                                 LATTICES["square4"]], ids=["triangular3", "square5x3", "cubic3",
                                                            "square4"])
@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: "+".join(str(r[0]) for r in s))
def test_potts_chains_are_bit_exact_against_replay(q, lat, sched):
    _compare(lat, q, sched)


def test_many_chains_on_a_larger_lattice():
    """300 chains (five workgroups of lanes) on a 24 x 16 triangular torus."""
    _compare(triangular(24, 16), 0, [[1, 1], [2, 1], [3, 1]], R=300, n_meas=(3, 2))


def _pull(series, exact_value):
    mean, err = batch_mean_error(series)
    return (mean - exact_value) / err


@pytest.mark.parametrize("updates", ["metropolis", "metropolis-random", "heatbath", "wolff", "sw",
                                     "metropolis-random+wolff"])
def test_square_3x3_ising_matches_enumeration(updates):
    lat = square(3)
    n, beta = lat.n_sites, 0.35
    sim = Simulation(Ising(lat), beta, updates, seed=12)
    sim.thermalize(2000)
    ts = sim.run(200_000)
    assert ts.meta["sweep_order"] == "sequential"
    assert abs(_pull(ts["E"].astype(float), exact.ising_average(lat.bonds, n, beta, 0.0,
                                                                 lambda e, m: e))) < 4.5
    assert abs(_pull(ts["m2"], exact.ising_average(lat.bonds, n, beta, 0.0,
                                                   lambda e, m: (m / n) ** 2))) < 4.5
    s = sim.state.astype(np.int64)
    assert sim.running["e_b"] == -(s[lat.bonds[:, 0]] * s[lat.bonds[:, 1]]).sum()
    assert sim.running["m"] == s.sum()


@pytest.mark.parametrize("updates", ["metropolis", "heatbath", "wolff", "sw"])
def test_triangular_3x3_ising_matches_enumeration(updates):
    lat = triangular(3)
    n, beta = lat.n_sites, 0.25
    sim = Simulation(Ising(lat), beta, updates, seed=3)
    sim.thermalize(1000)
    ts = sim.run(150_000)
    assert abs(_pull(ts["E"].astype(float), exact.ising_average(lat.bonds, n, beta, 0.0,
                                                                 lambda e, m: e))) < 4.5
    assert abs(_pull(ts["m2"], exact.ising_average(lat.bonds, n, beta, 0.0,
                                                   lambda e, m: (m / n) ** 2))) < 4.5


@pytest.mark.parametrize("updates", ["metropolis", "metropolis-random", "heatbath", "sw"])
def test_triangular_potts_matches_enumeration(updates):
    lat, q, beta = triangular(3), 3, 0.8
    n = lat.n_sites
    sim = Simulation(Potts(lat, q), beta, updates, seed=8)
    sim.thermalize(1000)
    ts = sim.run(100_000)
    assert abs(_pull(ts["E"].astype(float), exact.potts_average(lat.bonds, n, q, beta,
                                                                 lambda e, c: e))) < 4.5
