"""Statistical parity with the reference: GPU chains against exact answers.

* Ising on bipartite lattices against brute-force enumeration (the
  reference's own 3x3 is not bipartite; critmc tests/test_kernels.py:27-70).
* The plain CPU-runnable configuration: 64x64 at beta_c, Metropolis, cold
  start, seed 42, 10^4 + 10^5 sweeps; Gamma-method mean of e against
  Kaufman's finite torus, |z| < 3.
* Kaufman pulls at L = 16 across the critical region (critmc
  results/e01_validation.json temperatures) with the batched kernel.
* Potts against enumeration (critmc tests/test_kernels.py:78-92 analogue).
"""

import json
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import exact, stats

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _z(series, exact_value):
    mean, err = stats.batch_mean_error(series)
    return (mean - exact_value) / err


@pytest.mark.parametrize("updates,h", [("metropolis", 0.0), ("heatbath", 0.0), ("metropolis", 0.4),
                                       ("heatbath", 0.4), ("metropolis+heatbath", 0.0)])
@pytest.mark.parametrize("shape", [(4, 4), (2, 6)])
def test_ising_matches_enumeration(updates, h, shape):
    from critmc_amd import Ising, Simulation, hypercubic

    lat = hypercubic(shape)
    n = lat.n_sites
    beta = 0.35 if h == 0 else 0.3
    sim = Simulation(Ising(lat, h=h), beta, updates, seed=12)
    sim.thermalize(2000)
    ts = sim.run(200_000)
    zs = [_z(ts["E"], exact.ising_average(lat.bonds, n, beta, h, lambda e, m: e - h * m)),
          _z(ts["m2"], exact.ising_average(lat.bonds, n, beta, h, lambda e, m: (m / n) ** 2))]
    assert all(abs(z) < 4.5 for z in zs), zs
    assert sim.running["e_b"] == -int(np.sum(sim.state[lat.bonds[:, 0]] * sim.state[lat.bonds[:, 1]].astype(int)))
    assert sim.running["m"] == int(sim.state.astype(int).sum())


def test_plain_configuration_64_at_beta_c_against_kaufman():
    from critmc_amd import Ising, Simulation, square

    beta = float(np.log(1 + np.sqrt(2)) / 2)
    sim = Simulation(Ising(square(64)), beta, "metropolis", seed=42, start="ordered")
    sim.thermalize(10_000)
    ts = sim.run(100_000)
    g = stats.gamma_mean(ts["e"])
    u = exact.kaufman(beta, 64)["u"]
    z = (g.value - u) / g.error
    print(f"e = {g.value:.6f} +- {g.error:.2e} (tau_int {g.tau_int:.1f}) vs Kaufman {u:.6f}: z = {z:.2f}")
    assert abs(z) < 3, z


def test_kaufman_pulls_L16_batched():
    from critmc_amd.kernels import ising as K
    from critmc_amd.kernels.chains import Chains

    gold = json.loads((GOLDEN / "kaufman_L16.json").read_text())
    temps = np.array(gold["temperatures"])
    u = np.array(gold["u"])
    reps_per_t = 16
    R = len(temps) * reps_per_t
    betas = np.repeat(1.0 / temps, reps_per_t)
    dev = torch.device("cuda", 0)
    keys = np.random.default_rng(99).integers(0, 1 << 32, size=(R, 2), dtype=np.uint64)
    ch = Chains(keys, np.arange(R), None, dev)
    s = torch.empty((R, 256), dtype=torch.int8, device=dev)
    K.init_spins(s, (16, 16), ch, "random")
    e, m = K.measure(s, (16, 16), R)
    K.run(s, (16, 16), betas, 0.0, ch, np.array([[0, 1]]), 5000, e, m)
    e_out, _, _ = K.run(s, (16, 16), betas, 0.0, ch, np.array([[0, 1]]), 40_000, e, m)
    E = e_out.cpu().numpy().reshape(len(temps), reps_per_t, -1) / 256.0
    zs = []
    for t in range(len(temps)):
        means = E[t].mean(axis=1)  # independent replicas: their spread is the error
        mean, err = means.mean(), means.std(ddof=1) / np.sqrt(reps_per_t)
        zs.append((mean - u[t]) / err)
    zs = np.array(zs)
    assert np.all(np.abs(zs) < 4.5), zs
    assert abs(zs.mean()) < 4.5 / np.sqrt(len(zs)) * 1.5, zs


@pytest.mark.parametrize("shape,q,beta,updates", [((2, 4), 3, 0.8, "metropolis"),
                                                  ((2, 4), 3, 0.8, "heatbath"),
                                                  ((2, 2), 10, 1.4, "metropolis+heatbath")])
def test_potts_matches_enumeration(shape, q, beta, updates):
    from critmc_amd import Potts, Simulation, hypercubic

    lat = hypercubic(shape)
    n = lat.n_sites
    sim = Simulation(Potts(lat, q), beta, updates, seed=21)
    sim.thermalize(2000)
    ts = sim.run(150_000)
    z1 = _z(ts["E"], exact.potts_average(lat.bonds, n, q, beta, lambda e, c: e))
    m2 = lambda e, c: (q * np.sum((c / n) ** 2, axis=1) - 1) / (q - 1)  # noqa: E731
    z2 = _z(ts["m2"], exact.potts_average(lat.bonds, n, q, beta, m2))
    assert abs(z1) < 4.5 and abs(z2) < 4.5, (z1, z2)
    assert np.array_equal(sim.running["counts"], np.bincount(sim.state, minlength=q))
