"""GPU parity of batched EA parallel tempering against the CPU replay.

+-J couplings, bit-exact: spins, energies, permutations, labels, trips, label
histograms, swap counts and the int32 E / q series of every sample in an
uneven-size batch; the FP64 accumulators agree to rounding (sum order
differs).  Gaussian couplings: the same quantities bit-exact except the FP64
energies, E series and accumulators, which agree to 1e-9 (the device sums
energy changes in its reduction order; a swap decision could only differ if
its test fell within that rounding of the threshold).  Then the reference's
own statistical test through the ParallelTempering API, for both disorders.
"""

import numpy as np
import pytest

import oracle
from critmc_amd import hypercubic, square
from critmc_amd.glass import EdwardsAnderson, ParallelTempering, TemperingBatch, geometric_temperatures

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ndim,Ls,code,disorder", [
    (3, [4, 6, 8, 4], 2, "bimodal"), (2, [4, 8, 6], 0, "bimodal"), (3, [8, 12, 16], 2, "bimodal"),
    (3, [4, 6, 8, 4], 2, "gaussian"), (2, [4, 8, 6], 0, "gaussian"), (3, [8, 12], 0, "gaussian"),
    (3, [4, 18], 2, "bimodal")])  # L = 18 > 16: the per-configuration +-J path
def test_batch_is_bit_exact_against_replay(ndim, Ls, code, disorder):
    m = 5
    models = [EdwardsAnderson(hypercubic((L,) * ndim), disorder=disorder, seed=7 + s)
              for s, L in enumerate(Ls)]
    fp = disorder == "gaussian"
    temps = geometric_temperatures(0.8, 2.0, m)
    keys = [(1000 + s, 77) for s in range(len(Ls))]
    batch = TemperingBatch(models, temps, keys=keys, update="heatbath" if code == 2 else "metropolis")
    n_bins = 5
    runs = batch.run(n_bins, record=True)
    runs2 = batch.run(3, record=True)  # a second call continues the streams
    eng = batch.engine
    assert eng.multispin == (not fp and max(Ls) <= (32 if ndim == 2 else 16))
    for s, L in enumerate(Ls):
        N = L ** ndim
        rep = oracle.EaReplay(oracle.ea_init(m, N, keys[s]), models[s].device_couplings(), ndim, L,
                              1.0 / np.sort(temps)[::-1], keys[s], code=code,
                              multispin=eng.multispin)
        ref = rep.run((1 << n_bins) - 1, n_bins)
        ref2 = rep.run(7, 3)
        assert np.array_equal(eng.configurations(s), rep.spins)
        e = eng.energies.cpu().numpy().reshape(-1, 2 * m)[s]
        if fp:
            np.testing.assert_allclose(e, rep.energies, rtol=0, atol=1e-9)
        else:
            assert np.array_equal(e, rep.energies)
        assert np.array_equal(eng.perm.cpu().numpy().reshape(-1, 2 * m)[s], rep.perm)
        assert np.array_equal(eng.labels.cpu().numpy().reshape(-1, 2 * m)[s], rep.labels)
        for got, want in ((runs, ref), (runs2, ref2)):
            assert np.array_equal(got[s].trips, want["trips"].astype(np.int64))
            assert np.array_equal(got[s].label_hist, want["label_hist"].sum(axis=0))
            np.testing.assert_allclose(got[s].acc, want["acc"], rtol=1e-12, atol=1e-9 if fp else 1e-12)
        series = batch.last_series
        if fp:
            np.testing.assert_allclose(series["E"][s], ref2["E"], rtol=0, atol=1e-9)
        else:
            assert np.array_equal(series["E"][s], ref2["E"])
        assert np.array_equal(series["q"][s], ref2["q"])


@pytest.mark.parametrize("disorder", ["bimodal", "gaussian"])
def test_parallel_tempering_matches_enumeration(disorder):
    lat = square(4)
    ea = EdwardsAnderson(lat, disorder=disorder, seed=11)
    temps = geometric_temperatures(0.4, 2.5, 8)
    run = ParallelTempering(ea, temps, seed=5).run(16)
    eq = run.equilibrium()
    n = lat.n_sites
    states = np.arange(1 << n)
    spins = (2 * ((states[:, None] >> np.arange(n)) & 1) - 1).astype(np.float64)
    energies = -(spins[:, lat.bonds[:, 0]] * spins[:, lat.bonds[:, 1]]) @ ea.couplings
    b = lat.bonds
    for k, T in enumerate(run.temperatures):
        w = np.exp(-(energies - energies.min()) / T)
        w /= w.sum()
        corr = (spins * w[:, None]).T @ spins
        assert eq["e"][k] == pytest.approx(w @ energies / n, abs=0.01)
        assert eq["q2"][k] == pytest.approx(np.sum(corr ** 2) / n ** 2, abs=0.01)
        assert eq["ql"][k] == pytest.approx(np.mean(corr[b[:, 0], b[:, 1]] ** 2), abs=0.01)
    assert np.all(run.swap_rate > 0.3)
    assert run.trips.sum() > 100
