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
from critmc_amd._lib import set_option
from critmc_amd.glass import EdwardsAnderson, ParallelTempering, TemperingBatch, geometric_temperatures

pytestmark = pytest.mark.gpu


@pytest.fixture
def ea_launch_form():
    """Yields a setter for the multi-spin launch form (1: one fused launch per
    call, 0: three launches per step); the default is restored afterwards."""
    old = set_option("ea_fused", 1)
    yield lambda form: set_option("ea_fused", form)
    set_option("ea_fused", old)


def _check_against_replay(ndim, Ls, m, code, disorder, n_bins=5, spot=None, sweeps_per_step=1):
    """Run a batch on the device and compare the samples in ``spot`` (default: all)
    with the CPU replay from the same start: two calls, the second continuing."""
    models = [EdwardsAnderson(hypercubic((L,) * ndim), disorder=disorder, seed=7 + s)
              for s, L in enumerate(Ls)]
    fp = disorder == "gaussian"
    temps = geometric_temperatures(0.8, 2.0, m)
    keys = [(1000 + s, 77) for s in range(len(Ls))]
    batch = TemperingBatch(models, temps, keys=keys, update="heatbath" if code == 2 else "metropolis",
                           sweeps_per_step=sweeps_per_step)
    runs = batch.run(n_bins, record=True)
    runs2 = batch.run(3, record=True)  # a second call continues the streams
    eng = batch.engine
    assert eng.multispin == (not fp and m <= 32 and max(Ls) <= (32 if ndim == 2 else 16))
    for s in (range(len(Ls)) if spot is None else spot):
        L = Ls[s]
        N = L ** ndim
        rep = oracle.EaReplay(oracle.ea_init(m, N, keys[s]), models[s].device_couplings(), ndim, L,
                              1.0 / np.sort(temps)[::-1], keys[s], code=code,
                              sweeps_per_step=sweeps_per_step, multispin=eng.multispin)
        ref = rep.run((1 << n_bins) - 1, n_bins)
        ref2 = rep.run(7, 3)
        assert np.array_equal(eng.configurations(s), rep.spins), f"sample {s}: spins"
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
        swaps = eng.swap_acc.cpu().numpy().reshape(-1, 2, m - 1)[s]
        tries = eng.swap_try.cpu().numpy().reshape(-1, 2, m - 1)[s]
        assert np.array_equal(swaps, ref2["swap_acc"]) and np.array_equal(tries, ref2["swap_try"])
        series = batch.last_series
        if fp:
            np.testing.assert_allclose(series["E"][s], ref2["E"], rtol=0, atol=1e-9)
        else:
            assert np.array_equal(series["E"][s], ref2["E"])
        assert np.array_equal(series["q"][s], ref2["q"])


@pytest.mark.parametrize("ndim,Ls,code,disorder", [
    (3, [4, 6, 8, 4], 2, "bimodal"), (2, [4, 8, 6], 0, "bimodal"), (3, [8, 12, 16], 2, "bimodal"),
    (3, [4, 6, 8, 4], 2, "gaussian"), (2, [4, 8, 6], 0, "gaussian"), (3, [8, 12], 0, "gaussian"),
    (3, [4, 18], 2, "bimodal")])  # L = 18 > 16: the per-configuration +-J path
def test_batch_is_bit_exact_against_replay(ndim, Ls, code, disorder):
    _check_against_replay(ndim, Ls, 5, code, disorder)


@pytest.mark.parametrize("ndim,Ls,m,code,disorder", [
    # the spin-glass workload's ladder: 32 temperatures in [0.8, 2.0], heat bath,
    # L in {8, 12, 16}: every bit of the multi-spin word live, the full 32-plane
    # threshold table and all 528 parallel swap decisions
    (3, [8, 12, 16, 16, 8, 12], 32, 2, "bimodal"),
    (3, [16, 8], 32, 0, "bimodal"),
    (3, [8, 12], 32, 2, "gaussian"),
    (3, [8, 12, 16], 33, 2, "bimodal"),  # 33 temperatures: the per-configuration path
])
def test_full_ladder_is_bit_exact_against_replay(ndim, Ls, m, code, disorder):
    _check_against_replay(ndim, Ls, m, code, disorder, n_bins=4)


@pytest.mark.parametrize("form", [1, 0], ids=["fused", "three-launch"])
@pytest.mark.parametrize("ndim,Ls,m,code,sps", [
    (3, [16, 8, 12, 8, 16], 32, 2, 1),
    (3, [6, 4, 10], 7, 0, 2),            # two sweeps per PT step
    (2, [20, 32, 8, 18], 32, 2, 1),      # rows of 17-32 sites: 32-bit overlap patterns
    (2, [12, 16, 4], 9, 0, 3),
])
def test_multispin_launch_forms_are_bit_exact(ea_launch_form, form, ndim, Ls, m, code, sps):
    """The fused launch (every PT step of a call in one workgroup per sample) and
    the three-launch step give the replay's chains and series bit for bit."""
    ea_launch_form(form)
    _check_against_replay(ndim, Ls, m, code, "bimodal", n_bins=4, sweeps_per_step=sps)


def _recount(spins, J, L):
    """E = -sum over forward bonds of J s_i s_j for every configuration ([C, N] spins;
    J axis-major, x fastest)."""
    s = spins.reshape(-1, L, L, L).astype(np.int32)
    J = np.asarray(J).reshape(3, L, L, L)
    assert np.all(np.abs(J) == 1)
    J = J.astype(np.int64)
    e = np.zeros(s.shape[0], np.int64)
    for a, ax in enumerate((3, 2, 1)):  # bond axis a = x, y, z
        e -= (J[a][None] * s * np.roll(s, -1, axis=ax)).reshape(s.shape[0], -1).sum(axis=1)
    return e


def test_spin_glass_workload_size_properties():
    """512 +-J samples, L drawn from {8, 12, 16} with seed 7, 2 sets x 32 temperatures,
    heat bath + swaps: after 7 PT steps every sample's energies equal a recount from its
    spins, the last E / q series entries equal the final configurations' energies and
    overlaps, each set's permutation is a permutation with kofc its inverse, and spot
    samples match the CPU replay bit for bit."""
    from critmc_amd.rng import make_rng

    S, m = 512, 32
    Ls = make_rng(7).choice([8, 12, 16], size=S)
    models = [EdwardsAnderson(hypercubic((int(L),) * 3), disorder="bimodal", seed=make_rng(7, s))
              for s, L in enumerate(Ls)]
    temps = geometric_temperatures(0.8, 2.0, m)
    keys = [(7_000 + s, 99) for s in range(S)]
    batch = TemperingBatch(models, temps, keys=keys, update="heatbath")
    eng = batch.engine
    assert eng.multispin
    n_bins = 3
    batch.run(n_bins, record=True)
    n_steps = (1 << n_bins) - 1
    spins = eng.spins.cpu().numpy()
    energies = eng.energies.cpu().numpy().reshape(S, 2, m)
    perm = eng.perm.cpu().numpy().reshape(S, 2, m)
    kofc = eng.kofc.cpu().numpy().reshape(S, 2, m)
    E, q = batch.last_series["E"], batch.last_series["q"]
    assert E.shape == (S, 2, m, n_steps) and q.shape == (S, m, n_steps)
    for s in range(S):
        L, N = int(Ls[s]), int(Ls[s]) ** 3
        o = int(eng.spin_off[s])
        conf = spins[o:o + 2 * m * N].reshape(2, m, N)
        J = models[s].device_couplings()
        for r in range(2):
            assert np.array_equal(np.sort(perm[s, r]), np.arange(m)), f"sample {s}: perm"
            assert np.array_equal(kofc[s, r][perm[s, r]], np.arange(m)), f"sample {s}: kofc"
            assert np.array_equal(_recount(conf[r], J, L), energies[s, r]), f"sample {s}: E"
            assert np.array_equal(E[s, r, :, -1], energies[s, r][perm[s, r]])
        qa = conf[0][perm[s, 0]].astype(np.int32)
        qb = conf[1][perm[s, 1]].astype(np.int32)
        assert np.array_equal((qa * qb).sum(axis=1), q[s, :, -1]), f"sample {s}: q"
    for s in (0, 1, 255, 510, 511):
        L, N = int(Ls[s]), int(Ls[s]) ** 3
        rep = oracle.EaReplay(oracle.ea_init(m, N, keys[s]), models[s].device_couplings(), 3, L,
                              1.0 / np.sort(temps)[::-1], keys[s], code=2, multispin=True)
        ref = rep.run(n_steps, n_bins)
        assert np.array_equal(eng.configurations(s), rep.spins), f"sample {s}: spins vs replay"
        assert np.array_equal(E[s], ref["E"]) and np.array_equal(q[s], ref["q"])
        assert np.array_equal(perm[s].reshape(-1), rep.perm)


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
