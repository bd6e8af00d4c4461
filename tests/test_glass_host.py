"""Host-side spin-glass analysis: the temperature-ladder feedback step and the
disorder averages (critmc/glass.py:211-313), on synthetic inputs whose answers
are known in closed form.  No GPU needed."""

import numpy as np
import pytest

from critmc_amd.glass import DisorderAverage, TemperingRun, feedback_temperatures
from critmc_amd.kernels import ea as k_ea


def _hist_from_fraction(frac, visits=1_000_000):
    """Label histogram whose hot-end fraction at each temperature is ``frac``."""
    hot = np.round(visits * np.asarray(frac)).astype(np.int64)
    return np.column_stack([hot, visits - hot])


def test_feedback_keeps_the_ends_and_order():
    temps = np.array([0.6, 0.9, 1.2, 1.5, 1.8, 2.1])
    hist = np.array([[5, 95], [20, 80], [45, 55], [70, 30], [85, 15], [99, 1]])
    new = feedback_temperatures(temps, hist)
    assert new[0] == temps[0] and new[-1] == temps[-1]
    assert np.all(np.diff(new) > 0)


def test_feedback_fixed_point_is_an_even_fall_of_the_fraction():
    # equal rises of f across every interval are optimal whatever the spacing
    for temps in (np.geomspace(0.4, 2.5, 11), np.array([0.5, 0.6, 1.0, 1.1, 2.0])):
        frac = np.linspace(0.0, 1.0, len(temps))
        np.testing.assert_allclose(feedback_temperatures(temps, _hist_from_fraction(frac)), temps,
                                   rtol=1e-12)


def test_feedback_gathers_temperatures_at_a_bottleneck():
    temps = np.linspace(1.0, 4.0, 7)  # spacing 0.5; f jumps across [2.0, 2.5]
    frac = np.array([0.0, 0.02, 0.04, 0.92, 0.96, 0.98, 1.0])
    new = feedback_temperatures(temps, _hist_from_fraction(frac))
    between = lambda t: int(((t > 2.0) & (t < 2.5)).sum())  # noqa: E731
    assert between(temps) == 0 and between(new) >= 2
    assert np.all(np.diff(new) > 0)


def test_feedback_ignores_input_order_and_unvisited_temperatures():
    temps = np.geomspace(0.5, 2.0, 9)
    hist = _hist_from_fraction(np.linspace(0.0, 1.0, 9))
    perm = np.random.default_rng(3).permutation(9)
    np.testing.assert_allclose(feedback_temperatures(temps[perm], hist[perm]),
                               feedback_temperatures(temps, hist), rtol=0, atol=0)
    # a temperature never visited counts as f = 1/2; the result stays a valid ladder
    hist[4] = 0
    new = feedback_temperatures(temps, hist)
    assert np.all(np.isfinite(new)) and np.all(np.diff(new) > 0)


def _run(means: dict[str, np.ndarray], temps, n_sites=64, z=6) -> TemperingRun:
    """A run whose per-bin thermal means are ``means[name]`` ([bins, temps])."""
    bins = next(iter(means.values())).shape[0]
    acc = np.zeros((bins, len(temps), k_ea.N_OBS))
    count = 10.0
    cols = {"e": k_ea.OBS_E, "ql": k_ea.OBS_QL, "absq": k_ea.OBS_Q, "q2": k_ea.OBS_Q2,
            "q4": k_ea.OBS_Q4, "qk2": k_ea.OBS_QK2}
    for name, col in cols.items():
        acc[:, :, col] = count * means[name]
    acc[:, :, k_ea.OBS_COUNT] = count
    return TemperingRun(np.asarray(temps, float), n_sites, z, acc, np.ones(len(temps) - 1),
                        np.zeros((len(temps), 2)), np.zeros(2))


def _synthetic_runs(n_samples=40, bins=3, temps=(2.0, 1.5, 1.0), seed=0):
    rng = np.random.default_rng(seed)
    runs = []
    for _ in range(n_samples):
        q2 = rng.uniform(0.1, 0.5, size=(bins, len(temps)))
        means = {"e": rng.normal(-1.5, 0.1, size=q2.shape), "ql": rng.uniform(0.3, 0.7, size=q2.shape),
                 "absq": np.sqrt(q2), "q2": q2, "q4": q2 * q2 * rng.uniform(1.0, 2.5, size=q2.shape),
                 "qk2": q2 * rng.uniform(0.2, 0.6, size=q2.shape)}
        runs.append(_run(means, temps))
    return runs


def test_disorder_average_estimates_follow_their_definitions():
    runs = _synthetic_runs()
    avg = DisorderAverage(runs, L=4)
    last = {name: np.array([r.bin_means()[name][-1] for r in runs]) for name in ("e", "ql", "q2", "q4", "qk2")}
    q2, q4, qk2 = (last[k].mean(axis=0) for k in ("q2", "q4", "qk2"))
    obs = avg.observables(n_boot=20)
    np.testing.assert_allclose(obs["e"][0], last["e"].mean(axis=0), rtol=1e-14)
    np.testing.assert_allclose(obs["ql"][0], last["ql"].mean(axis=0), rtol=1e-14)
    np.testing.assert_allclose(obs["g"][0], (3 - q4 / q2 ** 2) / 2, rtol=1e-12)
    np.testing.assert_allclose(obs["chi_sg"][0], 64 * q2, rtol=1e-14)
    # xi_L / L with k_min = 2 pi / L: sqrt(chi(0) / chi(k_min) - 1) / (2 sin(pi / L)) / L
    np.testing.assert_allclose(obs["xi_L"][0], np.sqrt(q2 / qk2 - 1) / (2 * np.sin(np.pi / 4)) / 4,
                               rtol=1e-12)


def test_disorder_bootstrap_error_is_the_standard_error_of_the_mean():
    runs = _synthetic_runs(n_samples=60, seed=1)
    avg = DisorderAverage(runs, L=4)
    e_last = np.array([r.bin_means()["e"][-1] for r in runs])
    sem = e_last.std(axis=0, ddof=1) / np.sqrt(len(runs))
    _, err = avg.observables(n_boot=4000, seed=7)["e"]
    np.testing.assert_allclose(err, sem, rtol=0.1)
    # the same seed gives the same errors
    np.testing.assert_array_equal(avg.observables(n_boot=50, seed=2)["g"][1],
                                  avg.observables(n_boot=50, seed=2)["g"][1])


def test_identity_residuals_and_equilibration():
    temps = (2.0, 1.5, 1.0)
    runs = _synthetic_runs(n_samples=12, bins=4, temps=temps)
    avg = DisorderAverage(runs, L=4)
    res = avg.identity_residuals()
    assert res.shape == (4, 12, 3)
    r = runs[5].bin_means()
    np.testing.assert_allclose(res[:, 5], r["e"] + 3.0 * (1 - r["ql"]) / np.array(temps), rtol=1e-13)
    eq = avg.equilibration()
    np.testing.assert_allclose(eq["e"] - eq["e_from_ql"], res.mean(axis=1), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(eq["residual_error"], res.std(axis=1, ddof=1) / np.sqrt(12), rtol=1e-13)
    with pytest.raises(ValueError, match="no runs"):
        DisorderAverage([], L=4)
