"""The exact oracles, pinned by the reference's stored vectors and closed forms."""

import json
from pathlib import Path

import numpy as np
import pytest

from oracle import exact, stats

GOLDEN = Path(__file__).resolve().parent / "golden"


def _golden(name):
    return json.loads((GOLDEN / name).read_text())


def test_kaufman_matches_reference_vector_L16():
    g = _golden("kaufman_L16.json")  # critmc results/e01_validation.json (exact.u, exact.c)
    for T, u, c in zip(g["temperatures"], g["u"], g["c"]):
        k = exact.kaufman(1.0 / T, 16)
        assert k["u"] == pytest.approx(u, abs=1e-12)
        assert k["c"] == pytest.approx(c, abs=1e-11)


def test_kaufman_specific_heat_at_beta_c_matches_reference():
    g = _golden("kaufman_c_at_beta_c.json")  # critmc results/e03_ising2d_fss.json
    for L, c in g["c_exact_at_beta_c"].items():
        if int(L) > 128:
            continue
        assert exact.kaufman(g["beta_c"], int(L))["c"] == pytest.approx(c, rel=1e-11)


@pytest.mark.parametrize("shape", [(4, 4), (2, 6), (4, 6)])
@pytest.mark.parametrize("beta", [0.2, 0.4406868, 0.7])
def test_kaufman_matches_enumeration(shape, beta):
    from critmc_amd import hypercubic

    lat = hypercubic(shape)
    n = lat.n_sites
    e = exact.ising_average(lat.bonds, n, beta, 0.0, lambda e, m: e) / n
    assert exact.kaufman(beta, shape[0], shape[1])["u"] == pytest.approx(e, abs=1e-12)


def test_kaufman_onsager_limit():
    assert exact.kaufman(0.3, 400)["u"] == pytest.approx(exact.onsager_energy(0.3), abs=2e-9)


def test_enumeration_chain_closed_form():
    from critmc_amd import chain

    beta, N = 0.7, 10
    lat = chain(N)
    e = exact.ising_average(lat.bonds, N, beta, 0.0, lambda e, m: e)
    # Z = (2 cosh b)^N + (2 sinh b)^N
    t = np.tanh(beta)
    expect = -N * (t + t ** (N - 1)) / (1 + t ** N)
    assert e == pytest.approx(expect, rel=1e-12)


def test_potts_q2_is_ising_at_half_coupling():
    from critmc_amd import square

    lat = square(2)
    lat4 = square(4)
    q2 = exact.potts_average(lat4.bonds, 16, 2, 0.8, lambda e, c: e)
    # delta(s,s') = (1 + s s')/2  ->  E_potts = (E_ising - B) / 2 at beta_ising = beta / 2
    ising = exact.ising_average(lat4.bonds, 16, 0.4, 0.0, lambda e, m: e)
    assert q2 == pytest.approx((ising - lat4.n_bonds) / 2, rel=1e-12)
    assert lat.n_sites == 4


def test_on_ring_limits():
    beta = 1.1
    from scipy.special import iv

    assert exact.on_ring_energy(beta, 400, 2) == pytest.approx(-iv(1, beta) / iv(0, beta), rel=1e-9)
    assert exact.on_ring_energy(beta, 400, 3) == pytest.approx(-(1 / np.tanh(beta) - 1 / beta), rel=1e-9)
    # two-site ring = double bond: E/N = -<cos> at coupling 2 beta
    assert exact.on_ring_energy(beta, 2, 2) == pytest.approx(-iv(1, 2 * beta) / iv(0, 2 * beta), rel=1e-9)


def test_gamma_method_ar1_closed_form():
    rng = np.random.default_rng(3)
    phi, n = 0.8, 400_000
    x = np.empty(n)
    x[0] = 0
    eps = rng.standard_normal(n)
    for i in range(1, n):
        x[i] = phi * x[i - 1] + eps[i]
    r = stats.gamma_mean(x)
    tau = (1 + phi) / (2 * (1 - phi))
    assert r.tau_int == pytest.approx(tau, rel=0.08)
    var = 1 / (1 - phi ** 2)
    assert r.error == pytest.approx(np.sqrt(2 * tau * var / n), rel=0.08)


def test_gamma_method_matches_loop_transcription():
    rng = np.random.default_rng(11)
    x = np.cumsum(rng.standard_normal(3000)) * 0.05 + rng.standard_normal(3000)
    r = stats.gamma_mean(x)
    # direct loops over the same definitions
    n = len(x)
    d = x - x.mean()
    gam = [sum(d[i] * d[i + t] for i in range(n - t)) / (n - t) for t in range(r.window + 1)]
    c_f = gam[0] + 2 * sum(gam[1:])
    gam = [g + c_f / n for g in gam]
    c_f = gam[0] + 2 * sum(gam[1:])
    assert r.error == pytest.approx(np.sqrt(c_f / n), rel=1e-10)
    assert r.tau_int == pytest.approx(c_f / (2 * gam[0]), rel=1e-10)


def test_batch_means_error_closed_forms():
    """Batch means reproduce sigma / sqrt(N) for independent samples and the AR(1)
    error sqrt(2 tau_int var / N), tau_int = (1 + phi) / (2 (1 - phi)); the mean is
    the plain mean of the whole series."""
    from scipy.signal import lfilter

    rng = np.random.default_rng(11)
    n = 640_000
    iid = rng.normal(2.0, 3.0, n)
    mean, err = stats.batch_mean_error(iid)
    assert mean == iid.mean()
    assert err == pytest.approx(3.0 / np.sqrt(n), rel=0.25)  # 64 batches: ~9 % scatter
    phi = 0.9
    errs = []
    for _ in range(8):
        x = lfilter([1.0], [1.0, -phi], rng.standard_normal(n))
        errs.append(stats.batch_mean_error(x)[1])
    tau = (1 + phi) / (2 * (1 - phi))
    var = 1 / (1 - phi ** 2)
    assert np.mean(errs) == pytest.approx(np.sqrt(2 * tau * var / n), rel=0.1)
    # the first len % 64 samples stay out of the batches
    y = np.concatenate([[1e6], np.zeros(64 * 10)])
    assert stats.batch_mean_error(y)[1] == 0.0
