"""Single-histogram reweighting (critmc/reweight.py:40-156) on the host.

Exact samples of the 4x4 Ising energy drawn at beta0 and reweighted to a
nearby beta reproduce the exact <E> there; at beta0 the weights are uniform;
the Kish ESS of uniform weights is n; the estimators agree with direct
formulas for the moments they are built from.
"""

import numpy as np
import pytest
from scipy.special import logsumexp

from critmc_amd.reweight import Reweighter, canonical_observables, kish_ess
from oracle import exact


def _levels():
    from critmc_amd import square

    lat = square(4)
    e_b, m = exact.ising_states(lat.bonds, 16)
    return e_b.astype(float), m.astype(float)


def test_reweighting_exact_samples():
    e_all, m_all = _levels()
    beta0 = 0.35
    lw = -beta0 * e_all
    p = np.exp(lw - logsumexp(lw))
    rng = np.random.default_rng(0)
    idx = rng.choice(len(e_all), size=400_000, p=p)
    rw = Reweighter(e_all[idx], beta0, {"m2": (m_all[idx] / 16) ** 2})
    assert np.allclose(rw.weights(beta0), 1.0 / len(idx))
    assert rw.ess(beta0) == pytest.approx(len(idx))
    for b in (0.33, 0.37):
        w = np.exp(-b * e_all - logsumexp(-b * e_all))
        assert rw.mean(e_all[idx], b) == pytest.approx(w @ e_all, abs=0.05)
        assert rw.mean("m2", b) == pytest.approx(w @ (m_all / 16) ** 2, abs=0.005)
    lo, hi = rw.reliable_range(0.1)
    assert lo < beta0 < hi


def test_canonical_observables_match_moments():
    rng = np.random.default_rng(1)
    e = rng.normal(-300, 10, 5000)
    m = rng.uniform(0, 1, 5000)
    out = canonical_observables(e, m, m ** 2, m ** 4, 0.4, 0.4, 256)
    assert out["e"] == pytest.approx(e.mean() / 256)
    assert out["c"] == pytest.approx(0.16 * e.var() / 256)
    assert out["U"] == pytest.approx(1 - np.mean(m ** 4) / (3 * np.mean(m ** 2) ** 2))
    assert out["chi"] == pytest.approx(0.4 * 256 * (np.mean(m ** 2) - m.mean() ** 2))
    assert kish_ess(np.zeros(100)) == pytest.approx(100.0)


def test_beta_derivatives_match_finite_differences():
    """dU, dlogm, dlogm2 and Reweighter.derivative are the beta-derivatives of the
    reweighted U, ln<|m|>, ln<m^2> and <O>: compare with central differences."""
    rng = np.random.default_rng(2)
    n = 20_000
    e = rng.normal(-300.0, 12.0, n)
    m = np.clip(0.5 + 0.004 * (-300.0 - e) + rng.normal(0, 0.05, n), 0, 1)  # correlated with E
    beta0, beta, h = 0.42, 0.425, 1e-5

    def at(b):
        return canonical_observables(e, m, m ** 2, m ** 4, beta0, b, 256)

    lo, mid, hi = at(beta - h), at(beta), at(beta + h)
    assert mid["dU"] == pytest.approx((hi["U"] - lo["U"]) / (2 * h), rel=1e-5)
    assert mid["dlogm"] == pytest.approx((np.log(hi["absm"]) - np.log(lo["absm"])) / (2 * h), rel=1e-5)
    assert mid["dlogm2"] == pytest.approx((np.log(hi["m2"]) - np.log(lo["m2"])) / (2 * h), rel=1e-5)
    rw = Reweighter(e, beta0, {"m": m})
    fd = (rw.mean("m", beta + h) - rw.mean("m", beta - h)) / (2 * h)
    assert rw.derivative("m", beta) == pytest.approx(fd, rel=1e-5)
    # an estimator subset gives the same values as the full set
    from critmc_amd.reweight import observables_from_weights

    w = rw.weights(beta)
    sub = observables_from_weights(w, e, m, m ** 2, m ** 4, beta, 256, keys=["dU", "chi"])
    assert sub.keys() == {"dU", "chi"}
    assert sub["dU"] == mid["dU"] and sub["chi"] == mid["chi"]
    assert observables_from_weights(w, e, m, m ** 2, m ** 4, beta, 256, keys=[]) == {}
