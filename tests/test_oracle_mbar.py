"""Oracle MBAR (critmc/reweight.py:159-345) against closed forms and exact enumeration.

* One run: MBAR is single-histogram reweighting, ln Z(b)/Z(b0) = ln <e^{-(b-b0)E}>.
* Two runs at the same beta: f_1 = f_0 exactly (by symmetry of the objective).
* Exact samples: energies drawn i.i.d. from the enumerated 4x4 Ising energy
  distribution at K betas; the MBAR free energies match the exact ln Z ratios
  to within a few standard errors, and a planted wrong beta is rejected.
"""

import numpy as np
import pytest
from scipy.special import logsumexp

from oracle import exact
from oracle.mbar import Mbar


def _ising_levels():
    from critmc_amd import square

    lat = square(4)
    e_b, m = exact.ising_states(lat.bonds, 16)
    levels, deg = np.unique(e_b, return_counts=True)
    return levels.astype(float), deg.astype(float)


def _exact_log_z(levels, deg, beta):
    return logsumexp(np.log(deg) - beta * levels)


def _draw(levels, deg, beta, n, rng):
    lw = np.log(deg) - beta * levels
    p = np.exp(lw - logsumexp(lw))
    return rng.choice(levels, size=n, p=p)


def test_one_run_is_single_histogram_reweighting():
    rng = np.random.default_rng(1)
    e = rng.normal(-30.0, 5.0, size=5000).round()
    mb = Mbar([e], [0.4])
    mb.solve()
    for b in (0.35, 0.4, 0.45):
        want = logsumexp(-(b - 0.4) * e) - np.log(len(e))
        assert mb.log_z(b)[0] == pytest.approx(want, abs=1e-10)


def test_two_runs_at_one_beta_have_equal_free_energy():
    rng = np.random.default_rng(2)
    levels, deg = _ising_levels()
    e0 = _draw(levels, deg, 0.3, 4000, rng)
    e1 = _draw(levels, deg, 0.3, 6000, rng)
    f = Mbar([e0, e1], [0.3, 0.3]).solve()
    assert abs(f[1] - f[0]) < 1e-9


@pytest.mark.parametrize("g", [None, [1.0, 2.0, 1.5, 3.0]])
def test_exact_samples_recover_log_z(g):
    rng = np.random.default_rng(3)
    levels, deg = _ising_levels()
    betas = [0.2, 0.3, 0.4, 0.5]
    runs = [_draw(levels, deg, b, 20_000, rng) for b in betas]
    mb = Mbar(runs, betas, inefficiency=g)
    f = mb.solve()
    _, grad, hess = mb.objective(mb.fs)
    assert np.max(np.abs(grad[1:])) < 1e-10 * np.mean(mb.counts / mb.g)
    assert np.allclose(hess, hess.T) and np.all(np.linalg.eigvalsh(hess[1:, 1:]) > 0)
    exact_f = np.array([-(_exact_log_z(levels, deg, b) - _exact_log_z(levels, deg, betas[0]))
                        for b in betas])
    assert np.max(np.abs(f - exact_f)) < 0.02, (f, exact_f)
    # log Z between and beyond the runs
    for b in (0.25, 0.45):
        want = _exact_log_z(levels, deg, b) - _exact_log_z(levels, deg, betas[0])
        assert mb.log_z(b)[0] == pytest.approx(want, abs=0.02)


def test_planted_wrong_beta_is_rejected():
    rng = np.random.default_rng(4)
    levels, deg = _ising_levels()
    betas = [0.2, 0.35, 0.5]
    runs = [_draw(levels, deg, b, 20_000, rng) for b in betas]
    wrong = Mbar(runs, [0.2, 0.35 * 1.1, 0.5]).solve()
    exact_f = np.array([-(_exact_log_z(levels, deg, b) - _exact_log_z(levels, deg, 0.2))
                        for b in betas])
    assert np.max(np.abs(wrong - exact_f)) > 0.1
