"""CPU replay of EA parallel tempering (critmc/kernels/glass.py:33-163).

Pinned by: det_exp against exp (and the library's copy, bit for bit), energy
bookkeeping against a recount, swap/permutation invariants, and the
reference's own test -- PT on a 4x4 sample against 2^16-state enumeration
(critmc tests/test_glass.py:37-52), for Metropolis and heat bath, with +-J and
Gaussian couplings.
"""

import math

import numpy as np
import pytest

from critmc_amd import square
from critmc_amd.glass import EdwardsAnderson, geometric_temperatures
from oracle import exact


def test_det_exp_is_exp_and_matches_library(oracle_lib):
    from critmc_amd._lib import lib

    rng = np.random.default_rng(0)
    xs = np.concatenate([-rng.exponential(3.0, 4000), -rng.uniform(0, 700, 200),
                         rng.uniform(0, 60, 500), [0.0, -1e-300, -744.0]])
    for x in xs:
        d = oracle_lib.det_exp(x)
        assert d == pytest.approx(math.exp(x), rel=4e-16, abs=1e-320)
        assert lib().cm_det_exp(float(x)) == d


def _sample(seed=11, L=4, ndim=2, disorder="bimodal"):
    from critmc_amd import hypercubic

    lat = hypercubic((L,) * ndim)
    ea = EdwardsAnderson(lat, disorder=disorder, seed=seed)
    return lat, ea


@pytest.mark.parametrize("disorder", ["bimodal", "gaussian"])
@pytest.mark.parametrize("code", [0, 2])
def test_pt_bookkeeping(oracle_lib, code, disorder):
    lat, ea = _sample(3, 6, 3, disorder)
    m = 6
    betas = 1.0 / geometric_temperatures(0.8, 2.0, m)[::-1]
    rep = oracle_lib.EaReplay(oracle_lib.ea_init(m, lat.n_sites, (1, 2)), ea.device_couplings(), 3, 6,
                              betas, (4, 5), code=code)
    out = rep.run(63, 6)
    for r in range(2):
        for c in range(m):
            s = rep.spins[r, c]
            if disorder == "bimodal":
                assert rep.energies[r * m + c] == int(ea.energy(s))
            else:  # running FP64 sum since the last bin edge
                assert rep.energies[r * m + c] == pytest.approx(ea.energy(s), abs=1e-9)
        assert sorted(rep.perm[r * m:(r + 1) * m]) == list(range(m))
        assert all(rep.kofc[r * m + rep.perm[r * m + k]] == k for k in range(m))
    assert np.all(out["swap_try"] == 63)
    assert out["acc"][:, :, 6].sum() == 63 * m
    assert np.all(out["label_hist"].sum(axis=2) <= 63)


def _enum_expect(lat, ea, T):
    n = lat.n_sites
    states = np.arange(1 << n)
    spins = (2 * ((states[:, None] >> np.arange(n)) & 1) - 1).astype(np.float64)
    energies = -(spins[:, lat.bonds[:, 0]] * spins[:, lat.bonds[:, 1]]) @ ea.couplings
    w = np.exp(-(energies - energies.min()) / T)
    w /= w.sum()
    corr = (spins * w[:, None]).T @ spins
    b = lat.bonds
    return w @ energies / n, np.sum(corr ** 2) / n ** 2, np.mean(corr[b[:, 0], b[:, 1]] ** 2)


@pytest.mark.parametrize("disorder", ["bimodal", "gaussian"])
@pytest.mark.parametrize("code", [0, 2])
def test_pt_matches_enumeration(oracle_lib, code, disorder):
    lat, ea = _sample(11, disorder=disorder)
    temps = geometric_temperatures(0.4, 2.5, 8)[::-1]
    rep = oracle_lib.EaReplay(oracle_lib.ea_init(8, 16, (9, 9)), ea.device_couplings(), 2, 4,
                              1.0 / temps, (5, 6), code=code)
    out = rep.run((1 << 16) - 1, 16)
    acc = out["acc"][-1]
    cnt = acc[:, 6]
    for k, T in enumerate(temps):
        e, q2, ql = _enum_expect(lat, ea, T)
        assert acc[k, 0] / cnt[k] == pytest.approx(e, abs=0.01)
        assert acc[k, 3] / cnt[k] == pytest.approx(q2, abs=0.01)
        assert acc[k, 1] / cnt[k] == pytest.approx(ql, abs=0.01)
    rate = out["swap_acc"].sum(0) / out["swap_try"].sum(0)
    assert np.all(rate > 0.3)
    assert out["trips"].sum() > 100
