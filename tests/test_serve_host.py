"""Host-side logic of the request loop (no GPU): job stream and accounting."""

import numpy as np

from critmc_amd.serve import Job, poisson_jobs


def test_poisson_stream_shape():
    jobs = poisson_jobs(5000, rate=100.0, seed=1)
    arr = np.array([j.arrival for j in jobs])
    assert np.all(np.diff(arr) > 0)
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert abs(gaps.mean() - 0.01) < 0.001  # exponential gaps with mean 1/rate
    models = [j.model for j in jobs]
    for m in ("ising", "potts", "xy"):
        assert abs(models.count(m) / len(jobs) - 1 / 3) < 0.03
    Ls = np.array([j.L for j in jobs])
    assert set(Ls) <= {16, 32, 64, 128, 256, 512}
    sw = np.array([j.sweeps for j in jobs])
    assert sw.min() >= 1000 and sw.max() <= 100000
    assert abs(np.mean(np.log10(sw)) - 4.0) < 0.05  # log-uniform on [1e3, 1e5]
    bc = {"ising": np.log(1 + np.sqrt(2)) / 2, "potts": np.log(1 + np.sqrt(10)), "xy": 1 / 0.89294}
    for j in jobs:
        assert abs(j.beta / bc[j.model] - 1) <= 0.1 + 1e-12


def test_footprints():
    j = Job("potts", 64, 1.4, 1000, seed=0)
    assert j.state_bytes() == 64 * 64 + 64
    assert j.series_bytes() == 1000 * (4 + 40)
    x = Job("xy", 64, 1.1, 10, seed=0)
    assert x.state_bytes() == 64 * 64 * 8 + 64
