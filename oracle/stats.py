"""ORACLE (test infrastructure) -- error analysis used to judge statistical pulls.

``gamma_mean``: U. Wolff, "Monte Carlo errors with less errors", Comput.
Phys. Commun. 156, 143 (2004), for the mean of one series, as used for the
"Gamma-method mean of e" of the reference (critmc/stats/autocorr.py:363-522):
automatic window W = first W with exp(-W/tau_W) - tau_W/sqrt(W N) < 0,
tau_W = S / ln(1 + 1/G(W)), G(W) = sum_{t<=W} rho(t), then the bias
correction Gamma(t) += C_F / N.  Pinned in the tests by the AR(1) closed form
tau_int = (1+phi) / (2(1-phi)) and by a direct loop transcription.

``batch_mean_error``: 64 batch means, the reference test suite's own
estimator (critmc tests/conftest.py:6-17).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class GammaResult:
    value: float
    error: float
    tau_int: float
    window: int
    n: int


def _autocov(x: np.ndarray, t_max: int) -> np.ndarray:
    n = x.size
    size = 1 << int(np.ceil(np.log2(2 * n)))
    f = np.fft.rfft(x, size)
    acf = np.fft.irfft(f * np.conj(f), size)[: t_max + 1]
    return acf / (n - np.arange(t_max + 1))


def gamma_mean(x, S: float = 1.5) -> GammaResult:
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    mean = float(x.mean())
    d = x - mean
    t_max = n // 2
    gamma = _autocov(d, t_max)
    if not gamma[0] > 0:
        return GammaResult(mean, 0.0, 0.5, 0, n)
    w_opt = t_max
    g_int = 0.0
    for w in range(1, t_max + 1):
        g_int += gamma[w] / gamma[0]
        tau_w = S / np.log1p(1.0 / g_int) if g_int > 0 else 1e-300
        tau_w = max(tau_w, 1e-300)
        if np.exp(-w / tau_w) - tau_w / np.sqrt(w * n) < 0:
            w_opt = w
            break
    c_f = gamma[0] + 2.0 * gamma[1 : w_opt + 1].sum()
    gamma = gamma + c_f / n
    c_f = gamma[0] + 2.0 * gamma[1 : w_opt + 1].sum()
    return GammaResult(mean, float(np.sqrt(c_f / n)) if c_f > 0 else float("nan"),
                       float(c_f / (2.0 * gamma[0])), int(w_opt), n)


def batch_mean_error(x, n_batches: int = 64) -> tuple[float, float]:
    """Mean of ``x`` and its standard error from ``n_batches`` equal contiguous batches.

    Batch means of batches much longer than the autocorrelation time are
    independent, so their sample variance over ``n_batches`` estimates the
    variance of the mean.  The first ``len(x) mod n_batches`` samples, the
    least equilibrated, are left out of the batches (not of the mean).
    """
    x = np.asarray(x, dtype=np.float64)
    per_batch = len(x) // n_batches
    means = x[len(x) % n_batches:].reshape(n_batches, per_batch).mean(axis=1)
    return float(x.mean()), float(np.sqrt(np.var(means, ddof=1) / n_batches))


def pull(x, exact: float, method: str = "batch") -> tuple[float, float, float]:
    """(mean, error, z) of a series against an exact value."""
    if method == "gamma":
        r = gamma_mean(x)
        mean, err = r.value, r.error
    else:
        mean, err = batch_mean_error(x)
    z = (mean - exact) / err if err > 0 else 0.0
    return mean, err, z
