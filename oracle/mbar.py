"""ORACLE (test infrastructure only): numpy restatement of the reference's
multi-histogram (MBAR) estimator, critmc/reweight.py:159-345.

Distinct (energy, run weight) pairs with multiplicities (:210-223), the convex
objective with gradient and Hessian (:231-247), the chained single-histogram
initial guess (:249-257), Newton with backtracking in the gauge f_0 = 0
(:259-286), and log Z at a target beta (:313-320).  Pinned in
tests/test_oracle_mbar.py by closed forms (one run = single-histogram
reweighting; two runs at one beta; exact samples of a 4x4 Ising model drawn
from its enumerated energy distribution recover ln Z ratios).
"""

from __future__ import annotations

import numpy as np
from scipy.special import logsumexp


class Mbar:
    def __init__(self, energies, betas, inefficiency=None):
        self.betas = np.asarray(betas, dtype=np.float64)
        K = len(self.betas)
        self.counts = np.array([len(e) for e in energies], dtype=np.float64)
        self.g = np.ones(K) if inefficiency is None else np.asarray(inefficiency, dtype=float)
        self.energy = np.concatenate([np.asarray(e, dtype=np.float64) for e in energies])
        run = np.repeat(np.arange(K), self.counts.astype(int))
        self.sample_logw = -np.log(self.g)[run]
        self.run_index = run
        self.shift = float(self.energy.mean())
        order = np.lexsort((self.sample_logw, self.energy))
        e, lw = self.energy[order], self.sample_logw[order]
        first = np.ones(len(order), dtype=bool)
        first[1:] = (e[1:] != e[:-1]) | (lw[1:] != lw[:-1])
        start = np.flatnonzero(first)
        cnt = np.diff(start, append=len(order))
        self.u_energy = e[start] - self.shift
        self.u_logmult = np.log(cnt) + lw[start]
        self.log_n = np.log(self.counts / self.g)
        self.fs = np.zeros(K)

    def log_denominator(self, f):
        a = self.log_n[None, :] + f[None, :] - self.betas[None, :] * self.u_energy[:, None]
        return logsumexp(a, axis=1)

    def objective(self, f):
        ld = self.log_denominator(f)
        value = np.sum(np.exp(self.u_logmult) * ld) - np.sum(np.exp(self.log_n) * f)
        log_w = (self.log_n[None, :] + f[None, :] - self.betas[None, :] * self.u_energy[:, None]
                 - ld[:, None])
        w = np.exp(log_w)
        mult = np.exp(self.u_logmult)
        col = mult @ w
        grad = col - np.exp(self.log_n)
        hess = np.diag(col) - (w * mult[:, None]).T @ w
        return value, grad, hess

    def initial_guess(self):
        order = np.argsort(self.betas)
        f = np.zeros(len(self.betas))
        for a, b in zip(order[:-1], order[1:]):
            e = self.energy[self.run_index == a] - self.shift
            f[b] = f[a] - (logsumexp(-(self.betas[b] - self.betas[a]) * e) - np.log(len(e)))
        return f - f[0]

    def solve(self, rtol=1e-10, max_iter=100):
        if len(self.betas) == 1:
            return self.f
        f = self.initial_guess()
        scale = float(np.mean(self.counts / self.g))
        for _ in range(max_iter):
            value, grad, hess = self.objective(f)
            if np.max(np.abs(grad[1:])) < rtol * scale:
                break
            step = np.zeros_like(f)
            step[1:] = -np.linalg.solve(hess[1:, 1:], grad[1:])
            t, slope = 1.0, float(grad @ step)
            while True:
                trial = f + t * step
                if self.objective(trial)[0] <= value + 1e-4 * t * slope or t < 1e-12:
                    break
                t *= 0.5
            f = trial
        self.fs = f - f[0]
        return self.f

    @property
    def f(self):
        return self.fs + (self.betas - self.betas[0]) * self.shift

    def log_z(self, beta):
        beta = np.atleast_1d(np.asarray(beta, dtype=np.float64))
        ld = self.log_denominator(self.fs)
        out = np.array([logsumexp(self.u_logmult - b * self.u_energy - ld) for b in beta])
        return out - (beta - self.betas[0]) * self.shift
