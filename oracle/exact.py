"""ORACLE (test infrastructure) -- exact results the samplers are pulled against.

* Kaufman's finite torus, in the form of P. D. Beale, PRL 76, 78 (1996), as
  used by critmc/exact/kaufman.py:1-26.  Evaluated here in multiprecision
  (mpmath) with derivatives taken by mpmath's numerical differentiation, an
  evaluation route independent of the reference's analytic log-space one.
  Pinned by: the reference's stored vectors (results/e01_validation.json,
  results/e03_ising2d_fss.json), brute-force enumeration, the Onsager limit.
* Brute-force enumeration of small Ising / Potts / +-J lattices, restating
  the loops of critmc/exact/enumeration.py:25-136 in vectorised numpy.
  Pinned by closed forms (two-site, chain transfer matrix) in the tests.
* O(n) rings: transfer-operator energy of a periodic ring of N sites
  (critmc/exact/chains.py:23-78), from the Bessel-function eigenvalues.
"""

from __future__ import annotations

import itertools
from functools import lru_cache

import mpmath as mp
import numpy as np
from scipy import special

BETA_C = float(np.log(1.0 + np.sqrt(2.0)) / 2.0)


# ---------------------------------------------------------------------------
# Kaufman torus
def _ln_z(K, m: int, n: int):
    K = mp.mpf(K)
    c2, s2 = mp.cosh(2 * K), mp.sinh(2 * K)
    base = c2 * c2 / s2  # cosh 2K coth 2K
    gam = [2 * K + mp.log(mp.tanh(K))]
    for l in range(1, 2 * n):
        gam.append(mp.acosh(base - mp.cos(mp.pi * l / n)))
    z1 = z2 = z3 = z4 = mp.mpf(1)
    for r in range(n):
        go, ge = gam[2 * r + 1], gam[2 * r]
        z1 *= 2 * mp.cosh(m * go / 2)
        z2 *= 2 * mp.sinh(m * go / 2)
        z3 *= 2 * mp.cosh(m * ge / 2)
        z4 *= 2 * mp.sinh(m * ge / 2)
    return -mp.log(2) + (m * n / mp.mpf(2)) * mp.log(2 * s2) + mp.log(z1 + z2 + z3 + z4)


def kaufman(beta: float, L: int, M: int | None = None, dps: int = 40) -> dict[str, float]:
    """``lnZ``, energy ``u`` and specific heat ``c`` per site of the L x M torus."""
    m, n = int(L), int(L if M is None else M)
    with mp.workdps(dps):
        f = lambda k: _ln_z(k, m, n)  # noqa: E731
        b = mp.mpf(beta)
        lnz = f(b)
        d1 = mp.diff(f, b, 1)
        d2 = mp.diff(f, b, 2)
        N = m * n
        return {"lnZ": float(lnz), "u": float(-d1 / N), "c": float(b * b * d2 / N)}


def onsager_energy(beta: float) -> float:
    """Infinite-lattice energy per site (Onsager)."""
    K = beta
    k = 2.0 * np.sinh(2 * K) / np.cosh(2 * K) ** 2
    K1 = special.ellipk(k * k)
    return float(-1.0 / np.tanh(2 * K) * (1 + (2 * np.tanh(2 * K) ** 2 - 1) * (2 / np.pi) * K1))


# ---------------------------------------------------------------------------
# Enumeration
def _all_spins(n: int) -> np.ndarray:
    idx = np.arange(1 << n, dtype=np.int64)[:, None]
    return (((idx >> np.arange(n)) & 1) * 2 - 1).astype(np.int8)


@lru_cache(maxsize=16)
def _ising_levels_cached(bonds_key: bytes, n: int, weights_key: bytes | None):
    bonds = np.frombuffer(bonds_key, dtype=np.int64).reshape(-1, 2)
    s = _all_spins(n).astype(np.int64)
    if weights_key is None:
        e_b = -np.sum(s[:, bonds[:, 0]] * s[:, bonds[:, 1]], axis=1)
    else:
        J = np.frombuffer(weights_key, dtype=np.float64)
        e_b = -(s[:, bonds[:, 0]] * s[:, bonds[:, 1]]) @ J
    m = s.sum(axis=1)
    return e_b, m


def ising_states(bonds: np.ndarray, n: int, couplings: np.ndarray | None = None):
    """Bond energy and magnetisation of all 2^n states (n <= 24)."""
    if n > 24:
        raise ValueError("enumeration is limited to 24 spins")
    bonds = np.ascontiguousarray(bonds, dtype=np.int64)
    wk = None if couplings is None else np.ascontiguousarray(couplings, np.float64).tobytes()
    return _ising_levels_cached(bonds.tobytes(), n, wk)


def ising_average(bonds, n: int, beta: float, h: float, fn, couplings=None) -> float:
    """Exact <fn(E_b, M)> at (beta, h) with weight exp(-beta (E_b - h M))."""
    e_b, m = ising_states(bonds, n, couplings)
    x = -beta * (e_b - h * m)
    w = np.exp(x - x.max())
    return float(np.sum(w * fn(e_b.astype(np.float64), m.astype(np.float64))) / w.sum())


def potts_average(bonds, n: int, q: int, beta: float, fn) -> float:
    """Exact <fn(E, counts)> for the q-state Potts model, E = -#satisfied bonds."""
    if q ** n > 3 * 10 ** 7:
        raise ValueError("too many states to enumerate")
    bonds = np.asarray(bonds)
    states = np.array(list(itertools.product(range(q), repeat=n)), dtype=np.int8)[:, ::-1]
    e = -np.sum(states[:, bonds[:, 0]] == states[:, bonds[:, 1]], axis=1).astype(np.float64)
    counts = np.stack([(states == k).sum(axis=1) for k in range(q)], axis=1).astype(np.float64)
    x = -beta * e
    w = np.exp(x - x.max())
    return float(np.sum(w * fn(e, counts)) / w.sum())


# ---------------------------------------------------------------------------
# O(n) rings
def on_ring_energy(beta: float, n_sites: int, n_comp: int) -> float:
    """Energy per site of a periodic O(n) ring from the transfer-operator spectrum.

    The eigenvalues of the bond kernel exp(beta S.S') on the sphere are
    proportional to I_{l+n/2-1}(beta)/beta^{n/2-1} with degeneracy d_l;
    Z = sum_l d_l lambda_l^N.
    """
    return float(_on_ring(beta, n_sites, n_comp)[0])


def on_ring_specific_heat(beta: float, n_sites: int, n_comp: int) -> float:
    return float(_on_ring(beta, n_sites, n_comp)[1])


def _on_ring(beta: float, N: int, n: int, lmax: int = 80):
    nu0 = n / 2.0 - 1.0
    with mp.workdps(40):
        def lnz(b):
            tot = mp.mpf(0)
            for l in range(lmax):
                if n == 2:
                    deg = 1 if l == 0 else 2
                else:
                    deg = mp.binomial(l + n - 1, n - 1) - (mp.binomial(l + n - 3, n - 1) if l >= 2 else 0)
                lam = mp.besseli(l + nu0, b) / b ** nu0
                tot += deg * lam ** N
            return mp.log(tot)

        b = mp.mpf(beta)
        d1 = mp.diff(lnz, b, 1)
        d2 = mp.diff(lnz, b, 2)
    return -float(d1) / N, float(b * b * d2) / N
