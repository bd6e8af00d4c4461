r"""Edwards-Anderson spin glasses with parallel tempering on the GPU.

Interface of critmc/glass.py:45-313.  ``H = -sum_<ij> J_ij s_i s_j``; two
independent replica sets run at every temperature and give the overlap
``q``, the link overlap ``q_l`` and ``|q(k_min)|^2``, accumulated in
logarithmic bins of PT steps (bin b covers steps [2^b - 1, 2^(b+1) - 1)).

+-1 couplings (``disorder="bimodal"``) run on int8 couplings with exact
integer energies; any other couplings (``disorder="gaussian"``) run with FP64
couplings, fields and energies, recomputed at every bin edge as the reference
does.  :class:`TemperingBatch` advances many samples of uneven size in the same
launches (samples of one batch share the coupling kind); :class:`ParallelTempering`
is the reference's one-sample driver on top of it.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .kernels import ea as k_ea
from .lattice import Lattice, require_checkerboard
from .rng import chain_key, make_rng

__all__ = ["EdwardsAnderson", "TemperingRun", "ParallelTempering", "TemperingBatch",
           "DisorderAverage", "feedback_temperatures", "geometric_temperatures"]


class EdwardsAnderson:
    """One disorder realisation (critmc/glass.py:45-90)."""

    def __init__(self, lattice: Lattice, couplings: np.ndarray | None = None, *,
                 disorder: str = "gaussian", seed=None):
        self.lattice = lattice
        self.disorder = disorder
        if couplings is None:
            rng = make_rng(seed)
            if disorder == "gaussian":
                couplings = rng.standard_normal(lattice.n_bonds)
            elif disorder == "bimodal":
                couplings = rng.choice([-1.0, 1.0], size=lattice.n_bonds)
            else:
                raise ValueError("disorder must be 'gaussian' or 'bimodal'")
        couplings = np.asarray(couplings, dtype=np.float64)
        if couplings.shape != (lattice.n_bonds,):
            raise ValueError("need one coupling per bond")
        self.couplings = couplings
        self.jn = self._neighbour_couplings()

    def _neighbour_couplings(self) -> np.ndarray:
        lat = self.lattice
        jn = np.full(lat.nbr.shape, np.nan)
        axis = lat.bond_axis
        jn[lat.bonds[:, 0], 2 * axis] = self.couplings
        jn[lat.bonds[:, 1], 2 * axis + 1] = self.couplings
        if np.isnan(jn).any():
            raise ValueError("lattice neighbour table is not in forward/backward pair form")
        return jn

    @property
    def n_sites(self) -> int:
        return self.lattice.n_sites

    def energy(self, state: np.ndarray) -> float:
        b = self.lattice.bonds
        s = np.asarray(state, dtype=np.float64)
        return float(-np.sum(self.couplings * s[b[:, 0]] * s[b[:, 1]]))

    def device_couplings(self) -> np.ndarray:
        """Forward-bond couplings in the device order (axis-major, one per site and axis)."""
        require_checkerboard(self.lattice)
        if self.lattice.dim not in (2, 3) or len(set(self.lattice.shape)) != 1:
            raise ValueError("GPU tempering supports L x L and L x L x L lattices")
        return self.couplings.copy()


def geometric_temperatures(t_min: float, t_max: float, n: int) -> np.ndarray:
    """``T_min (T_max/T_min)^(k/(n-1))`` (critmc/glass.py:93-95)."""
    return t_min * (t_max / t_min) ** (np.arange(n) / (n - 1))


@dataclass
class TemperingRun:
    """Accumulated measurements of one sample (critmc/glass.py:98-131)."""

    temperatures: np.ndarray
    n_sites: int
    z: int
    acc: np.ndarray
    swap_rate: np.ndarray
    label_hist: np.ndarray
    trips: np.ndarray
    wall_time: float = 0.0
    meta: dict = field(default_factory=dict)

    def bin_means(self) -> dict[str, np.ndarray]:
        cnt = self.acc[:, :, k_ea.OBS_COUNT]
        with np.errstate(invalid="ignore", divide="ignore"):
            return {name: self.acc[:, :, col] / cnt for name, col in (
                ("e", k_ea.OBS_E), ("ql", k_ea.OBS_QL), ("absq", k_ea.OBS_Q),
                ("q2", k_ea.OBS_Q2), ("q4", k_ea.OBS_Q4), ("qk2", k_ea.OBS_QK2))}

    def equilibrium(self) -> dict[str, np.ndarray]:
        return {k: v[-1] for k, v in self.bin_means().items()}


class TemperingBatch:
    """Many samples advanced together (same dimension and ladder size, any even L)."""

    def __init__(self, models: list[EdwardsAnderson], temperatures, *, keys=None, seeds=None,
                 sweeps_per_step: int = 1, update: str = "heatbath", spins=None):
        if not models:
            raise ValueError("no samples")
        d = models[0].lattice.dim
        if any(md.lattice.dim != d for md in models):
            raise ValueError("all samples of a batch need the same dimension")
        temps = np.asarray(temperatures, dtype=np.float64)
        if temps.ndim == 1:
            temps = np.tile(temps, (len(models), 1))
        temps = -np.sort(-temps, axis=1)  # hot -> cold
        self.models = models
        self.temperatures = temps
        self.betas = 1.0 / temps
        self.sweeps_per_step = int(sweeps_per_step)
        self.update = update
        if keys is None:
            seeds = seeds if seeds is not None else [None] * len(models)
            keys = [chain_key(make_rng(sd)) for sd in seeds]
        self.engine = k_ea.EaBatch([md.lattice.L for md in models], d,
                                   [md.device_couplings() for md in models], self.betas, keys,
                                   spins=spins)

    def run(self, n_bins: int, *, record: bool = False) -> list[TemperingRun]:
        n_steps = (1 << n_bins) - 1
        t0 = time.perf_counter()
        out = self.engine.run(n_steps, n_bins, update=self.update,
                              sweeps_per_step=self.sweeps_per_step, record=record)
        acc = out["acc"].cpu().numpy()
        sacc = out["swap_acc"].cpu().numpy().sum(axis=1)
        stry = out["swap_try"].cpu().numpy().sum(axis=1)
        hist = out["label_hist"].cpu().numpy().sum(axis=1).astype(np.int64)
        trips = out["trips"].cpu().numpy().astype(np.int64)
        wall = time.perf_counter() - t0
        self.last_series = {k: out[k].cpu().numpy() for k in ("E", "q") if k in out}
        runs = []
        for s, md in enumerate(self.models):
            with np.errstate(invalid="ignore", divide="ignore"):
                rate = sacc[s] / stry[s]
            runs.append(TemperingRun(self.temperatures[s].copy(), md.n_sites, md.lattice.z,
                                     acc[s], rate, hist[s], trips[s], wall,
                                     {"n_bins": n_bins, "sweeps_per_step": self.sweeps_per_step,
                                      "update": self.update, "backend": "hip"}))
        return runs


class ParallelTempering:
    """Parallel tempering of one sample with two replica sets (critmc/glass.py:134-208).

    The start configurations are drawn from ``seed`` exactly as the reference
    does; the chain's Philox key is drawn next from the same generator.
    ``update`` is the reference's Metropolis by default; ``"heatbath"`` is the
    heat-bath rule with the local field.
    """

    def __init__(self, model: EdwardsAnderson, temperatures, *, seed=None,
                 sweeps_per_step: int = 1, update: str = "metropolis"):
        temps = np.sort(np.asarray(temperatures, dtype=np.float64))[::-1]
        self.model = model
        self.temperatures = temps
        self.betas = 1.0 / temps
        self.sweeps_per_step = int(sweeps_per_step)
        self.rng = make_rng(seed)
        m, n = len(temps), model.n_sites
        spins = (2 * self.rng.integers(0, 2, size=(2, m, n)) - 1).astype(np.int8)
        key = chain_key(self.rng)
        self._batch = TemperingBatch([model], temps, keys=[key], sweeps_per_step=sweeps_per_step,
                                     update=update, spins=[spins])

    @property
    def spins(self) -> np.ndarray:
        return self._batch.engine.configurations(0)

    @property
    def energies(self) -> np.ndarray:
        return self._batch.engine.energies.cpu().numpy().reshape(2, -1).astype(np.float64)

    @property
    def perm(self) -> np.ndarray:
        return self._batch.engine.perm.cpu().numpy().reshape(2, -1).astype(np.int64)

    @property
    def labels(self) -> np.ndarray:
        return self._batch.engine.labels.cpu().numpy().reshape(2, -1).astype(np.int64)

    def run(self, n_bins: int) -> TemperingRun:
        return self._batch.run(n_bins)[0]


def feedback_temperatures(temperatures: np.ndarray, label_hist: np.ndarray) -> np.ndarray:
    """One feedback-optimisation step of the ladder (critmc/glass.py:211-242).

    Katzgraber, Trebst, Huse & Troyer, J. Stat. Mech. P03018 (2006).  ``f(T)``
    is the fraction of configurations at T whose last visited end was the
    hottest one (``label_hist[k] = (visits from the hot end, from the cold
    end)``); it is taken as 1/2 where T saw neither, and made non-decreasing
    in T.  The optimal temperature density is sqrt(df/dT / dT), whose integral
    over an interval of the current ladder is sqrt(df): the new temperatures
    sit at equal steps of that cumulative weight.  The end points stay, and
    the result is in ascending order.
    """
    temps = np.asarray(temperatures, dtype=np.float64)
    order = np.argsort(temps)
    ladder = temps[order]
    hot, cold = (np.asarray(label_hist, dtype=np.float64)[order, c] for c in (0, 1))
    visited = hot + cold
    frac = np.divide(hot, visited, out=np.full_like(hot, 0.5), where=visited > 0)
    frac = np.maximum.accumulate(frac)
    cumulative = np.append(0.0, np.cumsum(np.sqrt(np.clip(np.diff(frac), 1e-6, None))))
    profile = cumulative / cumulative[-1]
    new = np.interp(np.linspace(0.0, 1.0, ladder.size), profile, ladder)
    new[[0, -1]] = ladder[[0, -1]]
    return new


def _estimates(mean: dict[str, np.ndarray], n_sites: int, L: int) -> dict[str, np.ndarray]:
    """Binder ratio, susceptibility and xi_L / L from disorder means (any leading axes):
    g = (3 - [q^4] / [q^2]^2) / 2, chi_SG = N [q^2],
    xi_L = sqrt([q^2] / [|q(k_min)|^2] - 1) / (2 sin(k_min / 2)), k_min = 2 pi / L."""
    q2, q4, qk2 = mean["q2"], mean["q4"], mean["qk2"]
    with np.errstate(invalid="ignore", divide="ignore"):
        ratio = np.clip(q2 / qk2 - 1.0, 0.0, None)
        xi = np.sqrt(ratio) / (2.0 * np.sin(np.pi / L))
        g = 0.5 * (3.0 - q4 / (q2 * q2))
    return {"e": mean["e"], "ql": mean["ql"], "g": g, "chi_sg": n_sites * q2, "xi_L": xi / L}


class DisorderAverage:
    """Disorder averages over samples run at common temperatures (critmc/glass.py:245-313).

    ``samples[name]`` holds each sample's last-bin (equilibrium) thermal
    average, shape ``(samples, temperatures)``.
    """

    def __init__(self, runs: list[TemperingRun], L: int):
        if not runs:
            raise ValueError("no runs")
        self.runs = runs
        self.L = int(L)
        first = runs[0]
        self.temperatures = first.temperatures
        self.n_sites = first.n_sites
        self.z = first.z
        per_run = [r.equilibrium() for r in runs]
        self.samples = {name: np.stack([eq[name] for eq in per_run]) for name in per_run[0]}

    @staticmethod
    def _observables(s, n_sites: int, L: int) -> dict[str, np.ndarray]:
        """Estimates from per-sample values ``s[name]`` (samples first)."""
        return _estimates({name: v.mean(axis=0) for name, v in s.items()}, n_sites, L)

    def observables(self, n_boot: int = 500, seed: int = 0):
        """``{name: (estimate, bootstrap error)}``; the bootstrap resamples disorder
        realisations, all ``n_boot`` resamples drawn and estimated at once."""
        full = self._observables(self.samples, self.n_sites, self.L)
        n = len(self.runs)
        picks = make_rng(seed).integers(0, n, size=(n_boot, n))
        resampled = _estimates({name: v[picks].mean(axis=1) for name, v in self.samples.items()},
                               self.n_sites, self.L)
        return {name: (full[name], resampled[name].std(axis=0, ddof=1)) for name in full}

    def _bins(self, name: str) -> np.ndarray:
        """Per-bin thermal means of every sample, shape ``(samples, bins, temperatures)``."""
        return np.stack([r.bin_means()[name] for r in self.runs])

    def identity_residuals(self) -> np.ndarray:
        """``e + (z/2)(1 - q_l)/T`` per bin and sample, shape ``(bins, samples, temperatures)``.

        The Gaussian-disorder identity (Katzgraber, Palassini & Young, PRB 63,
        184422 (2001)) makes its disorder average vanish in equilibrium.
        """
        residual = self._bins("e") + 0.5 * self.z * (1.0 - self._bins("ql")) / self.temperatures
        return residual.swapaxes(0, 1)

    def equilibration(self) -> dict[str, np.ndarray]:
        """Per-bin disorder means of e and of -(z/2)(1 - q_l)/T, and the standard
        error of their difference over samples."""
        residual = self.identity_residuals()
        n = residual.shape[1]
        return {"e": self._bins("e").mean(axis=0),
                "e_from_ql": -0.5 * self.z * (1.0 - self._bins("ql").mean(axis=0)) / self.temperatures,
                "residual_error": residual.std(axis=1, ddof=1) / np.sqrt(n)}
