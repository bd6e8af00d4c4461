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

    Katzgraber, Trebst, Huse & Troyer, J. Stat. Mech. P03018 (2006): with f(T)
    the fraction of configurations at T that last visited the hottest end, new
    temperatures sit at equal quantiles of sum sqrt(df); end points are kept.
    """
    temps = np.asarray(temperatures, dtype=np.float64)
    order = np.argsort(temps)
    t = temps[order]
    up = label_hist[order, 0].astype(float)
    down = label_hist[order, 1].astype(float)
    with np.errstate(invalid="ignore", divide="ignore"):
        f = up / (up + down)
    f = np.maximum.accumulate(np.nan_to_num(f, nan=0.5))
    df = np.maximum(np.diff(f), 1e-6)
    cum = np.concatenate([[0.0], np.cumsum(np.sqrt(df))])
    cum /= cum[-1]
    new = np.interp(np.linspace(0.0, 1.0, len(t)), cum, t)
    new[0], new[-1] = t[0], t[-1]
    return new


class DisorderAverage:
    """Disorder averages over samples at common temperatures (critmc/glass.py:245-313)."""

    def __init__(self, runs: list[TemperingRun], L: int):
        if not runs:
            raise ValueError("no runs")
        self.runs = runs
        self.L = int(L)
        self.temperatures = runs[0].temperatures
        self.n_sites = runs[0].n_sites
        self.z = runs[0].z
        eq = [r.equilibrium() for r in runs]
        self.samples = {k: np.array([e[k] for e in eq]) for k in eq[0]}

    @staticmethod
    def _observables(s, n_sites: int, L: int) -> dict[str, np.ndarray]:
        q2, q4, qk2 = s["q2"].mean(axis=0), s["q4"].mean(axis=0), s["qk2"].mean(axis=0)
        kmin = 2.0 * np.pi / L
        with np.errstate(invalid="ignore"):
            xi = np.sqrt(np.clip(q2 / qk2 - 1.0, 0.0, None)) / (2.0 * np.sin(kmin / 2.0))
        return {"e": s["e"].mean(axis=0), "ql": s["ql"].mean(axis=0),
                "g": 0.5 * (3.0 - q4 / q2 ** 2), "chi_sg": n_sites * q2, "xi_L": xi / L}

    def observables(self, n_boot: int = 500, seed: int = 0):
        rng = make_rng(seed)
        full = self._observables(self.samples, self.n_sites, self.L)
        n = len(self.runs)
        boots = {k: [] for k in full}
        for _ in range(n_boot):
            idx = rng.integers(0, n, size=n)
            b = self._observables({k: v[idx] for k, v in self.samples.items()}, self.n_sites, self.L)
            for k in full:
                boots[k].append(b[k])
        return {k: (full[k], np.std(np.array(boots[k]), axis=0, ddof=1)) for k in full}

    def identity_residuals(self) -> np.ndarray:
        e = np.array([r.bin_means()["e"] for r in self.runs])
        ql = np.array([r.bin_means()["ql"] for r in self.runs])
        res = e + (self.z / 2.0) * (1.0 - ql) / self.temperatures[None, None, :]
        return np.transpose(res, (1, 0, 2))

    def equilibration(self) -> dict[str, np.ndarray]:
        e = np.array([r.bin_means()["e"] for r in self.runs]).mean(axis=0)
        ql = np.array([r.bin_means()["ql"] for r in self.runs]).mean(axis=0)
        res = self.identity_residuals()
        return {"e": e, "e_from_ql": -(self.z / 2.0) * (1.0 - ql) / self.temperatures[None, :],
                "residual_error": res.std(axis=1, ddof=1) / np.sqrt(res.shape[1])}
