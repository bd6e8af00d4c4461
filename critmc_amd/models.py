"""Spin models bound to the HIP kernels.

Mirrors critmc/models.py:32-351: a model creates states, validates update
schedules, advances a chain through its kernel (``advance``, the plug-in seam
of models.py:58-68) and turns raw measurement streams into per-site
observables (``derive``).  States are created on the host exactly as the
reference does (same generator calls, so the same seed gives the same start)
and then live on the GPU; ``advance`` takes the device state and the chain's
Philox record where the reference takes the numpy array and ``Generator``.

Conventions (J = k_B = 1): Ising ``H = -sum s_i s_j - h sum s_i``; Potts
``H = -sum delta(s_i, s_j)``; O(n) ``H = -sum S_i . S_j``.
"""

from __future__ import annotations

import math
from abc import ABC, abstractmethod
from dataclasses import dataclass, field
from typing import Any

import numpy as np
import torch

from .device import device
from .kernels import ising as k_ising
from .kernels import on as k_on
from .kernels import potts as k_potts
from .kernels import table as k_table
from .kernels.chains import Chains
from .kernels.common import CLUSTER_CODES, GPU_UPDATES, METROPOLIS_RANDOM, N_AUX, UPDATE_CODES
from .lattice import Lattice, checkerboard_ok, make_lattice, require_checkerboard

__all__ = ["Model", "Ising", "Potts", "ON", "XY", "Heisenberg", "build_model", "derive",
           "DeviceState"]


@dataclass
class DeviceState:
    """A chain's device-resident state: spins plus running integer bookkeeping."""

    spins: torch.Tensor
    extra: dict[str, torch.Tensor] = field(default_factory=dict)


class Model(ABC):
    kind: str = ""
    updates: tuple[str, ...] = ()

    def __init__(self, lattice: Lattice):
        self.lattice = lattice

    @property
    def n_sites(self) -> int:
        return self.lattice.n_sites

    @abstractmethod
    def initial_state(self, start: str, rng: np.random.Generator) -> np.ndarray: ...

    @abstractmethod
    def energy(self, state: np.ndarray) -> float: ...

    @abstractmethod
    def begin(self, state: np.ndarray) -> dict[str, Any]: ...

    @abstractmethod
    def to_device(self, state: np.ndarray, running: dict[str, Any]) -> DeviceState: ...

    @abstractmethod
    def from_device(self, ds: DeviceState) -> np.ndarray: ...

    @abstractmethod
    def advance(self, ds: DeviceState, running: dict[str, Any], beta: float, chains: Chains,
                schedule: np.ndarray, n_meas: int, ws: Any) -> dict[str, np.ndarray]:
        """Run ``n_meas`` schedule applications on the GPU and return the raw streams."""

    def workspace(self) -> Any:
        return None

    def check_updates(self, names: list[str]) -> None:
        """Same errors as critmc/models.py:70-77, plus the GPU path's own limits."""
        bad = [n for n in names if n not in self.updates]
        if bad:
            gpu = GPU_UPDATES.get(self.kind, ())
            unsupported = [n for n in bad if n in UPDATE_CODES and n not in gpu]
            if unsupported and len(unsupported) == len(bad):
                raise ValueError(
                    f"{type(self).__name__} updates {unsupported} have no checkerboard GPU kernel; "
                    f"available: {list(self.updates)}"
                )
            raise ValueError(
                f"{type(self).__name__} does not support updates {bad}; "
                f"available: {list(self.updates)}"
            )
        self._check_lattice()

    def _check_lattice(self) -> None:
        require_checkerboard(self.lattice)

    def sweep_order(self, schedule) -> str:
        """``"checkerboard"``, or ``"sequential"`` where the schedule runs in the
        reference's own site orders (typewriter / random sites)."""
        return "checkerboard"

    def spec(self) -> dict[str, Any]:
        return {"model": self.kind, "lattice": self.lattice.name, "L": self.lattice.L}

    def meta(self) -> dict[str, Any]:
        return {**self.spec(), "N": self.n_sites, "dim": self.lattice.dim,
                "n_bonds": self.lattice.n_bonds}

    def __repr__(self) -> str:
        return f"{type(self).__name__}({self.lattice!r})"


class _AnyLattice:
    """Discrete models run on any pair-form lattice: the checkerboard kernels
    where they apply, the neighbour-table kernels otherwise (and for
    ``metropolis-random``, whose sites are drawn at random)."""

    def _check_lattice(self) -> None:
        if not checkerboard_ok(self.lattice):
            k_table.check_lattice(self.lattice)

    def _sequential(self, schedule) -> bool:
        codes = np.asarray(schedule)[:, 0]
        return not checkerboard_ok(self.lattice) or bool((codes == METROPOLIS_RANDOM).any())

    def sweep_order(self, schedule) -> str:
        return "sequential" if self._sequential(schedule) else "checkerboard"

    def _table_ws(self, ds: DeviceState, q: int):
        if "table_ws" not in ds.extra:
            ds.extra["table_ws"] = k_table.Workspace(self.n_sites, ds.spins.shape[0], q,
                                                     ds.spins.device)
        return k_table.neighbours(self.lattice, ds.spins.device), ds.extra["table_ws"]


def _ising_bond_energy(s: np.ndarray, bonds: np.ndarray) -> int:
    return -int(np.sum(s[bonds[:, 0]].astype(np.int64) * s[bonds[:, 1]]))


class Ising(_AnyLattice, Model):
    """Ising model with optional uniform field ``h`` (critmc/models.py:95-162)."""

    kind = "ising"
    updates = GPU_UPDATES["ising"]

    def __init__(self, lattice: Lattice, h: float = 0.0):
        super().__init__(lattice)
        self.h = float(h)
        if not math.isfinite(self.h):
            raise ValueError(f"field h must be finite, got {h}")

    def check_updates(self, names):
        cluster = {"wolff", "sw", "swendsen-wang"}
        if self.h != 0.0 and cluster.intersection(names):
            raise ValueError("cluster updates require zero field h")
        super().check_updates(names)

    def initial_state(self, start, rng):
        n = self.n_sites
        if start == "random":
            return (2 * rng.integers(0, 2, size=n) - 1).astype(np.int8)
        if start in ("ordered", "up"):
            return np.ones(n, dtype=np.int8)
        if start == "down":
            return -np.ones(n, dtype=np.int8)
        raise ValueError(f"unknown start {start!r}")

    def energy(self, state):
        s = np.asarray(state)
        return float(_ising_bond_energy(s, self.lattice.bonds)) - self.h * float(s.sum())

    def begin(self, state):
        s = np.asarray(state)
        return {"e_b": _ising_bond_energy(s, self.lattice.bonds), "m": int(s.astype(np.int64).sum())}

    def to_device(self, state, running):
        dev = device()
        spins = torch.from_numpy(np.ascontiguousarray(state, dtype=np.int8)).to(dev).reshape(1, -1)
        e = torch.tensor([running["e_b"]], dtype=torch.int32, device=dev)
        m = torch.tensor([running["m"]], dtype=torch.int32, device=dev)
        return DeviceState(spins, {"e": e, "m": m})

    def from_device(self, ds):
        return ds.spins.reshape(-1).cpu().numpy().copy()

    def workspace(self):
        return k_ising.Workspace(1)

    def advance(self, ds, running, beta, chains, schedule, n_meas, ws):
        if self._sequential(schedule):
            nbr, tws = self._table_ws(ds, 0)
            e_out, m_out, aux_d = k_table.ising_run(ds.spins, nbr, [float(beta)], self.h, chains,
                                                    schedule, n_meas, ds.extra["e"],
                                                    ds.extra["m"], ws=tws)
            aux = aux_d[0].cpu().numpy()
        elif np.isin(np.asarray(schedule)[:, 0], CLUSTER_CODES).any():
            if "cluster_ws" not in ds.extra:
                ds.extra["cluster_ws"] = k_ising.ClusterWorkspace(self.lattice.shape, 1)
            e_out, m_out, aux_d = k_ising.run_cluster(
                ds.spins, self.lattice.shape, [float(beta)], self.h, chains, schedule, n_meas,
                ds.extra["e"], ds.extra["m"], ws=ds.extra["cluster_ws"])
            aux = aux_d[0].cpu().numpy()
        else:
            e_out, m_out, acc = k_ising.run(ds.spins, self.lattice.shape, [float(beta)], self.h,
                                            chains, schedule, n_meas, ds.extra["e"],
                                            ds.extra["m"], ws=ws, form=_FORM[0])
            aux = np.zeros((n_meas, N_AUX), dtype=np.int64)
            aux[:, 0] = acc[0].cpu().numpy()
        e_b = e_out[0].cpu().numpy().astype(np.int64)
        m = m_out[0].cpu().numpy().astype(np.int64)
        running["e_b"] = int(e_b[-1])
        running["m"] = int(m[-1])
        energy = e_b - self.h * m if self.h != 0.0 else e_b
        return {"E": energy, "Eb": e_b, "M": m, "aux": aux}

    def spec(self):
        return {**super().spec(), "h": self.h}


def _potts_bond_energy(s: np.ndarray, bonds: np.ndarray) -> int:
    return -int(np.sum(s[bonds[:, 0]] == s[bonds[:, 1]]))


class Potts(_AnyLattice, Model):
    """q-state Potts model (critmc/models.py:165-222)."""

    kind = "potts"
    updates = GPU_UPDATES["potts"]

    def __init__(self, lattice: Lattice, q: int):
        super().__init__(lattice)
        if q < 2 or q > 127:
            raise ValueError(f"q must be in [2, 127], got {q}")
        self.q = int(q)

    def initial_state(self, start, rng):
        n = self.n_sites
        if start == "random":
            return rng.integers(0, self.q, size=n).astype(np.int8)
        if start in ("ordered", "up"):
            return np.zeros(n, dtype=np.int8)
        raise ValueError(f"unknown start {start!r}")

    def energy(self, state):
        return float(_potts_bond_energy(np.asarray(state), self.lattice.bonds))

    def begin(self, state):
        s = np.asarray(state)
        return {"energy": _potts_bond_energy(s, self.lattice.bonds),
                "counts": np.bincount(s, minlength=self.q).astype(np.int64)}

    def to_device(self, state, running):
        dev = device()
        spins = torch.from_numpy(np.ascontiguousarray(state, dtype=np.int8)).to(dev).reshape(1, -1)
        e = torch.tensor([running["energy"]], dtype=torch.int32, device=dev)
        c = torch.from_numpy(np.asarray(running["counts"], dtype=np.int32)).to(dev).reshape(1, -1)
        return DeviceState(spins, {"e": e, "counts": c})

    def from_device(self, ds):
        return ds.spins.reshape(-1).cpu().numpy().copy()

    def workspace(self):
        return k_potts.Workspace(1)

    def advance(self, ds, running, beta, chains, schedule, n_meas, ws):
        if self._sequential(schedule):
            nbr, tws = self._table_ws(ds, self.q)
            e_out, counts, aux_d = k_table.potts_run(ds.spins, nbr, self.q, [float(beta)], chains,
                                                     schedule, n_meas, ds.extra["e"],
                                                     ds.extra["counts"], ws=tws)
            aux = aux_d[0].cpu().numpy()
        elif np.isin(np.asarray(schedule)[:, 0], CLUSTER_CODES).any():
            if "cluster_ws" not in ds.extra:
                ds.extra["cluster_ws"] = k_potts.ClusterWorkspace(self.lattice.shape, 1, self.q)
            e_out, counts, aux_d = k_potts.run_cluster(
                ds.spins, self.lattice.shape, self.q, [float(beta)], chains, schedule, n_meas,
                ds.extra["e"], ds.extra["counts"], ws=ds.extra["cluster_ws"])
            aux = aux_d[0].cpu().numpy()
        else:
            e_out, acc, counts = k_potts.run(ds.spins, self.lattice.shape, self.q, [float(beta)],
                                             chains, schedule, n_meas, ds.extra["e"],
                                             ds.extra["counts"], ws=ws)
            aux = np.zeros((n_meas, N_AUX), dtype=np.int64)
            aux[:, 0] = acc[0].cpu().numpy()
        e = e_out[0].cpu().numpy().astype(np.int64)
        c = counts[0].cpu().numpy().astype(np.int64)
        running["energy"] = int(e[-1])
        running["counts"] = c[-1].copy()
        return {"E": e, "counts": c, "aux": aux}

    def spec(self):
        return {**super().spec(), "q": self.q}


class ON(_AnyLattice, Model):
    """Classical O(n) model with unit spins (critmc/models.py:225-299); GPU path n = 2, 3.

    On bipartite chain / square / cubic lattices the spins live on the device in
    FP32, colour-split; elsewhere they are FP64 in site order and every update
    runs in the reference's own site orders (``cm_on_table_run``)."""

    kind = "on"
    updates = GPU_UPDATES["on"]

    def __init__(self, lattice: Lattice, n: int):
        super().__init__(lattice)
        if n < 2:
            raise ValueError(f"n must be at least 2, got {n} (n = 1 is the Ising model)")
        if n > 3:
            raise ValueError(f"the GPU O(n) kernels support n = 2 and 3, got {n}")
        self.n = int(n)

    def initial_state(self, start, rng):
        if start == "random":
            s = rng.standard_normal((self.n_sites, self.n))
            s /= np.linalg.norm(s, axis=1, keepdims=True)
            return s
        if start in ("ordered", "up"):
            s = np.zeros((self.n_sites, self.n))
            s[:, 0] = 1.0
            return s
        raise ValueError(f"unknown start {start!r}")

    def energy(self, state):
        b = self.lattice.bonds
        s = np.asarray(state, dtype=np.float64)
        return float(-np.sum(s[b[:, 0]] * s[b[:, 1]]))

    def begin(self, state):
        return {"step": 1.0}

    def to_device(self, state, running):
        if not checkerboard_ok(self.lattice):
            s = torch.from_numpy(np.ascontiguousarray(state, dtype=np.float64)).to(device())
            return DeviceState(s.reshape(1, self.n_sites, self.n))
        s = torch.from_numpy(np.ascontiguousarray(state, dtype=np.float32)).to(device())
        return DeviceState(k_on.to_split(s.reshape(1, self.n_sites, self.n), self.lattice.shape, self.n))

    def from_device(self, ds):
        if not checkerboard_ok(self.lattice):
            return ds.spins.reshape(self.n_sites, self.n).cpu().numpy().copy()
        sites = k_on.to_sites(ds.spins, self.lattice.shape, self.n)
        return sites.reshape(self.n_sites, self.n).cpu().numpy().astype(np.float64)

    def workspace(self):
        if not checkerboard_ok(self.lattice):
            return None
        return k_on.Workspace(self.lattice.shape, 1, self.n)

    def _advance_sequential(self, ds, running, beta, chains, schedule, n_meas):
        dev = ds.spins.device
        nbr = k_table.neighbours(self.lattice, dev)
        if "on_table_ws" not in ds.extra:
            ds.extra["on_table_ws"] = k_table.OnWorkspace(self.n_sites, 1, dev)
        split = checkerboard_ok(self.lattice)
        # on a checkerboard lattice (metropolis-random rows), an FP64 site-order copy for the call
        s64 = (k_on.to_sites(ds.spins, self.lattice.shape, self.n).double() if split else ds.spins)
        e, m, sc, ss, aux = k_table.on_run(s64, nbr, self.n, [float(beta)], [float(running["step"])],
                                           chains, schedule, n_meas, ws=ds.extra["on_table_ws"])
        if split:
            ds.spins.copy_(k_on.to_split(s64.float(), self.lattice.shape, self.n))
        return {"E": e[0].cpu().numpy(), "M": m[0].cpu().numpy(), "stiff_c": sc[0].cpu().numpy(),
                "stiff_s": ss[0].cpu().numpy(), "aux": aux[0].cpu().numpy()}

    def advance(self, ds, running, beta, chains, schedule, n_meas, ws):
        if self._sequential(schedule):
            return self._advance_sequential(ds, running, beta, chains, schedule, n_meas)
        cws = None
        if np.isin(np.asarray(schedule)[:, 0], CLUSTER_CODES).any():
            if "cluster_ws" not in ds.extra:
                ds.extra["cluster_ws"] = k_on.ClusterWorkspace(self.lattice.shape, 1)
            cws = ds.extra["cluster_ws"]
        acc, e, m, sc, ss, aux_d = k_on.run(ds.spins, self.lattice.shape, self.n, [float(beta)],
                                            [float(running["step"])], chains, schedule, n_meas,
                                            ws=ws, cluster_ws=cws, with_aux=True)
        aux = aux_d[0].cpu().numpy()
        return {"E": e[0].cpu().numpy(), "M": m[0].cpu().numpy(), "stiff_c": sc[0].cpu().numpy(),
                "stiff_s": ss[0].cpu().numpy(), "aux": aux}

    def spec(self):
        return {**super().spec(), "n": self.n}


def XY(lattice: Lattice) -> ON:
    """Plane rotator, O(2)."""
    return ON(lattice, 2)


def Heisenberg(lattice: Lattice) -> ON:
    """Classical Heisenberg model, O(3)."""
    return ON(lattice, 3)


# kernel form used by single-chain models: 0 auto, 1 plain (one launch per half sweep), 2 batched
_FORM = [0]


def set_kernel_form(form: int) -> None:
    if form not in (0, 1, 2):
        raise ValueError("form must be 0 (auto), 1 (plain) or 2 (batched)")
    _FORM[0] = int(form)


def build_model(spec: dict[str, Any]) -> Model:
    """Inverse of :meth:`Model.spec` (critmc/models.py:312-322)."""
    lattice = make_lattice(spec["lattice"], int(spec["L"]))
    kind = spec["model"]
    if kind == "ising":
        return Ising(lattice, h=float(spec.get("h", 0.0)))
    if kind == "potts":
        return Potts(lattice, int(spec["q"]))
    if kind == "on":
        return ON(lattice, int(spec["n"]))
    raise ValueError(f"unknown model kind {kind!r}")


def derive(raw: dict[str, np.ndarray], meta: dict[str, Any]) -> dict[str, np.ndarray]:
    """Per-site observables (critmc/models.py:325-351)."""
    n = float(meta["N"])
    kind = meta["model"]
    out: dict[str, np.ndarray] = {"e": np.asarray(raw["E"], dtype=np.float64) / n}
    if kind == "ising":
        m = raw["M"] / n
        out["m"] = m
        out["m2"] = m * m
    elif kind == "potts":
        q = int(meta["q"])
        rho = raw["counts"] / n
        out["m2"] = np.clip((q * np.sum(rho * rho, axis=1) - 1.0) / (q - 1.0), 0.0, None)
        out["mmax"] = (q * rho.max(axis=1) - 1.0) / (q - 1.0)
    elif kind == "on":
        mv = raw["M"] / n
        out["m2"] = np.sum(mv * mv, axis=1)
    else:
        raise ValueError(f"unknown model kind {kind!r}")
    out["absm"] = np.sqrt(out["m2"])
    out["m4"] = out["m2"] ** 2
    return out

