"""Periodic lattices with the reference's layout contract.

Interface of critmc/lattice.py:27-221: a ``Lattice`` carries a neighbour table
``nbr[N, z]`` whose columns come in opposite pairs (``2a`` = +e_a,
``2a+1`` = -e_a, lattice.py:129-134), a forward bond list ``bonds`` built
axis-major with ``bond_axis`` (:136-138), and a two-colouring ``sublattice``
when every extent is even (:140-141).  Sites are numbered with the first
coordinate running fastest, ``i = x0 + L0 (x1 + L1 x2 ...)`` (:118-128).

The checkerboard kernels derive neighbours from ``shape``; lattices they cannot
sweep (odd extents, triangular) run on the neighbour-table kernels, which read
``nbr`` from the device.
"""

from __future__ import annotations

from dataclasses import dataclass
from math import prod

import numpy as np

__all__ = ["Lattice", "hypercubic", "chain", "square", "cubic", "triangular", "make_lattice",
           "checkerboard_ok", "require_checkerboard"]


@dataclass(frozen=True, eq=False)
class Lattice:
    name: str
    shape: tuple[int, ...]
    nbr: np.ndarray
    bonds: np.ndarray
    bond_axis: np.ndarray
    positions: np.ndarray
    coords: np.ndarray
    primitive: np.ndarray
    sublattice: np.ndarray | None

    @property
    def n_sites(self) -> int:
        return int(self.nbr.shape[0])

    @property
    def z(self) -> int:
        return int(self.nbr.shape[1])

    @property
    def dim(self) -> int:
        return len(self.shape)

    @property
    def n_bonds(self) -> int:
        return int(self.bonds.shape[0])

    @property
    def L(self) -> int:
        return int(self.shape[0])

    @property
    def bipartite(self) -> bool:
        return self.sublattice is not None

    @property
    def hypercubic(self) -> bool:
        return self.name in ("chain", "square", "cubic") or self.name.startswith("hypercubic")

    def site(self, *x: int) -> int:
        if len(x) != self.dim:
            raise ValueError(f"expected {self.dim} coordinates, got {len(x)}")
        idx = 0
        for c, n in reversed(list(zip(x, self.shape))):
            idx = idx * n + (int(c) % n)
        return idx

    def reciprocal(self) -> np.ndarray:
        return 2.0 * np.pi * np.linalg.inv(self.primitive).T

    def kmin(self, axis: int = 0) -> np.ndarray:
        return self.reciprocal()[axis] / self.shape[axis]

    def __repr__(self) -> str:
        return f"Lattice({self.name}, {'x'.join(map(str, self.shape))}, N={self.n_sites}, z={self.z})"


def _extents(shape) -> tuple[int, ...]:
    shape = tuple(int(n) for n in shape)
    if not shape:
        raise ValueError("lattice needs at least one dimension")
    if any(n < 2 for n in shape):
        raise ValueError(f"every linear extent must be >= 2, got {shape}")
    return shape


def _coords(shape: tuple[int, ...]) -> np.ndarray:
    n = prod(shape)
    idx = np.arange(n, dtype=np.int64)
    out = np.empty((n, len(shape)), dtype=np.int64)
    for a, ext in enumerate(shape):
        out[:, a] = idx % ext
        idx = idx // ext
    return out


def _index(coords: np.ndarray, shape: tuple[int, ...]) -> np.ndarray:
    idx = np.zeros(coords.shape[0], dtype=np.int64)
    for a in reversed(range(len(shape))):
        idx = idx * shape[a] + coords[:, a]
    return idx


def hypercubic(shape, name: str | None = None) -> Lattice:
    """Periodic hypercubic lattice (critmc/lattice.py:118-155)."""
    shape = _extents(shape)
    d = len(shape)
    xyz = _coords(shape)
    n = xyz.shape[0]
    nbr = np.empty((n, 2 * d), dtype=np.int64)
    for a in range(d):
        for col, step in ((2 * a, 1), (2 * a + 1, -1)):
            moved = xyz.copy()
            moved[:, a] = (moved[:, a] + step) % shape[a]
            nbr[:, col] = _index(moved, shape)
    sites = np.arange(n, dtype=np.int64)
    bonds = np.concatenate([np.column_stack([sites, nbr[:, 2 * a]]) for a in range(d)])
    bond_axis = np.repeat(np.arange(d, dtype=np.int64), n)
    even = all(ext % 2 == 0 for ext in shape)
    sub = (xyz.sum(axis=1) % 2).astype(np.int8) if even else None
    if name is None:
        name = {1: "chain", 2: "square", 3: "cubic"}.get(d, f"hypercubic{d}d")
    return Lattice(name, shape, nbr, bonds, bond_axis, xyz.astype(np.float64), xyz,
                   np.eye(d), sub)


def chain(L: int) -> Lattice:
    return hypercubic((L,), name="chain")


def square(L: int, M: int | None = None) -> Lattice:
    return hypercubic((L, L if M is None else M), name="square")


def cubic(L: int) -> Lattice:
    return hypercubic((L, L, L), name="cubic")


def triangular(L: int, M: int | None = None) -> Lattice:
    """Triangular torus (critmc/lattice.py:173-213); not bipartite, so no checkerboard sweeps."""
    shape = _extents((L, L if M is None else M))
    xyz = _coords(shape)
    n = xyz.shape[0]
    nbr = np.empty((n, 6), dtype=np.int64)
    for a, (dx, dy) in enumerate(((1, 0), (0, 1), (1, -1))):
        for col, sgn in ((2 * a, 1), (2 * a + 1, -1)):
            moved = np.column_stack([(xyz[:, 0] + sgn * dx) % shape[0],
                                     (xyz[:, 1] + sgn * dy) % shape[1]])
            nbr[:, col] = _index(moved, shape)
    sites = np.arange(n, dtype=np.int64)
    bonds = np.concatenate([np.column_stack([sites, nbr[:, 2 * a]]) for a in range(3)])
    bond_axis = np.repeat(np.arange(3, dtype=np.int64), n)
    prim = np.array([[1.0, 0.0], [0.5, np.sqrt(3.0) / 2.0]])
    return Lattice("triangular", shape, nbr, bonds, bond_axis, xyz @ prim, xyz, prim, None)


_BUILDERS = {"chain": chain, "square": square, "cubic": cubic, "triangular": triangular}


def make_lattice(name: str, L: int) -> Lattice:
    try:
        return _BUILDERS[name](L)
    except KeyError:
        raise ValueError(f"unknown lattice {name!r}; choose from {sorted(_BUILDERS)}") from None


def checkerboard_ok(lattice: Lattice) -> bool:
    """Whether the checkerboard kernels can sweep this lattice: hypercubic, d <= 3,
    every extent even."""
    return lattice.hypercubic and lattice.dim <= 3 and lattice.bipartite


def require_checkerboard(lattice: Lattice) -> None:
    """Checkerboard sweeps need a bipartite hypercubic lattice of dimension <= 3."""
    if not lattice.hypercubic or lattice.dim > 3:
        raise ValueError(f"checkerboard GPU sweeps support chain/square/cubic lattices, not {lattice!r}")
    if not lattice.bipartite:
        raise ValueError(
            f"checkerboard GPU sweeps need every extent even (bipartite lattice); got {lattice!r}"
        )
