"""critmc_amd -- MI355X-native lattice-update engine with critmc's API.

Hand-written HIP kernels (gfx950) advance Ising, Potts, Edwards-Anderson and
O(n) chains with checkerboard sweeps and cluster updates, and solve the MBAR
reweighting equations on their energy series; the Python layer mirrors
critmc's ``Simulation`` / ``ParallelTempering`` / runner / reweighting
interfaces and reaches the kernels through the C-ABI library
``libcritmc_hip.so`` via ctypes.
"""

from .flathist import (EnergyLevels, Multicanonical, MulticanonicalBatch, WangLandau,
                       WangLandauBatch)
from .lattice import Lattice, chain, cubic, hypercubic, make_lattice, square, triangular
from .models import ON, DeviceState, Heisenberg, Ising, Model, Potts, XY, build_model, derive
from .reweight import MultiHistogram, Reweighter, canonical_observables, kish_ess
from .rng import chain_key, make_rng, seed_sequence, spawn_rngs
from .sim import Simulation, TimeSeries, parse_updates

__version__ = "0.1.0"

__all__ = [
    "Lattice", "chain", "cubic", "hypercubic", "make_lattice", "square", "triangular",
    "DeviceState", "Ising", "Model", "Potts", "ON", "XY", "Heisenberg", "build_model", "derive",
    "chain_key", "make_rng", "seed_sequence", "spawn_rngs",
    "Simulation", "TimeSeries", "parse_updates",
    "MultiHistogram", "Reweighter", "canonical_observables", "kish_ess",
    "EnergyLevels", "WangLandau", "WangLandauBatch", "Multicanonical", "MulticanonicalBatch",
]
