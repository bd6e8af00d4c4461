"""Update codes and diagnostics columns (critmc/kernels/common.py:16-40).

The schedule vocabulary and the ``aux`` column layout are the reference's, so
schedules and result files mean the same thing in both code bases.
"""

from __future__ import annotations

METROPOLIS = 0
METROPOLIS_RANDOM = 1
HEATBATH = 2
WOLFF = 3
SWENDSEN_WANG = 4
OVERRELAX = 5

UPDATE_CODES = {
    "metropolis": METROPOLIS,
    "metropolis-random": METROPOLIS_RANDOM,
    "heatbath": HEATBATH,
    "wolff": WOLFF,
    "sw": SWENDSEN_WANG,
    "swendsen-wang": SWENDSEN_WANG,
    "overrelax": OVERRELAX,
}

CLUSTER_CODES = (WOLFF, SWENDSEN_WANG)

AUX_ACCEPTED = 0
AUX_WOLFF_SIZE = 1
AUX_SW_C2 = 2
AUX_SW_NCLUSTERS = 3
N_AUX = 4

# codes the GPU kernels implement, per model kind
GPU_UPDATES = {
    "ising": ("metropolis", "metropolis-random", "heatbath", "wolff", "sw", "swendsen-wang"),
    "potts": ("metropolis", "metropolis-random", "heatbath", "wolff", "sw", "swendsen-wang"),
    "on": ("metropolis", "metropolis-random", "heatbath", "overrelax", "wolff", "sw",
           "swendsen-wang"),
}
