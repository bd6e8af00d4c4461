"""Per-chain random-stream state shared by every kernel family.

A chain is identified by its Philox key ``(k0, k1)`` and a stream id, and
advanced by its sweep counter (the number of local sweeps it has performed).
The counters live on the device, where the kernels advance them, and are
mirrored on the host.
"""

from __future__ import annotations

import numpy as np
import torch

from ..device import device, from_numpy_u32, from_numpy_u64, to_numpy_u64


class Chains:
    def __init__(self, keys, stream_ids=None, sweeps=None, dev: torch.device | None = None):
        keys = np.asarray(keys, dtype=np.uint64).reshape(-1, 2)
        if np.any(keys >= (1 << 32)):
            raise ValueError("Philox key words must fit in 32 bits")
        self.R = keys.shape[0]
        self.keys_h = keys.astype(np.uint32)
        sids = np.zeros(self.R, np.uint32) if stream_ids is None else np.asarray(stream_ids)
        self.sids_h = np.ascontiguousarray(sids, dtype=np.uint32).reshape(self.R)
        sw = np.zeros(self.R, np.uint64) if sweeps is None else np.asarray(sweeps)
        self.sweeps_h = np.ascontiguousarray(sw, dtype=np.uint64).reshape(self.R).copy()
        self.dev = dev or device()
        self.keys_d = from_numpy_u32(self.keys_h, self.dev)
        self.sids_d = from_numpy_u32(self.sids_h, self.dev)
        self.sweeps_d = from_numpy_u64(self.sweeps_h, self.dev)

    def advance(self, n: int) -> None:
        """Record on the host that every chain performed ``n`` more sweeps."""
        self.sweeps_h += np.uint64(n)

    def sync(self) -> None:
        """Refresh the host mirror from the device counters."""
        self.sweeps_h = to_numpy_u64(self.sweeps_d).copy()

    def subset(self, idx) -> Chains:
        idx = np.asarray(idx)
        return Chains(self.keys_h[idx], self.sids_h[idx], self.sweeps_h[idx], self.dev)
