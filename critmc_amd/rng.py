"""Reproducible streams keyed by ``(seed, *keys)``.

Same semantics as critmc/rng.py:28-71: a stream is a root seed plus a tuple of
keys, mapped through ``numpy.random.SeedSequence`` spawn keys, so any chain of
a batch can be reproduced in isolation.  Host-side draws (initial states,
disorder) use ``PCG64DXSM`` exactly as the reference does; the GPU sweeps use a
counter-based Philox4x32-10 stream whose 64-bit key is taken from the same
generator (:func:`chain_key`).
"""

from __future__ import annotations

import zlib
from collections.abc import Hashable

import numpy as np

__all__ = ["seed_sequence", "make_rng", "spawn_rngs", "chain_key"]


def _key_to_int(key: Hashable) -> int:
    if isinstance(key, (bool, np.bool_)):
        return int(key)
    if isinstance(key, (int, np.integer)):
        if key < 0:
            raise ValueError("integer stream keys must be non-negative")
        return int(key)
    text = key if isinstance(key, str) else repr(key)
    return zlib.crc32(text.encode("utf-8"))


def seed_sequence(seed: int | None, *keys: Hashable) -> np.random.SeedSequence:
    return np.random.SeedSequence(entropy=seed, spawn_key=tuple(_key_to_int(k) for k in keys))


def make_rng(seed=None, *keys: Hashable) -> np.random.Generator:
    if isinstance(seed, np.random.Generator):
        if keys:
            raise ValueError("stream keys cannot be applied to an existing Generator")
        return seed
    if isinstance(seed, np.random.SeedSequence):
        ss = seed
        if keys:
            ss = np.random.SeedSequence(
                entropy=ss.entropy, spawn_key=ss.spawn_key + tuple(_key_to_int(k) for k in keys)
            )
    else:
        ss = seed_sequence(seed, *keys)
    return np.random.Generator(np.random.PCG64DXSM(ss))


def spawn_rngs(seed: int | None, n: int, *keys: Hashable) -> list[np.random.Generator]:
    return [make_rng(seed, *keys, i) for i in range(n)]


def chain_key(rng: np.random.Generator) -> tuple[int, int]:
    """Philox key of a chain: two 32-bit words drawn from its host generator."""
    k = rng.integers(0, 1 << 32, size=2, dtype=np.uint64)
    return int(k[0]), int(k[1])
