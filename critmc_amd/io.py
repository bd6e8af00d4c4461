"""Result files: arrays in ``.npz`` plus a JSON ``__meta__`` entry.

The on-disk format is the reference's (critmc/io.py:13-48): every array under
its own name, compressed, and one extra 0-d text entry ``__meta__`` holding
the metadata as JSON.  Files written here therefore load in the reference's
analysis stack and vice versa.
"""

from __future__ import annotations

import json
import math
from functools import singledispatch
from pathlib import Path, PurePath
from typing import Any

import numpy as np

__all__ = ["to_jsonable", "save_npz", "load_npz", "save_json", "load_json"]

META = "__meta__"


@singledispatch
def to_jsonable(obj: Any) -> Any:
    """A JSON-ready copy of ``obj``.

    Mappings get string keys, lists and tuples become lists, numpy arrays and
    scalars become plain Python values, paths become strings, and a
    non-finite Python float becomes its text (``"nan"``, ``"inf"``).
    Anything else is returned as it is.
    """
    return obj


@to_jsonable.register(dict)
def _(obj: dict) -> dict:
    return {str(key): to_jsonable(value) for key, value in obj.items()}


@to_jsonable.register(list)
@to_jsonable.register(tuple)
def _(obj) -> list:
    return [to_jsonable(value) for value in obj]


@to_jsonable.register(np.ndarray)
def _(obj: np.ndarray) -> Any:
    return obj.tolist()


@to_jsonable.register(np.generic)
def _(obj: np.generic) -> Any:
    return obj.item()


@to_jsonable.register(PurePath)
def _(obj: PurePath) -> str:
    return str(obj)


@to_jsonable.register(float)
def _(obj: float) -> Any:
    return obj if math.isfinite(obj) else repr(obj)


def _with_parent(path) -> Path:
    target = Path(path)
    target.parent.mkdir(parents=True, exist_ok=True)
    return target


def save_npz(path, arrays: dict[str, np.ndarray], meta: dict[str, Any]) -> Path:
    """Write ``arrays`` and ``meta`` to one compressed ``.npz`` file."""
    if META in arrays:
        raise ValueError(f"array name {META!r} is reserved")
    entries = {name: np.asarray(values) for name, values in arrays.items()}
    entries[META] = np.array(json.dumps(to_jsonable(meta)))
    target = _with_parent(path)
    np.savez_compressed(target, **entries)
    return target


def load_npz(path) -> tuple[dict[str, np.ndarray], dict[str, Any]]:
    """Arrays and metadata of a file written by :func:`save_npz` (no pickles)."""
    with np.load(path, allow_pickle=False) as archive:
        names = list(archive.files)
        meta = json.loads(str(archive[META])) if META in names else {}
        arrays = {name: archive[name] for name in names if name != META}
    return arrays, meta


def save_json(path, obj: Any) -> Path:
    """``obj`` as indented UTF-8 JSON with a final newline."""
    target = _with_parent(path)
    target.write_text(json.dumps(to_jsonable(obj), indent=2) + "\n", encoding="utf-8")
    return target


def load_json(path) -> Any:
    return json.loads(Path(path).read_text(encoding="utf-8"))
