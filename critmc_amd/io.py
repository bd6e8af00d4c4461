"""Result files: arrays in ``.npz`` plus a JSON ``__meta__`` entry.

The format is the reference's (critmc/io.py:13-48) so that files written here
load in the reference's analysis stack and vice versa.
"""

from __future__ import annotations

import json
from pathlib import Path
from typing import Any

import numpy as np

__all__ = ["to_jsonable", "save_npz", "load_npz", "save_json", "load_json"]

META = "__meta__"


def to_jsonable(obj: Any) -> Any:
    if isinstance(obj, dict):
        return {str(k): to_jsonable(v) for k, v in obj.items()}
    if isinstance(obj, (list, tuple)):
        return [to_jsonable(v) for v in obj]
    if isinstance(obj, np.ndarray):
        return obj.tolist()
    if isinstance(obj, np.generic):
        return obj.item()
    if isinstance(obj, Path):
        return str(obj)
    if isinstance(obj, float) and not np.isfinite(obj):
        return str(obj)
    return obj


def save_npz(path, arrays: dict[str, np.ndarray], meta: dict[str, Any]) -> Path:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    if META in arrays:
        raise ValueError(f"array name {META!r} is reserved")
    payload = {k: np.asarray(v) for k, v in arrays.items()}
    payload[META] = np.array(json.dumps(to_jsonable(meta)))
    np.savez_compressed(path, **payload)
    return path


def load_npz(path) -> tuple[dict[str, np.ndarray], dict[str, Any]]:
    with np.load(path, allow_pickle=False) as data:
        arrays = {k: data[k] for k in data.files if k != META}
        meta = json.loads(str(data[META])) if META in data.files else {}
    return arrays, meta


def save_json(path, obj: Any) -> Path:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(to_jsonable(obj), indent=2) + "\n", encoding="utf-8")
    return path


def load_json(path) -> Any:
    return json.loads(Path(path).read_text(encoding="utf-8"))
