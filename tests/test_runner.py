"""Runner cache semantics (critmc tests/test_runner.py:10-60), host side only."""

import numpy as np
import pytest

from critmc_amd.runner import StaleCacheError, Task, run_tasks
from critmc_amd.sim import TimeSeries

SPEC = {"model": "ising", "lattice": "square", "L": 8}


def _cached(task, out_dir, value=1.0):
    from dataclasses import asdict

    ts = TimeSeries({"E": np.full(4, value), "M": np.zeros(4)},
                    {"model": "ising", "N": 64, "task": asdict(task)})
    ts.save(out_dir / task.filename)


def test_filename_and_reuse(tmp_path):
    t = Task(SPEC, 0.4, "metropolis", 10, 4, seed=3, keys=("a", 1))
    assert t.filename == "ising_L8_b0.40000000_a_1.npz"
    assert Task(SPEC, 0.4, "metropolis", 10, 4, seed=3, name="x").filename == "x.npz"
    _cached(t, tmp_path, 7.0)
    (tmp_path / "left.part.npz").write_bytes(b"")
    out = run_tasks([t], out_dir=tmp_path, progress=False)  # served from the cache, no GPU needed
    assert np.all(out[0]["E"] == 7.0)
    assert not (tmp_path / "left.part.npz").exists()


@pytest.mark.parametrize("field,value", [("seed", 4), ("beta", 0.41), ("n_meas", 5), ("keys", ("a", 2))])
def test_stale_cache_is_refused(tmp_path, field, value):
    t = Task(SPEC, 0.4, "metropolis", 10, 4, seed=3, keys=("a", 1), name="same")
    _cached(t, tmp_path)
    kw = dict(model=SPEC, beta=0.4, updates="metropolis", n_therm=10, n_meas=4, seed=3,
              keys=("a", 1), name="same")
    kw[field] = value
    with pytest.raises(StaleCacheError):
        run_tasks([Task(**kw)], out_dir=tmp_path, progress=False)


def test_file_without_task_record_is_refused(tmp_path):
    t = Task(SPEC, 0.4, "metropolis", 10, 4, seed=3, name="bare")
    TimeSeries({"E": np.zeros(2)}, {"model": "ising", "N": 64}).save(tmp_path / "bare.npz")
    with pytest.raises(StaleCacheError, match="does not record"):
        run_tasks([t], out_dir=tmp_path, progress=False)
