"""Runner on the GPU: batched groups equal single runs bit for bit, order and caching."""

import numpy as np
import pytest

from critmc_amd.runner import Task, run_task, run_tasks

pytestmark = pytest.mark.gpu


def test_batched_group_equals_single_runs(tmp_path):
    spec = {"model": "ising", "lattice": "square", "L": 32}
    tasks = [Task(spec, b, "metropolis+heatbath", 50, 200, seed=5, keys=("b", k))
             for k, b in enumerate([0.38, 0.42, 0.44, 0.47])]
    tasks.append(Task({"model": "potts", "lattice": "square", "L": 8, "q": 3}, 0.9, "heatbath",
                      20, 100, seed=2))
    batched = run_tasks(tasks, out_dir=tmp_path, progress=False)
    assert batched[0].meta["batched_with"] == 4
    for t, ts in zip(tasks, batched):
        single = run_task(t)
        for k in ("E", "M", "counts"):
            if k in single.raw:
                assert np.array_equal(ts[k], single[k]), (t, k)
    again = run_tasks(tasks, out_dir=tmp_path, progress=False)  # from the cache
    for a, b in zip(batched, again):
        assert np.array_equal(a["E"], b["E"])
