"""The launch-shape autotuner: it picks one of its candidates, caches the choice,
and tuning cannot change results, because every candidate runs the same chains
bit for bit."""

import numpy as np
import pytest
import torch

from critmc_amd import tune
from critmc_amd.kernels import ising as KI
from critmc_amd.kernels import potts as KP
from critmc_amd.kernels.chains import Chains

pytestmark = pytest.mark.gpu


def test_ising_candidates_cover_forms_and_layouts():
    c64 = tune.ising_candidates(64)
    assert {(c["form"], c["layout"]) for c in c64} == {("batched", "int8"), ("batched", "packed"),
                                                      ("wide", "int8"), ("wide", "packed")}
    assert all(c["form"] == "batched" for c in tune.ising_candidates(32))
    assert {c["form"] for c in tune.ising_candidates(192)} == {"batched", "wide"}
    assert tune.ising_candidates(96) == []  # neither form tiles L0 = 96


def test_every_ising_candidate_runs_the_same_chains():
    dev = torch.device("cuda", 0)
    L, R = 64, 3
    s0 = torch.empty((R, L * L), dtype=torch.int8, device=dev)
    KI.init_spins(s0, (L, L), Chains(np.arange(2 * R, dtype=np.uint64).reshape(R, 2) + 5,
                                     np.arange(R), None, dev))
    betas = [0.40, 0.44, 0.48]
    ref = None
    for c in tune.ising_candidates(L):
        final, series = tune.run_ising_candidate(c, s0, L, betas, 6)
        got = (final.cpu().numpy(), [t.cpu().numpy() for t in series])
        if ref is None:
            ref = got
            continue
        assert np.array_equal(got[0], ref[0]), c
        for a, b in zip(got[1], ref[1]):
            assert np.array_equal(a, b), c


def test_every_potts_candidate_runs_the_same_chains():
    dev = torch.device("cuda", 0)
    L, R, q = 64, 3, 10
    s0 = torch.empty((R, L * L), dtype=torch.int8, device=dev)
    KP.init_states(s0, (L, L), q, Chains(np.arange(2 * R, dtype=np.uint64).reshape(R, 2) + 9,
                                         np.arange(R), None, dev), "random")
    for code in (0, 2):
        ref = None
        for c in tune.potts_candidates(L):
            final, series = tune.run_potts_candidate(c, s0, L, q, [1.3, 1.4, 1.5], code, 5)
            got = (final.cpu().numpy(), [t.cpu().numpy() for t in series])
            if ref is None:
                ref = got
                continue
            assert np.array_equal(got[0], ref[0]), (code, c)
            for a, b in zip(got[1], ref[1]):
                assert np.array_equal(a, b), (code, c)


def test_tuners_pick_a_candidate_and_cache_it(tmp_path):
    cache = tmp_path / "tune.json"
    best = tune.tune_ising(64, 16, sweeps=4, cache_file=cache)
    keys = {(c["form"], c["block"], c["layout"]) for c in best["candidates"]}
    assert (best["form"], best["block"], best["layout"]) in keys
    assert best["rate"] == max(c["rate"] for c in best["candidates"]) > 0
    assert tune.tune_ising(64, 16, sweeps=4, cache_file=cache) is best
    tune._CACHE.clear()
    assert tune.tune_ising(64, 16, sweeps=4, cache_file=cache)["rate"] == best["rate"]  # from disk
    pb = tune.tune_potts(64, 16, 10, sweeps=3)
    assert (pb["form"], pb["block"]) in {(c["form"], c["block"]) for c in pb["candidates"]}
    assert pb["rate"] > 0
