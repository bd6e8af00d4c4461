"""Per-replica measurement caps (continuous batching) against the CPU replay.

A launch of n_meas measurements with caps c[r] must leave replica r exactly
where c[r] measurements alone would: same state, same first c[r] series
entries, sweep counter advanced by c[r] schedules, running totals updated.
Replicas with cap 0 are untouched.
"""

import numpy as np
import pytest
import torch

import oracle
from critmc_amd.kernels import ising as KI
from critmc_amd.kernels import potts as KP
from critmc_amd.kernels.chains import Chains

pytestmark = pytest.mark.gpu


def _chains(R, seed, dev):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 1 << 32, size=(R, 2), dtype=np.uint64)
    sids = rng.integers(0, 1 << 20, size=R, dtype=np.uint64)
    sweeps = rng.integers(0, 1 << 36, size=R, dtype=np.uint64)
    return Chains(keys, sids, sweeps, dev), keys, sids, sweeps


CAPS = [3, 0, 6, 1, 5]


@pytest.mark.parametrize("layout,form", [("int8", KI.FORM_BATCHED), ("packed", KI.FORM_BATCHED),
                                         ("int8", KI.FORM_WIDE), ("packed", KI.FORM_WIDE)])
def test_ising_caps_match_replay(layout, form):
    dev = torch.device("cuda", 0)
    shape, R, n_meas = (64, 64), len(CAPS), max(CAPS)
    chains, keys, sids, sw0 = _chains(R, 11, dev)
    n = shape[0] * shape[1]
    s = torch.empty((R, n), dtype=torch.int8, device=dev)
    KI.init_spins(s, shape, chains, "random")
    s0 = s.cpu().numpy()
    e, m = KI.measure(s, shape, R)
    betas = np.linspace(0.3, 0.6, R)
    sched = np.array([[0, 1], [2, 2]])
    caps = torch.tensor(CAPS, dtype=torch.int64, device=dev)
    lay = KI.LAYOUT_PACKED if layout == "packed" else KI.LAYOUT_INT8
    st = KI.pack(s, shape, R) if lay == KI.LAYOUT_PACKED else s
    e_out, m_out, acc = KI.run(st, shape, betas, 0.0, chains, sched, n_meas, e, m, layout=lay,
                               form=form, n_meas_rep=caps)
    if lay == KI.LAYOUT_PACKED:
        s = KI.unpack(st, shape, R)
    got = s.cpu().numpy()
    e_out, m_out, acc = (t.cpu().numpy() for t in (e_out, m_out, acc))
    sweeps = chains.sweeps_d.cpu().numpy().view(np.uint64)
    for r, c in enumerate(CAPS):
        rep = oracle.IsingReplay(s0[r], shape, betas[r], 0.0, keys[r], int(sids[r]), int(sw0[r]))
        ref = rep.run(sched, c) if c else None
        assert np.array_equal(got[r], rep.s), f"replica {r}"
        if c:
            assert np.array_equal(e_out[r, :c], ref["Eb"])
            assert np.array_equal(m_out[r, :c], ref["M"])
            assert np.array_equal(acc[r, :c], ref["accepted"])
            assert int(e[r]) == int(ref["Eb"][-1]) and int(m[r]) == int(ref["M"][-1])
        assert int(sweeps[r]) == int(sw0[r]) + 3 * c


@pytest.mark.parametrize("shape,grid", [((32, 32), 2), ((128, 128), -1), ((4, 4, 6), -1)],
                         ids=["per-replica", "wide-2d", "wide-3d"])
def test_potts_caps_match_replay(shape, grid):
    dev = torch.device("cuda", 0)
    q, R, n_meas = 5, len(CAPS), max(CAPS)
    chains, keys, sids, sw0 = _chains(R, 12, dev)
    n = int(np.prod(shape))
    s = torch.empty((R, n), dtype=torch.int8, device=dev)
    KP.init_states(s, shape, q, chains, "random")
    s0 = s.cpu().numpy()
    e, cnt = KP.measure(s, shape, q, R)
    betas = np.linspace(0.8, 1.4, R)
    sched = np.array([[0, 1], [2, 1]])
    caps = torch.tensor(CAPS, dtype=torch.int64, device=dev)
    e_out, acc, counts = KP.run(s, shape, q, betas, chains, sched, n_meas, e, cnt, grid=grid,
                                n_meas_rep=caps)
    got = s.cpu().numpy()
    e_out, acc, counts = (t.cpu().numpy() for t in (e_out, acc, counts))
    sweeps = chains.sweeps_d.cpu().numpy().view(np.uint64)
    e_h, cnt_h = e.cpu().numpy(), cnt.cpu().numpy()
    for r, c in enumerate(CAPS):
        rep = oracle.PottsReplay(s0[r], shape, q, betas[r], keys[r], int(sids[r]), int(sw0[r]))
        ref = rep.run(sched, c) if c else None
        assert np.array_equal(got[r], rep.s), f"replica {r}"
        if c:
            assert np.array_equal(e_out[r, :c], ref["E"])
            assert np.array_equal(acc[r, :c], ref["accepted"])
            assert np.array_equal(counts[r, :c], ref["counts"])
            assert int(e_h[r]) == int(ref["E"][-1])
            assert np.array_equal(cnt_h[r], ref["counts"][-1])
        assert int(sweeps[r]) == int(sw0[r]) + 2 * c


@pytest.mark.parametrize("shape", [(64, 48), (512, 40)], ids=["one-tile", "four-tiles"])
@pytest.mark.parametrize("sched", [[[0, 1], [5, 2]], [[0, 1]], [[5, 1], [0, 1]]],
                         ids=["metropolis+2or", "metropolis", "or+metropolis"])
def test_on_caps_match_alone(shape, sched):
    """XY replicas with caps give what each replica gives alone with n_meas = its cap:
    the same FP32 spins, acceptance counts and FP64 measurements, bit for bit (the
    measurement partial sums are formed per replica in both cases)."""
    from critmc_amd.kernels import on as KO

    dev = torch.device("cuda", 0)
    R, n_meas = len(CAPS), max(CAPS)
    chains, keys, sids, sw0 = _chains(R, 13, dev)
    N = int(np.prod(shape))
    s = torch.empty((R, N, 2), dtype=torch.float32, device=dev)
    KO.init_spins(s, shape, 2, chains)
    s0 = s.clone()
    betas = np.linspace(0.8, 1.3, R)
    steps = np.linspace(0.5, 1.5, R)
    sched = np.array(sched)
    caps = torch.tensor(CAPS, dtype=torch.int64, device=dev)
    out = KO.run(s, shape, 2, betas, steps, chains, sched, n_meas, n_meas_rep=caps)
    out = [t.cpu().numpy() for t in out]
    sweeps = chains.sweeps_d.cpu().numpy().view(np.uint64)
    per_meas = int(sched[:, 1].sum())
    for r, c in enumerate(CAPS):
        assert int(sweeps[r]) == int(sw0[r]) + per_meas * c
        if c == 0:
            assert torch.equal(s[r], s0[r]), f"replica {r} moved"
            continue
        one = Chains(keys[r:r + 1], sids[r:r + 1], sw0[r:r + 1], dev)
        s1 = s0[r:r + 1].clone()
        alone = KO.run(s1, shape, 2, betas[r:r + 1], steps[r:r + 1], one, sched, c)
        assert torch.equal(s[r], s1[0]), f"replica {r}"
        for got, want in zip(out, alone):
            assert np.array_equal(got[r, :c], want.cpu().numpy()[0]), f"replica {r}"
