"""GPU parity of the checkerboard Ising kernels against the CPU replay.

Bit-exact: spins, E_b, M and accepted-count series of every chain, for the
plain (one launch per half sweep) and batched (LDS-resident) forms, int8 and
bit-packed layouts, 1D/2D/3D lattices, both update rules, with and without a
field.  Then size-independent properties at full size and statistical pulls
against exact results through the reference's Simulation API.
"""

import numpy as np
import pytest
import torch

import oracle
from critmc_amd.kernels import ising as K
from critmc_amd.kernels.chains import Chains

pytestmark = pytest.mark.gpu


def _dev():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch.device("cuda", 0)


def _chains(R, seed, dev):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 1 << 32, size=(R, 2), dtype=np.uint64)
    sids = rng.integers(0, 1 << 20, size=R, dtype=np.uint64)
    sweeps = rng.integers(0, 1 << 36, size=R, dtype=np.uint64) if seed % 2 else np.zeros(R, np.uint64)
    return Chains(keys, sids, sweeps, dev), keys, sids, sweeps


def test_device_philox_matches_oracle():
    from critmc_amd._lib import lib
    from critmc_amd.device import ptr, stream

    dev = _dev()
    out = torch.empty(4 * 1000, dtype=torch.int32, device=dev)
    rc = lib().cm_philox_fill(ptr(out), 1000, 0xA4093822, 0x299F31D0, 123, 0x85A308D3, 0x13198A2E,
                              0x03707344, stream())
    assert rc == 0
    got = out.cpu().numpy().view(np.uint32).reshape(-1, 4)
    for b in (0, 1, 517, 999):
        want = oracle.philox([123 + b, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])
        assert np.array_equal(got[b], want)


@pytest.mark.parametrize("shape", [(64, 64), (6, 10), (4, 4, 4), (10,)])
def test_device_init_and_measure_match_oracle(shape):
    dev = _dev()
    R = 3
    chains, keys, sids, _ = _chains(R, 4, dev)
    n = int(np.prod(shape))
    s = torch.empty((R, n), dtype=torch.int8, device=dev)
    K.init_spins(s, shape, chains, "random")
    e, m = K.measure(s, shape, R)
    host = s.cpu().numpy()
    for r in range(R):
        want = oracle.ising_init(n, keys[r], int(sids[r]))
        assert np.array_equal(host[r], want)
        assert (int(e[r]), int(m[r])) == oracle.ising_energy(want, shape)


def _replay(s0, shape, beta, h, key, sid, sweep, sched, n_meas):
    r = oracle.IsingReplay(s0, shape, beta, h, key, sid, sweep)
    out = r.run(sched, n_meas)
    return r.s, out


CASES = [
    # shape, form, layout
    ((4, 4), K.FORM_BATCHED, 0),
    ((16, 16), K.FORM_BATCHED, 0),
    ((64, 64), K.FORM_BATCHED, 0),
    ((128, 32), K.FORM_BATCHED, 0),
    ((256, 256), K.FORM_BATCHED, 0),
    ((512, 34), K.FORM_BATCHED, 0),   # row-parity split, a ragged last stride
    ((128, 128), K.FORM_BATCHED, 1),  # row-parity split, two words per row
    ((64, 256), K.FORM_BATCHED, 0),   # row-parity split, one word per row
    ((384, 64), K.FORM_BATCHED, 0),   # six words per row: no split
    ((64, 64), K.FORM_BATCHED, 1),
    ((256, 16), K.FORM_BATCHED, 1),
    ((6, 10), K.FORM_BATCHED, 0),
    ((4, 4), K.FORM_PLAIN, 0),
    ((64, 64), K.FORM_PLAIN, 0),
    ((100, 6), K.FORM_PLAIN, 0),
    ((8,), K.FORM_PLAIN, 0),
    ((4, 6, 8), K.FORM_PLAIN, 0),
    ((34, 2, 2), K.FORM_PLAIN, 0),
    # the wide form off 2D rows of 64-site segments: the plain form's words,
    # every replica in each launch (auto takes it on these int8 lattices)
    ((8,), K.FORM_WIDE, 0),
    ((100, 6), K.FORM_WIDE, 0),
    ((6, 10), K.FORM_WIDE, 0),
    ((4, 6, 8), K.FORM_WIDE, 0),
    ((34, 2, 2), K.FORM_WIDE, 0),
    ((16, 16, 16), K.FORM_AUTO, 0),
    ((64, 4, 6), K.FORM_WIDE, 0),     # cubic, 64-site row segments: whole segments per lane
    ((128, 6, 4), K.FORM_AUTO, 0),    # two segments per row
    ((64, 64), K.FORM_WIDE, 0),
    ((128, 32), K.FORM_WIDE, 0),
    ((64, 64), K.FORM_WIDE, 1),
    ((256, 16), K.FORM_WIDE, 1),
    ((192, 6), K.FORM_WIDE, 0),
]


@pytest.mark.parametrize("shape,form,layout", CASES)
@pytest.mark.parametrize("sched,h", [([[0, 1]], 0.0), ([[2, 1]], 0.0), ([[0, 2], [2, 1]], 0.25)])
def test_run_is_bit_exact_against_replay(shape, form, layout, sched, h):
    dev = _dev()
    R = 2 if form == K.FORM_PLAIN else 4
    seed = sum(shape) + len(sched)
    chains, keys, sids, sweeps0 = _chains(R, seed, dev)
    n = int(np.prod(shape))
    rng = np.random.default_rng(seed)
    s0 = (2 * rng.integers(0, 2, size=(R, n)) - 1).astype(np.int8)
    betas = rng.uniform(0.2, 0.9, size=R)
    s = torch.from_numpy(s0.copy()).to(dev)
    e_run, m_run = K.measure(s, shape, R)
    n_meas = 7
    state = K.pack(s, shape, R) if layout == 1 else s
    e_out, m_out, acc = K.run(state, shape, betas, h, chains, np.array(sched), n_meas, e_run,
                              m_run, layout=layout, form=form)
    final = K.unpack(state, shape, R) if layout == 1 else state
    final = final.cpu().numpy()
    e_out, m_out, acc = (t.cpu().numpy() for t in (e_out, m_out, acc))
    for r in range(R):
        s_ref, ref = _replay(s0[r], shape, betas[r], h, keys[r], int(sids[r]), int(sweeps0[r]),
                             sched, n_meas)
        assert np.array_equal(final[r], s_ref), f"spins differ (replica {r})"
        assert np.array_equal(e_out[r], ref["Eb"])
        assert np.array_equal(m_out[r], ref["M"])
        assert np.array_equal(acc[r], ref["accepted"])
    chains.sync()
    per = sum(c for _, c in sched) * n_meas
    assert np.array_equal(chains.sweeps_h, sweeps0 + np.uint64(per))


def test_plain_batched_and_wide_forms_agree():
    dev = _dev()
    shape, R = (64, 32), 2
    out = []
    for form in (K.FORM_PLAIN, K.FORM_BATCHED, K.FORM_WIDE):
        chains, *_ = _chains(R, 21, dev)
        s = torch.empty((R, 64 * 32), dtype=torch.int8, device=dev)
        K.init_spins(s, shape, chains, "random")
        e, m = K.measure(s, shape, R)
        series = K.run(s, shape, [0.44, 0.47], 0.0, chains, np.array([[0, 1]]), 20, e, m, form=form)
        out.append((s.cpu().numpy(), [t.cpu().numpy() for t in series]))
    for other in out[1:]:
        assert np.array_equal(out[0][0], other[0])
        for a, b in zip(out[0][1], other[1]):
            assert np.array_equal(a, b)


def test_full_size_batch_properties():
    """4096 x 256^2 int8, the benchmark batch: bookkeeping, layout and sweep invariants."""
    dev = _dev()
    shape, R = (256, 256), 4096
    chains, *_ = _chains(R, 1234, dev)
    s = torch.empty((R, 256 * 256), dtype=torch.int8, device=dev)
    K.init_spins(s, shape, chains, "random")
    e0, m0 = K.measure(s, shape, R)
    betas = np.random.default_rng(1234).uniform(0.40, 0.48, size=R)
    packed = K.pack(s, shape, R)
    chains_p = Chains(chains.keys_h, chains.sids_h, chains.sweeps_h, dev)
    e_run, m_run = e0.clone(), m0.clone()
    e_out, m_out, acc = K.run(s, shape, betas, 0.0, chains, np.array([[0, 1]]), 5, e_run, m_run)
    e_rp, m_rp = e0.clone(), m0.clone()
    pe, pm, pa = K.run(packed, shape, betas, 0.0, chains_p, np.array([[0, 1]]), 5, e_rp, m_rp,
                       layout=K.LAYOUT_PACKED)
    # the packed image advances exactly like the int8 one
    assert torch.equal(K.unpack(packed, shape, R), s)
    assert torch.equal(e_out, pe) and torch.equal(m_out, pm) and torch.equal(acc, pa)
    # incremental bookkeeping equals a recount on every replica
    e_chk, m_chk = K.measure(s, shape, R)
    assert torch.equal(e_chk, e_run) and torch.equal(m_chk, m_run)
    assert torch.equal(e_out[:, -1], e_run)
    # spot-check two replicas against the replay
    s_np = s.cpu().numpy()
    init = torch.empty((R, 256 * 256), dtype=torch.int8, device=dev)
    K.init_spins(init, shape, Chains(chains.keys_h, chains.sids_h, np.zeros(R, np.uint64), dev))
    init_np = init.cpu().numpy()
    for r in (0, 4095):
        s_ref, ref = _replay(init_np[r], shape, betas[r], 0.0, chains.keys_h[r], int(chains.sids_h[r]),
                             0, [[0, 1]], 5)
        assert np.array_equal(s_np[r], s_ref)
        assert np.array_equal(e_out[r].cpu().numpy(), ref["Eb"])


def test_simulation_chunks_and_bookkeeping():
    from critmc_amd import Ising, Simulation, square

    runs = []
    for chunks in ((100,), (30, 70), (1, 99)):
        sim = Simulation(Ising(square(32)), 0.44, "metropolis+heatbath", seed=7)
        parts = [sim.run(k) for k in chunks]
        runs.append((np.concatenate([p["E"] for p in parts]), sim.state))
        e_b, m = oracle.ising_energy(sim.state, (32, 32))
        assert sim.running == {"e_b": e_b, "m": m}
    for E, st in runs[1:]:
        assert np.array_equal(E, runs[0][0]) and np.array_equal(st, runs[0][1])


def test_simulation_matches_replay_from_reference_start():
    from critmc_amd import Ising, Simulation, make_rng, square
    from critmc_amd.rng import chain_key

    sim = Simulation(Ising(square(16), h=0.1), 0.5, "heatbath", seed=42)
    ts = sim.run(50)
    rng = make_rng(42)
    s0 = Ising(square(16)).initial_state("random", rng)
    key = chain_key(rng)
    s_ref, ref = _replay(s0, (16, 16), 0.5, 0.1, key, 0, 0, [[2, 1]], 50)
    assert np.array_equal(sim.state, s_ref)
    assert np.array_equal(ts["Eb"], ref["Eb"])
    assert np.allclose(ts["E"], ref["Eb"] - 0.1 * ref["M"])
