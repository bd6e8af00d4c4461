"""cm_mbar_objective_noalloc: the MBAR objective with its partials in a static
device array instead of a caller workspace, against cm_mbar_objective.

Both entry points run the same kernels with the same launch shape, so F, the
gradient, the Hessian and ln D(E_u) must agree bit for bit, on two calls in a
row, the second smaller than the first (the static array keeps the first
call's partials; the second call must not read them).
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")


def _inputs(U: int, K: int, seed: int, dev):
    rng = np.random.default_rng(seed)
    betas = np.sort(rng.uniform(0.30, 0.60, size=K))
    energy = rng.normal(0.0, 40.0, size=U)
    logmult = np.log(rng.integers(1, 50, size=U).astype(np.float64)) - np.log(3.0)
    log_n = np.log(rng.integers(1000, 5000, size=K).astype(np.float64))
    f = rng.normal(0.0, 2.0, size=K)
    f -= f[0]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)  # noqa: E731
    return t(energy), t(logmult), t(betas), t(log_n), t(f)


def _with_workspace(lib, ptr, stream, args, U, K, dev):
    E, lm, b, ln, f = args
    nbytes = int(lib.cm_mbar_workspace_bytes(K, 1))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    out = torch.empty(1 + K + K * K, dtype=torch.float64, device=dev)
    ld = torch.empty(U, dtype=torch.float64, device=dev)
    rc = lib.cm_mbar_objective(ptr(E), ptr(lm), U, ptr(b), ptr(ln), ptr(f), K, ptr(out), ptr(ld),
                               ptr(ws), nbytes, stream())
    assert rc == 0, lib.cm_last_error()
    return out, ld


def _no_alloc(lib, ptr, stream, args, U, K, dev):
    E, lm, b, ln, f = args
    out = torch.full((1 + K + K * K,), float("nan"), dtype=torch.float64, device=dev)
    ld = torch.full((U,), float("nan"), dtype=torch.float64, device=dev)
    rc = lib.cm_mbar_objective_noalloc(ptr(E), ptr(lm), U, ptr(b), ptr(ln), ptr(f), K, ptr(out),
                                       ptr(ld), stream())
    assert rc == 0, lib.cm_last_error()
    return out, ld


def test_noalloc_rejects_more_runs_than_its_scratch_holds():
    from critmc_amd._lib import CM_EINVAL, lib

    L = lib()
    P = ctypes.c_void_p
    rc = L.cm_mbar_objective_noalloc(P(0), P(0), 10, P(0), P(0), P(0), 65, P(0), P(0), P(0))
    assert rc == CM_EINVAL
    assert b"1..64" in L.cm_last_error()
    rc = L.cm_mbar_objective_noalloc(P(0), P(0), 0, P(0), P(0), P(0), 4, P(0), P(0), P(0))
    assert rc == CM_EINVAL


@pytest.mark.gpu
@pytest.mark.parametrize("sizes", [((40_000, 64), (3_000, 7)), ((5_000, 12), (1, 1))])
def test_noalloc_matches_workspace_entry_point_on_two_calls(sizes):
    from critmc_amd._lib import lib
    from critmc_amd.device import device, ptr, stream

    dev = device()
    L = lib()
    for call, (U, K) in enumerate(sizes):
        print(f"no-alloc MBAR call {call}: U={U} K={K}", flush=True)
        args = _inputs(U, K, 100 + call, dev)
        ref_out, ref_ld = _with_workspace(L, ptr, stream, args, U, K, dev)
        got_out, got_ld = _no_alloc(L, ptr, stream, args, U, K, dev)
        torch.cuda.synchronize()
        assert torch.isfinite(ref_out).all()
        assert torch.equal(got_out, ref_out), f"call {call}: objective differs"
        assert torch.equal(got_ld, ref_ld), f"call {call}: ln D(E_u) differs"
