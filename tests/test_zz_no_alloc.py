"""pyr_compact_plan_static (the kill-order plan on static device scratch, for
callers that cannot allocate) against pyr_compact_plan, through the C ABI:
two calls in a row, the second with a smaller input, bit-exact move lists and
counts; sizes beyond the static limit are rejected before any device work."""

import ctypes as C

import numpy as np
import pytest
import torch

from pyrticle_amd import _lib

pytestmark = pytest.mark.gpu

STATIC_MAX = 1048576      # PYR_STATIC_PLAN_MAX in include/pyrticle_mi355x.h


def _plan(fn, dead):
    n = dead.shape[0]
    frm = torch.full((max(n, 1),), -1, dtype=torch.int32, device=dead.device)
    to = torch.full((max(n, 1),), -1, dtype=torch.int32, device=dead.device)
    n_new, n_moves = C.c_int64(-1), C.c_int64(-1)
    _lib.sync_stream()
    rc = fn(n, _lib.ptr(dead), _lib.ptr(frm), _lib.ptr(to), C.byref(n_new), C.byref(n_moves))
    _lib.check(rc)
    k = int(n_moves.value)
    return int(n_new.value), k, frm[:k].cpu().numpy(), to[:k].cpu().numpy()


def _dead(n, frac, seed, cuda):
    rng = np.random.default_rng(seed)
    return torch.from_numpy((rng.uniform(size=n) < frac).astype(np.uint8)).to(cuda)


def test_static_plan_matches_allocating_plan(cuda):
    L = _lib.lib()
    cases = [(600_000, 0.1, 1), (3_000, 0.3, 2), (STATIC_MAX, 0.02, 3), (1, 1.0, 4),
             (5_000, 1.0, 5), (7_000, 0.0, 6)]
    for n, frac, seed in cases:
        print("static plan case n=%d dead fraction %.2f" % (n, frac), flush=True)
        dead = _dead(n, frac, seed, cuda)
        want = _plan(L.pyr_compact_plan, dead)
        got = _plan(L.pyr_compact_plan_static, dead)
        assert got[0] == want[0] and got[1] == want[1], (n, got[:2], want[:2])
        assert np.array_equal(got[2], want[2]) and np.array_equal(got[3], want[3]), n
        assert want[0] == n - int(dead.sum())


def test_static_plan_rejects_oversize(cuda):
    L = _lib.lib()
    dead = torch.zeros(16, dtype=torch.uint8, device=cuda)
    n_new, n_moves = C.c_int64(-1), C.c_int64(-1)
    rc = L.pyr_compact_plan_static(STATIC_MAX + 1, _lib.ptr(dead), _lib.ptr(dead),
                                   _lib.ptr(dead), C.byref(n_new), C.byref(n_moves))
    assert rc == _lib.PYR_ERR_INVALID_ARG
    assert b"PYR_STATIC_PLAN_MAX" in L.pyr_last_error()
    assert n_new.value == -1 and n_moves.value == -1
    # the entry point still works after the rejected call
    d = _dead(1000, 0.5, 7, cuda)
    got, want = _plan(L.pyr_compact_plan_static, d), _plan(L.pyr_compact_plan, d)
    assert got[:2] == want[:2] and np.array_equal(got[2], want[2])
    assert np.array_equal(got[3], want[3])
