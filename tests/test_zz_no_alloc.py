"""gatRangeQueryTestHandlerNoAlloc: the reference-signature range handler for
callers that cannot allocate, its bitmap in the library's fixed device scratch.

It is checked against gatRangeQueryTestHandler (cudaRangeQueryTestHandler's
drop-in, kernel.cu:1282-1312) through the C ABI on two calls in a row, the
second smaller than the first, so a bitmap left over from the first call would
show in the second; both are also checked against a NumPy restatement of the
inclusive containment test (MBB.cpp:21-26).  Bar: bit-exact."""
import ctypes

import numpy as np
import pytest

from gat_amd import _native as N, api

from tests.devbuf import DevBuf

pytestmark = pytest.mark.gpu

GAT_ERR_CAPACITY = 4


def _case(seed, npoints, ntraj, nq, ntasks):
    rng = np.random.default_rng(seed)
    xy = rng.random((npoints, 2), dtype=np.float32)
    tid = rng.integers(1, ntraj, npoints).astype(np.uint32)
    pid = rng.integers(0, 2048, npoints).astype(np.uint32)
    lo = rng.random((nq, 2), dtype=np.float32) * np.float32(0.8)
    side = rng.random((nq, 1), dtype=np.float32) * np.float32(0.2)
    boxes = np.concatenate([lo, lo + side], axis=1).astype(np.float32)
    start = rng.integers(0, npoints, ntasks)
    count = np.minimum(rng.integers(1, 700, ntasks), npoints - start)
    tasks = np.stack([start, count, rng.integers(0, nq, ntasks)], axis=1).astype(np.uint32)
    want = np.zeros((nq, ntraj), np.uint8)
    for s, c, q in tasks:
        p = xy[s:s + c]
        b = boxes[q]
        hit = (p[:, 0] >= b[0]) & (p[:, 0] <= b[2]) & (p[:, 1] >= b[1]) & (p[:, 1] <= b[3])
        want[q, tid[s:s + c][hit]] = 1
    return xy, (tid << 11) | pid, boxes, tasks, want


def _run(h, fn, case, ntraj, nq):
    xy, ids, boxes, tasks, _ = case
    dxy, did, dboxes, dtasks = DevBuf.of(xy), DevBuf.of(ids), DevBuf.of(boxes), DevBuf.of(tasks)
    store = N.PointStore(dxy.ptr, did.ptr, None, len(xy), 11)
    out = np.full(nq * ntraj, 7, np.uint8)
    N.check_hip(fn(ctypes.byref(store), dtasks.ptr, len(tasks), dboxes.ptr, api._ptr(out), ntraj, nq, None))
    return out.reshape(nq, ntraj)


def test_no_alloc_handler_matches_allocating_handler(gpu):
    h = N.hip()
    # 100 x 20,001 slots (the reference's 20k-trajectory batch), then a smaller
    # call: 9 x 5,000 slots, which reuses the start of the same scratch.
    for seed, npoints, ntraj, nq, ntasks in ((1, 400_000, 20_001, 100, 3_000), (2, 50_000, 5_000, 9, 400)):
        case = _case(seed, npoints, ntraj, nq, ntasks)
        ref = _run(h, h.gatRangeQueryTestHandler, case, ntraj, nq)
        got = _run(h, h.gatRangeQueryTestHandlerNoAlloc, case, ntraj, nq)
        assert ref.sum() > 0
        assert (ref == case[4]).all(), "allocating handler against the containment restatement"
        assert (got == ref).all(), f"no-alloc handler differs on the {nq} x {ntraj} call"


def test_no_alloc_handler_limits(gpu):
    h = N.hip()
    case = _case(3, 1_000, 100, 2, 10)
    xy, ids, boxes, tasks, _ = case
    dxy, did, dboxes, dtasks = DevBuf.of(xy), DevBuf.of(ids), DevBuf.of(boxes), DevBuf.of(tasks)
    store = N.PointStore(dxy.ptr, did.ptr, None, len(xy), 11)
    big = np.zeros(16, np.uint8)
    # A bitmap above the scratch is refused before any device work.
    ntraj = (64 << 20) // 2 + 1
    rc = h.gatRangeQueryTestHandlerNoAlloc(ctypes.byref(store), dtasks.ptr, len(tasks), dboxes.ptr, api._ptr(big),
                                           ntraj, 2, None)
    assert rc == GAT_ERR_CAPACITY
    assert b"scratch" in N.hip().gat_last_error()
    # Exactly the scratch size is accepted.
    out = np.zeros(64 << 20, np.uint8)
    N.check_hip(h.gatRangeQueryTestHandlerNoAlloc(ctypes.byref(store), dtasks.ptr, len(tasks), dboxes.ptr,
                                                  api._ptr(out), (64 << 20) // 2, 2, None))
    want = _run(h, h.gatRangeQueryTestHandler, case, 100, 2)
    got = out.reshape(2, (64 << 20) // 2)
    assert (got[:, :100] == want).all() and got[:, 100:].sum() == 0
