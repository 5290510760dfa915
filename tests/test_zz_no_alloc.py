"""colo_scan_slots_noalloc (the entry point that allocates nothing) against
colo_scan_slots, through the C ABI: bit-exact offsets and totals on calls
in a row, each later one smaller than the one before (the static workspace
keeps the previous call's contents), tile-boundary sizes, n = 0, and the
rejection of n above COLO_SCAN_NOALLOC_MAX_N."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def ctx(colo_mod, datagen):
    _say("no-alloc scan: context")
    labels, x, y, S, p = datagen.generate("c1", N=4000)
    fc = np.bincount(labels, minlength=p["F"]).astype(np.uint32)
    return colo_mod.Context(x.astype(np.float32), y.astype(np.float32), fc, np.array([0, 0, S, S], np.float32), p["r"])


def _scan(colo_mod, ctx, fn, slots, n):
    torch = ctx.torch
    with torch.cuda.stream(ctx.tstream):
        offs = torch.full((n + 1,), -1, dtype=torch.int64, device="cuda")
        tot = C.c_uint64(0)
        rc = getattr(colo_mod.lib(), fn)(ctx.h, C.c_void_p(slots.data_ptr()), n, C.c_void_p(offs.data_ptr()),
                                         C.byref(tot), C.c_void_p(ctx.stream))
        assert rc == 0, colo_mod.lib().colo_last_error(ctx.h)
        host = offs.cpu().numpy().view(np.uint64)
    return host, tot.value


def test_noalloc_scan_matches_scan_slots(colo_mod, ctx):
    torch = ctx.torch
    rng = np.random.default_rng(3)
    for n in (5_000_000, 1_234_567, 4097, 2049, 2048, 1, 0):
        _say(f"no-alloc scan n={n}")
        vals = rng.integers(0, 4000, size=max(n, 1)).astype(np.uint32)
        with torch.cuda.stream(ctx.tstream):
            slots = torch.from_numpy(vals.view(np.int32)).cuda()
        ref, rt = _scan(colo_mod, ctx, "colo_scan_slots", slots, n)
        got, gt = _scan(colo_mod, ctx, "colo_scan_slots_noalloc", slots, n)
        np.testing.assert_array_equal(got, ref)
        assert gt == rt
        expect = np.concatenate([[0], np.cumsum(vals[:n].astype(np.uint64))]).astype(np.uint64)
        np.testing.assert_array_equal(got, expect)


def test_noalloc_scan_rejects_oversize(colo_mod, ctx):
    _say("no-alloc scan: n above the limit is rejected before any launch")
    torch = ctx.torch
    with torch.cuda.stream(ctx.tstream):
        buf = torch.zeros(16, dtype=torch.int64, device="cuda")
    n = colo_mod.SCAN_NOALLOC_MAX_N + 1
    tot = C.c_uint64(0)
    rc = colo_mod.lib().colo_scan_slots_noalloc(ctx.h, C.c_void_p(buf.data_ptr()), n, C.c_void_p(buf.data_ptr()),
                                                C.byref(tot), C.c_void_p(ctx.stream))
    assert rc == 6
    assert b"exceeds" in colo_mod.lib().colo_last_error(ctx.h)
    assert int(buf.abs().sum()) == 0
