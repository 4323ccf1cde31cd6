"""The allocation-free entry point ps5_sort_device_noalloc, called through
the C ABI, against ps5_sort_device: bit-identical permutation and LCP arrays
on two calls in a row (the second with a smaller input, so the static
scratch still holds the first call's larger contents), lcp[0] untouched, and
sizes beyond the scratch's limit rejected before any GPU work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

POISON = 0xA5A5A5A5


def _sort(ps5, ctx, arena, offs, noalloc):
    lib = ps5.load()
    n = len(offs)
    da = ps5.DeviceBuffer(arena.size)
    do = ps5.DeviceBuffer(n * 4)
    dp = ps5.DeviceBuffer(n * 4)
    dl = ps5.DeviceBuffer(n * 4)
    try:
        da.upload(arena)
        do.upload(np.ascontiguousarray(offs, np.uint32))
        dp.upload(np.full(n, POISON, np.uint32))
        dl.upload(np.full(n, POISON, np.uint32))
        fn = lib.ps5_sort_device_noalloc if noalloc else lib.ps5_sort_device
        name = "ps5_sort_device_noalloc" if noalloc else "ps5_sort_device"
        ps5._check(fn(ctx.h, da.ptr, int(arena.size), do.ptr, n, dp.ptr, dl.ptr), name)
        return dp.download(np.uint32, n), dl.download(np.uint32, n)
    finally:
        for b in (da, do, dp, dl):
            b.free()


def test_noalloc_matches_sort_device(ps5, ctx, oracle):
    lib = ps5.load()
    nmax = int(lib.ps5_noalloc_max_strings())
    assert nmax == 1 << 20
    # call 1: exactly the limit (random ASCII, lengths 1..40); call 2: a
    # smaller, duplicate-heavy input (digits, lengths 0..12)
    a1, o1 = ps5.gen_uniform(0x0A110C, 24 << 20, 1, 40, ps5.ALPHA_ASCII94)
    assert len(o1) >= nmax
    o1 = o1[:nmax]
    a2, o2 = ps5.gen_uniform(0x0A110D, 2 << 20, 0, 12, ps5.ALPHA_DIGITS)
    assert len(o2) < nmax
    for k, (a, o) in enumerate([(a1, o1), (a2, o2)]):
        print(f"noalloc call {k + 1}: {len(o)} strings", flush=True)
        pn, ln = _sort(ps5, ctx, a, o, noalloc=True)
        pd, ld = _sort(ps5, ctx, a, o, noalloc=False)
        assert np.array_equal(pn, pd), f"call {k + 1}: permutation differs from ps5_sort_device"
        assert np.array_equal(ln[1:], ld[1:]), f"call {k + 1}: LCPs differ from ps5_sort_device"
        assert ln[0] == POISON, "lcp[0] must stay untouched"
        r, msg = oracle.verify(a, o, pn, ln)
        assert r == 0, msg


def test_noalloc_rejects_oversize(ps5, ctx):
    lib = ps5.load()
    nmax = int(lib.ps5_noalloc_max_strings())
    # the size check comes before any pointer is used: nothing runs on the GPU
    dummy = ps5.DeviceBuffer(64)
    try:
        rc = lib.ps5_sort_device_noalloc(ctx.h, dummy.ptr, 64, dummy.ptr, nmax + 1, dummy.ptr, dummy.ptr)
        assert rc != 0
        assert b"allocation-free limit" in lib.ps5_last_error()
    finally:
        dummy.free()
