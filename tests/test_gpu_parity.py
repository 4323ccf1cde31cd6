"""GPU parity: the HIP path (through the C ABI) against the oracle.

Bit-exact bar: the permutation (ties in input order) and every lcp[i], i >= 1.
Small and medium inputs are compared with the oracle's brute-force and pS5
outputs; full-size inputs with the oracle's verifier, which checks
permutation, order, tie order and exact LCPs -- properties that determine the
output uniquely."""
import os
import sys

import numpy as np
import pytest

from strsets import exhaustive, expected, pack, random_strings, unpack

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_against_brute(ctx, oracle, arena, offs, with_lcp=True):
    perm, lcp = ctx.sort(arena, offs, with_lcp=with_lcp)
    bp, bl = oracle.brute_sort(arena, offs)
    assert np.array_equal(perm, bp), f"perm differs at {np.argmax(perm != bp)}"
    if with_lcp:
        bad = np.nonzero(lcp[1:] != bl[1:])[0]
        assert bad.size == 0, f"lcp differs at {bad[:5] + 1}: got {lcp[bad[:5] + 1]} want {bl[bad[:5] + 1]}"


SMALL = [
    ("one", [b"abc"]),
    ("two_equal", [b"ab", b"ab"]),
    ("all_empty", [b""] * 50),
    ("all_equal_long", [b"x" * 37] * 300),
    ("prefix_chain", [b"a" * i for i in range(60, -1, -1)]),
    ("exhaustive_ab4", exhaustive(b"ab", 4)),
    ("exhaustive_abc5", exhaustive(b"abc", 5)),
    ("high_bytes", random_strings(3000, 0, 20, bytes([1, 2, 127, 128, 200, 255]), 1)),
    ("alnum_16_19", random_strings(5000, 16, 19, b"0123456789ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz", 2)),
    ("long_ac", random_strings(4000, 0, 120, b"AC", 3)),
    ("equal_8_prefix", [b"abcdefgh" + s for s in random_strings(2000, 0, 9, b"xy", 4)]),
]


@pytest.mark.parametrize("name,strs", SMALL, ids=[c[0] for c in SMALL])
def test_small_sets(ctx, oracle, name, strs):
    say("case", name, len(strs))
    arena, offs = pack(strs)
    check_against_brute(ctx, oracle, arena, offs)


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 100, 1023, 1024, 1025, 5000, 65536, 65537, 200000])
def test_size_classes(ctx, oracle, n):
    """root item as tiny / small / medium / large segment"""
    say("size", n)
    strs = random_strings(n, 0, 12, b"abcd", n)
    arena, offs = pack(strs)
    check_against_brute(ctx, oracle, arena, offs)


@pytest.mark.parametrize("letters", [b"01", b"ACGT", b"0123456789",
                                     b"0123456789ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz",
                                     bytes(range(1, 256))],
                         ids=["random2", "random4", "random10", "random62", "random255"])
def test_reference_random_inputs(ps5, ctx, oracle, letters):
    """the reference's own artificial inputs (src/tools/input.hpp:425-468), 4 MiB"""
    say("reference random", len(letters))
    arena, offs = ps5.gen_reference_random(4 * 2**20, letters)
    check_against_brute(ctx, oracle, arena, offs)


def test_without_lcp(ctx, oracle):
    strs = random_strings(30000, 0, 30, b"abc", 9)
    arena, offs = pack(strs)
    check_against_brute(ctx, oracle, arena, offs, with_lcp=False)


def test_random10mb_full(ps5, ctx, oracle):
    """10 MiB, L in [8,32] over '0'..'9': against the pS5 restatement"""
    say("random10mb")
    arena, offs = ps5.workload("random10mb")
    perm, lcp = ctx.sort(arena, offs)
    op, ol = oracle.ps5_sort(arena, offs)
    assert np.array_equal(perm, op)
    assert np.array_equal(lcp[1:], ol[1:])


def test_ascii_scaled(ps5, ctx, oracle):
    """ASCII 33..126, L in [1,64], 128 MiB (1/16 of the 2 GiB workload)"""
    say("ascii 128MiB")
    arena, offs = ps5.workload("ascii2gb", scale=1 / 16)
    perm, lcp = ctx.sort(arena, offs)
    r, msg = oracle.verify(arena, offs, perm, lcp)
    assert r == 0, msg
    op, ol = oracle.ps5_sort(arena, offs)
    assert np.array_equal(perm, op) and np.array_equal(lcp[1:], ol[1:])


def test_highlcp_scaled(ps5, ctx, oracle):
    """1M windows of 100 ACGT characters with 5% 2 kb repeats (1/20 scale)"""
    say("highlcp 1/20")
    arena, offs = ps5.workload("highlcp", scale=0.05)
    perm, lcp = ctx.sort(arena, offs)
    op, ol = oracle.ps5_sort(arena, offs)
    assert np.array_equal(perm, op)
    assert np.array_equal(lcp[1:], ol[1:])
    assert lcp.max() == 100


def test_repeated_calls_shrinking(ctx, oracle):
    """buffers are reused across calls; results must not depend on history"""
    for n, seed in [(300000, 1), (5000, 2), (70000, 3), (10, 4)]:
        strs = random_strings(n, 0, 16, b"ab", seed)
        arena, offs = pack(strs)
        check_against_brute(ctx, oracle, arena, offs)


def test_host_entry_contiguous_and_packed(ps5, oracle):
    """Contest-shaped entry on host pointers: lcp[0] untouched, lcp[1..] exact"""
    strs = random_strings(20000, 0, 10, b"xyz", 12)
    out, order, lcp = ps5.sort_strings(strs, with_lcp=True)
    ep, el = expected(strs)
    assert order == list(ep)
    assert lcp[0] == 42
    assert np.array_equal(lcp[1:].astype(np.uint32), el[1:])


@pytest.mark.parametrize("letters,hi,n", [(b"ab", 8, 20000), (b"ACGT", 30, 150000),
                                          (bytes(range(1, 256)), 3, 5000)], ids=["ab", "acgt", "bytes"])
def test_dist_chars_device(ctx, oracle, letters, hi, n):
    """distinguishing characters on the device equal the oracle's; cache[0] untouched"""
    strs = random_strings(n, 0, hi, letters, n + hi)
    arena, offs = pack(strs)
    perm, lcp, cache = ctx.sort_cache(arena, offs)
    bp, bl = oracle.brute_sort(arena, offs)
    assert np.array_equal(perm, bp) and np.array_equal(lcp[1:], bl[1:])
    want = oracle.dist_chars(arena, offs, bp, bl)
    assert cache[0] == 0xEE
    assert np.array_equal(cache[1:], want[1:])


def test_host_entry_lcp_cache(ps5, oracle):
    """func_lcp_cache_type shape: lcp[0] and cache[0] untouched"""
    strs = random_strings(30000, 0, 14, b"xyz0", 21)
    out, order, lcp, cache = ps5.sort_strings(strs, with_cache=True)
    arena, offs = pack(strs)
    bp, bl = oracle.brute_sort(arena, offs)
    assert order == list(bp)
    assert lcp[0] == 42 and cache[0] == 0xEE
    assert np.array_equal(lcp[1:].astype(np.uint32), bl[1:])
    assert np.array_equal(cache[1:], oracle.dist_chars(arena, offs, bp, bl)[1:])


def golden_files():
    if not os.path.isdir(GOLDEN):
        return []
    return sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))


@pytest.mark.parametrize("fname", golden_files())
def test_golden(ctx, fname):
    z = np.load(os.path.join(GOLDEN, fname))
    perm, lcp = ctx.sort(z["arena"], z["offs"])
    assert np.array_equal(perm, z["perm"])
    assert np.array_equal(lcp[1:], z["lcp"][1:])


@pytest.mark.slow
def test_ascii2gb_full(ps5, ctx, oracle):
    """the full 2 GiB workload: verified exactly by the oracle's verifier"""
    say("ascii2gb full")
    arena, offs = ps5.workload("ascii2gb")
    perm, lcp = ctx.sort(arena, offs)
    r, msg = oracle.verify(arena, offs, perm, lcp)
    assert r == 0, msg


def _check_batch(oracle, arena, offs64, set_first, perm, lcp):
    for s in range(len(set_first) - 1):
        b, e = int(set_first[s]), int(set_first[s + 1])
        if e == b:
            continue
        o = offs64[b:e]
        op, ol = oracle.brute_sort(arena, o)
        assert np.array_equal(perm[b:e], op), f"set {s}: permutation differs"
        assert np.array_equal(lcp[b + 1:e], ol[1:]), f"set {s}: lcp differs"


def test_batch_mixed_sets(ps5, ctx, oracle):
    """many independent sets of uneven sizes and alphabets in one batch"""
    say("batch 200 sets")
    arena, offs, sf = ps5.gen_batch(0x5EED0004, 200, 1, 30000)
    perm, lcp = ctx.sort_batch(arena, offs, sf)
    _check_batch(oracle, arena, offs, sf, perm, lcp)


def test_batch_edge_sets(ps5, ctx, oracle):
    """empty sets, single strings, duplicates and a large set side by side"""
    from strsets import pack
    sets = [[], [b"z"], [b"ab"] * 70, random_strings(5000, 0, 20, b"ab", 5), [], [b""] * 3,
            random_strings(70000, 1, 12, b"xyz", 6)]
    flat = [s for st in sets for s in st]
    arena, offs32 = pack(flat)
    sf = np.zeros(len(sets) + 1, np.uint64)
    sf[1:] = np.cumsum([len(st) for st in sets])
    perm, lcp = ctx.sort_batch(arena, offs32.astype(np.uint64), sf)
    _check_batch(oracle, arena, offs32.astype(np.uint64), sf, perm, lcp)


@pytest.mark.parametrize("budget_mb,slots", [(4096, 4), (2, 2)])
def test_serve_mix_equals_alone(ps5, ctx, oracle, budget_mb, slots):
    """requests served concurrently (continuous batching) give exactly the
    result each gets when sorted alone -- with a large budget (many requests
    per batch) and a tiny one (requests wait for memory, one batch each)"""
    say("serve", budget_mb, slots)
    reqs = []
    rng = np.random.default_rng(77)
    for i in range(24):
        size = int(np.exp(rng.uniform(np.log(10_000), np.log(3_000_000))))
        if i % 3 == 0:
            a, o = ps5.gen_prefixed(1000 + i, size)
        elif i % 3 == 1:
            a, o = ps5.gen_uniform(2000 + i, size, 0, 20, b"ACGT")
        else:
            a, o = ps5.gen_uniform(3000 + i, size, 1, 64, ps5.ALPHA_ASCII94)
        reqs.append((a, o))
    alone = [ctx.sort(a, o) for a, o in reqs]
    srv = ps5.Server(0, hbm_budget=budget_mb << 20, slots=slots)
    try:
        ids = [srv.submit(a, o) for a, o in reqs]
        mixed = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    say("serve stats", st)
    assert st["completed"] == len(reqs)
    for (a, o), (pa, la), (pm, lm, ms) in zip(reqs, alone, mixed):
        assert np.array_equal(pa, pm)
        assert np.array_equal(la[1:], lm[1:])
        assert oracle.verify(a, o, pm, lm)[0] == 0
