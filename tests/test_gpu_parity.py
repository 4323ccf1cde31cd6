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


@pytest.mark.parametrize("copies", [2, 3, 7, 40, 70])
def test_duplicate_groups_in_blocks(ctx, oracle, copies):
    """many small groups of identical long strings that share 8-character
    prefixes with others: blocks hold many short runs of equal keys (packed
    into shared wave passes, up to 64 per block) and runs longer than a wave;
    identical strings must keep input order"""
    say("duplicate groups", copies)
    rng = np.random.default_rng(100 + copies)
    base = [bytes(rng.choice(list(b"ACGT"), int(rng.integers(24, 90))).tolist()) for _ in range(120000 // copies)]
    strs = [b for b in base for _ in range(copies)]
    order = rng.permutation(len(strs))
    strs = [strs[i] for i in order]
    arena, offs = pack(strs)
    perm, lcp = ctx.sort(arena, offs)
    r, msg = oracle.verify(arena, offs, perm, lcp)
    assert r == 0, msg
    op, ol = oracle.ps5_sort(arena, offs)
    assert np.array_equal(perm, op) and np.array_equal(lcp[1:], ol[1:])


@pytest.mark.parametrize("groups,seed", [(12, 5), (400, 6)])
def test_tail_blocks_deep_runs(ctx, oracle, groups, seed):
    """groups of 65..300 strings sharing prefixes of 20..600 characters, with
    short differing tails and exact duplicates: in the round whose blocks are
    its last work, each block sorts its long runs again at every deeper depth
    itself (one round in all for the small set, which is a single block)"""
    rng = np.random.default_rng(seed)
    strs = []
    for g in range(groups):
        prefix = bytes(rng.integers(97, 99, size=int(rng.integers(20, 600))).astype(np.uint8))
        tails = [bytes(rng.integers(97, 100, size=int(rng.integers(0, 12))).astype(np.uint8))
                 for _ in range(int(rng.integers(65, 300)))]
        strs += [prefix + t for t in tails] + [prefix] * int(rng.integers(0, 3))
    rng.shuffle(strs)
    say("tail blocks", groups, len(strs))
    arena, offs = pack(strs)
    check_against_brute(ctx, oracle, arena, offs)


def shaped_set(seed):
    """A seeded random shape: string count (log-uniform in 30..300 000),
    alphabet (1 to 255 symbols, high bytes included), length range, a common
    prefix on part of the set and groups of identical copies."""
    rng = np.random.default_rng(7000 + seed)
    n = int(np.exp(rng.uniform(np.log(30), np.log(300000))))
    alphabets = [b"a", b"ab", b"abc", b"ACGT", b"0123456789", bytes(range(97, 123)), bytes(range(33, 127)),
                 bytes(range(1, 256))]
    alpha = alphabets[int(rng.integers(len(alphabets)))]
    lo = int(rng.integers(0, 20))
    hi = lo + int(rng.integers(0, 120))
    strs = random_strings(n, lo, hi, alpha, 7000 + seed)
    if rng.random() < 0.5:      # a shared prefix (up to 47 characters) on a fraction of the set
        pre = bytes(rng.choice(list(alpha), int(rng.integers(1, 48))).tolist())
        frac = rng.random()
        strs = [pre + s if rng.random() < frac else s for s in strs]
    if rng.random() < 0.4:      # every string a copy of one of fewer base strings
        base = strs[:max(1, n // int(rng.integers(2, 100)))]
        strs = [base[i] for i in rng.integers(0, len(base), n)]
    return strs


@pytest.mark.parametrize("seed", range(24))
def test_random_shapes(ctx, oracle, seed):
    """seeded sweep over set shapes; the oracle's verifier fixes the output
    uniquely (order, tie order, permutation, every LCP)"""
    strs = shaped_set(seed)
    say("shape", seed, len(strs), max(len(s) for s in strs))
    arena, offs = pack(strs)
    perm, lcp = ctx.sort(arena, offs)
    r, msg = oracle.verify(arena, offs, perm, lcp)
    assert r == 0, msg
    if len(strs) <= 20000:
        check_against_brute(ctx, oracle, arena, offs)


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


def test_host_entry_out_shapes(ps5, oracle):
    """parallel_sample_sortBTCUI_out (func_type: separate output, copied back)
    and parallel_sample_sort_out_base (strings[] untouched, sorted pointers
    in output[]), both against the oracle's order"""
    import ctypes
    strs = random_strings(40000, 0, 11, b"mno", 31)
    blob = b"".join(x + b"\0" for x in strs) + b"\0" * 16
    buf = ctypes.create_string_buffer(blob, len(blob))
    base = ctypes.addressof(buf)
    starts, pos = [], 0
    for x in strs:
        starts.append(base + pos)
        pos += len(x) + 1
    n = len(strs)
    arena, offs = pack(strs)
    bp, _ = oracle.brute_sort(arena, offs)
    want = [starts[i] for i in bp]
    lib = ps5.load()
    arr = (ctypes.c_void_p * n)(*starts)
    lib.mi355x_ps5_sort_out(arr, n)
    assert [arr[i] for i in range(n)] == want
    arr = (ctypes.c_void_p * n)(*starts)
    out = (ctypes.c_void_p * n)()
    lib.mi355x_ps5_sort_out_base(arr, out, n)
    assert [arr[i] for i in range(n)] == starts, "strings[] must be left as it was"
    assert [out[i] for i in range(n)] == want
    # the same for an input the entry has to pack (pointers in descending
    # order; identical strings then keep that reversed input order)
    rev = starts[::-1]
    rp, _ = oracle.brute_sort(*pack(strs[::-1]))
    arr = (ctypes.c_void_p * n)(*rev)
    out = (ctypes.c_void_p * n)()
    lib.mi355x_ps5_sort_out_base(arr, out, n)
    assert [arr[i] for i in range(n)] == rev
    assert [out[i] for i in range(n)] == [rev[i] for i in rp]


@pytest.mark.parametrize("layout", ["every_other", "page_gaps", "far_gap", "descending"])
def test_host_entry_pointer_layouts(ps5, oracle, layout):
    """pointers into one buffer: with gaps of up to a page the span is copied
    as is (gap bytes unread by the sort), otherwise the strings are packed;
    both give the exact order and LCPs"""
    import ctypes
    rng = np.random.default_rng(77)
    strs = random_strings(30000, 0, 12, b"pqr", 5)
    gaps = {"every_other": lambda i: 1 + (i % 2) * 7, "page_gaps": lambda i: int(rng.integers(0, 4000)),
            "far_gap": lambda i: 70000 if i == 15000 else 0, "descending": lambda i: 0}[layout]
    blob, starts = bytearray(), []
    for i, x in enumerate(strs):
        blob += b"\xAB" * gaps(i)          # filler the sort must never look at
        starts.append(len(blob))
        blob += x + b"\0"
    blob += b"\0" * 16
    buf = ctypes.create_string_buffer(bytes(blob), len(blob))
    base = ctypes.addressof(buf)
    ptrs = [base + o for o in starts]
    if layout == "descending":
        ptrs = ptrs[::-1]
        strs = strs[::-1]
    n = len(ptrs)
    arr = (ctypes.c_void_p * n)(*ptrs)
    lcp = np.full(n, 42, np.uint64)
    lib = ps5.load()
    assert lib.ps5_sort_strings_cache(arr, n, lcp.ctypes.data_as(ctypes.c_void_p), None, None) == 0
    pos = {p: i for i, p in enumerate(ptrs)}
    order = [pos[arr[i]] for i in range(n)]
    ep, el = expected(strs)
    assert order == list(ep)
    assert lcp[0] == 42 and np.array_equal(lcp[1:].astype(np.uint32), el[1:])


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


@pytest.mark.parametrize("kind", ["random4", "repeats"])
def test_host_entry_suffixes(ps5, oracle, kind):
    """suffixes of one text (psstest --suffix): the host entry copies the text
    once instead of packing every suffix; deep LCPs take many rounds"""
    rng = np.random.default_rng(5)
    if kind == "random4":
        text = bytes(rng.choice(list(b"ACGT"), 20000).tolist())
    else:
        unit = bytes(rng.choice(list(b"ab"), 37).tolist())
        text = unit * 60 + bytes(rng.choice(list(b"abc"), 500).tolist())
    text = text[:-1] + b"\0"   # the last byte is forced to NUL as psstest does
    order, lcp = ps5.sort_suffixes(text)
    arena = np.frombuffer(text + b"\0" * 16, np.uint8).copy()
    offs = np.arange(len(text), dtype=np.uint32)
    bp, bl = oracle.brute_sort(arena, offs)
    assert np.array_equal(order, bp.astype(np.int64))
    assert lcp[0] == 42 and np.array_equal(lcp[1:].astype(np.uint32), bl[1:])


def test_runner_suffix_and_compressed(tmp_path):
    """psstest_gpu reads a gzip file whose name gives the decompressed size,
    as lines and as suffixes, and every contestant's output checks out"""
    import gzip
    import subprocess
    rng = np.random.default_rng(8)
    lines = [bytes(rng.choice(list(b"abcdef"), int(rng.integers(0, 30))).tolist()) for _ in range(3000)]
    raw = b"\n".join(lines) + b"\n"
    path = tmp_path / f"lines.{len(raw)}.gz"
    path.write_bytes(gzip.compress(raw))
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "parallel-string-sorting_amd", "psstest_gpu")
    for extra in ([], ["--suffix"]):
        r = subprocess.run([exe, "-a", "mi355x", *extra, str(path)], capture_output=True, text=True, timeout=120)
        say(r.stdout[-2000:], r.stderr[-2000:])
        assert r.returncode == 0
        res = [ln for ln in r.stdout.splitlines() if ln.startswith("RESULT")]
        algos = sorted(ln.split("algo=")[1].split("\t")[0] for ln in res)
        assert algos == ["mi355x/lcp_mergesort_8way", "mi355x/parallel_sample_sortBTCU2",
                         "mi355x/parallel_sample_sortBTCU2_lcp", "mi355x/parallel_sample_sortBTCU2_lcp_cache",
                         "mi355x/parallel_sample_sortBTCU2_out"]
        assert all("status=ok" in ln for ln in res)
        count = len(raw) if extra else raw.count(b"\n")
        assert all(f"string_count={count}" in ln for ln in res)


@pytest.mark.parametrize("tuning", [(64, 1, 2), (96, 4, 4), (512, 7, 8), (1024, 10, 32)])
def test_tuning_does_not_change_output(ps5, oracle, tuning):
    """every shape setting -- tiny trees with many levels, huge blocks, little
    oversampling -- sorts to the same, exact output"""
    say("tuning", tuning)
    c = ps5.Context(0)
    try:
        c.set_tuning(*tuning)
        assert c.get_tuning() == tuning
        arena, offs = ps5.workload("ascii2gb", scale=1 / 256)
        perm, lcp = c.sort(arena, offs)
        op, ol = oracle.ps5_sort(arena, offs)
        assert np.array_equal(perm, op) and np.array_equal(lcp[1:], ol[1:])
        strs = random_strings(50000, 0, 10, b"ab", 31)
        a2, o2 = pack(strs)
        p2, l2 = c.sort(a2, o2)
        bp, bl = oracle.brute_sort(a2, o2)
        assert np.array_equal(p2, bp) and np.array_equal(l2[1:], bl[1:])
    finally:
        c.close()


@pytest.mark.parametrize("shape", [(2048, 2048, 1, 1), (4096, 6144, 2048, 2), (9216, 12032, 512, 4),
                                   (6144, 9216, 65536, 8)])
@pytest.mark.parametrize("planning", [0, 1])
def test_shape_does_not_change_output(ps5, oracle, shape, planning):
    """tile shapes (chunk limits, the smallest level's tile count) and the
    leaf passes' key width, with host- and device-planned levels, sort to the
    same exact output: random ASCII, short binary strings with many
    duplicates, and windows with long shared prefixes"""
    say("shape", shape, "planning", planning)
    c = ps5.Context(0)
    try:
        c.set_shape(*shape)
        c.set_planning(planning)
        assert c.get_shape() == shape
        arena, offs = ps5.workload("ascii2gb", scale=1 / 256)
        perm, lcp = c.sort(arena, offs)
        op, ol = oracle.ps5_sort(arena, offs)
        assert np.array_equal(perm, op) and np.array_equal(lcp[1:], ol[1:])
        strs = random_strings(50000, 0, 10, b"ab", 31)
        a2, o2 = pack(strs)
        p2, l2 = c.sort(a2, o2)
        bp, bl = oracle.brute_sort(a2, o2)
        assert np.array_equal(p2, bp) and np.array_equal(l2[1:], bl[1:])
        a3, o3 = ps5.workload("highlcp", scale=0.01)
        p3, l3 = c.sort(a3, o3)
        assert oracle.verify(a3, o3, p3, l3)[0] == 0
    finally:
        c.close()


def test_shape_rejects_bad_values(ps5):
    c = ps5.Context(0)
    try:
        for bad in [(2047, 12032, 512, 8), (9217, 12032, 512, 8), (9216, 12033, 512, 8), (9216, 12032, 0, 8),
                    (9216, 12032, 512, 0), (9216, 12032, 512, 9)]:
            with pytest.raises(ps5.Ps5Error):
                c.set_shape(*bad)
        assert c.get_shape() == (9216, 12032, 512, 8)
    finally:
        c.close()


def test_tuning_rejects_bad_values(ps5):
    c = ps5.Context(0)
    try:
        for bad in [(63, 10, 32), (1025, 10, 32), (256, 0, 32), (256, 11, 32), (256, 10, 3), (256, 10, 34)]:
            with pytest.raises(ps5.Ps5Error):
                c.set_tuning(*bad)
    finally:
        c.close()


def test_autotune_picks_a_valid_setting(ps5, oracle):
    """the tuner leaves a setting in the context and the sort after it is exact"""
    arena, offs = ps5.workload("random10mb")
    n = len(offs)
    c = ps5.Context(0)
    bufs = [ps5.DeviceBuffer(arena.size), ps5.DeviceBuffer(4 * n), ps5.DeviceBuffer(4 * n), ps5.DeviceBuffer(4 * n)]
    try:
        da, do, dp, dl = bufs
        da.upload(arena)
        do.upload(offs)
        best, ms = c.autotune_device(da, int(arena.size), do, n, dp, dl)
        say("autotune", best, ms)
        assert c.get_tuning() == best and ms > 0
        assert 64 <= best[0] <= 1024 and 1 <= best[1] <= 10 and 2 <= best[2] <= 32
        c.sort_device(da, int(arena.size), do, n, dp, dl)
        perm, lcp = dp.download(np.uint32, n), dl.download(np.uint32, n)
        assert oracle.verify(arena, offs, perm, lcp)[0] == 0
    finally:
        for b in bufs:
            b.free()
        c.close()


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


@pytest.mark.slow
def test_large_root_duplicates(ps5, ctx, oracle):
    """1 GiB of 0..12-character strings over 'ab' (153 M strings, almost all
    duplicated): a root large enough for the longest root chunks, with heavy
    equal buckets; verified exactly by the oracle's verifier"""
    say("large root duplicates")
    arena, offs = ps5.gen_uniform(0xB17, 1 << 30, 0, 12, b"ab")
    perm, lcp = ctx.sort(arena, offs)
    r, msg = oracle.verify(arena, offs, perm, lcp)
    assert r == 0, msg


@pytest.mark.slow
def test_highlcp_full(ps5, ctx, oracle):
    """the high-LCP workload at full size: 20 M windows of 100 characters of a
    2^29-base ACGT text with 5% repeated 2 kb segments; verified exactly"""
    say("highlcp full")
    arena, offs = ps5.workload("highlcp")
    assert len(offs) == 20_000_000
    perm, lcp = ctx.sort(arena, offs)
    r, msg = oracle.verify(arena, offs, perm, lcp)
    assert r == 0, msg
    assert lcp.max() == 100            # identical windows exist


@pytest.mark.slow
def test_batch_full(ps5, ctx, oracle):
    """the batch workload at full size: 4096 sets, string counts log-uniform
    in [1000, 2e6], lengths U[4,48], alphabets ACGT / digits / bytes 33..126
    (1.1e9 strings, about 30 GB) in one batched pass; every set verified
    exactly, and a sample of sets identical to sorting that set alone"""
    say("batch full: generating")
    arena, offs, sf = ps5.gen_batch(0x5EED0004, 4096, 1000, 2_000_000, 4, 48)
    say("batch full:", int(sf[-1]), "strings,", arena.size, "bytes")
    perm, lcp = ctx.sort_batch(arena, offs, sf)
    say("batch full: sorted; verifying every set")
    for s in range(4096):
        b, e = int(sf[s]), int(sf[s + 1])
        r, msg = oracle.verify(arena, offs[b:e], perm[b:e], lcp[b:e])
        assert r == 0, f"set {s}: {msg}"
    # the same sets sorted one by one through the plain device path
    for s in range(0, 4096, 256):
        b, e = int(sf[s]), int(sf[s + 1])
        base = int(offs[b])
        local = (offs[b:e] - base).astype(np.uint32)
        end = int(offs[e - 1]) + 64
        sub = np.concatenate([arena[base:end], np.zeros(16, np.uint8)])
        pp, pl = ctx.sort(sub, local)
        assert np.array_equal(pp, perm[b:e]), f"set {s}: batched and alone differ"
        assert np.array_equal(pl[1:], lcp[b + 1:e]), f"set {s}: LCPs differ"


@pytest.mark.slow
def test_serve_full_sizes(ps5, ctx, oracle):
    """serving at the workload's request sizes: log-uniform in [10 KB,
    256 MB] (the largest forced to 256 MB) of prefix-heavy strings (1000
    host prefixes, Zipf 1.1), Poisson arrivals; every result identical to the
    request sorted alone and verified exactly"""
    rng = np.random.default_rng(0x5EED0005)
    sizes = np.exp(rng.uniform(np.log(10_000), np.log(256 << 20), 40)).astype(np.int64)
    sizes[int(np.argmax(sizes))] = 256 << 20
    reqs = [ps5.gen_prefixed(0x5EED0005 * 1000 + i, int(sz)) for i, sz in enumerate(sizes)]
    say("serve full:", len(reqs), "requests,", sum(int(a.size) for a, _ in reqs), "bytes")
    alone = [ctx.sort(a, o) for a, o in reqs]
    srv = ps5.Server(0, hbm_budget=8 << 30, slots=4, max_batch_strings=8 << 20)
    try:
        import time
        gaps = rng.exponential(0.004, len(reqs))
        ids = []
        for (a, o), g in zip(reqs, gaps):
            time.sleep(float(g))
            ids.append(srv.submit(a, o))
        mixed = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    say("serve full stats", st)
    for (a, o), (pa, la), (pm, lm, _) in zip(reqs, alone, mixed):
        assert np.array_equal(pa, pm) and np.array_equal(la[1:], lm[1:])
        assert oracle.verify(a, o, pm, lm)[0] == 0


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


@pytest.mark.parametrize("budget_mb,slots,max_batch", [(4096, 4, 0), (2, 2, 0), (512, 3, 1_000_000)])
def test_serve_mix_equals_alone(ps5, ctx, oracle, budget_mb, slots, max_batch):
    """requests served concurrently (continuous batching) give exactly the
    result each gets when sorted alone -- with a large budget (many requests
    per batch), a tiny one (requests wait for memory, one batch each) and
    pre-allocated slots whose batches are capped by string count"""
    say("serve", budget_mb, slots, max_batch)
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
    srv = ps5.Server(0, hbm_budget=budget_mb << 20, slots=slots, max_batch_strings=max_batch)
    try:
        ids = [srv.submit(a, o) for a, o in reqs]
        mixed = [srv.wait(i) for i in ids]
        st = srv.stats()
        ledger = srv.ledger()
    finally:
        srv.close()
    say("serve stats", st, "ledger", ledger)
    assert st["completed"] == len(reqs)
    # admission ran on the device: every request admitted, every byte released
    assert ledger == {"admitted": len(reqs), "in_use_bytes": 0}
    for (a, o), (pa, la), (pm, lm, ms) in zip(reqs, alone, mixed):
        assert np.array_equal(pa, pm)
        assert np.array_equal(la[1:], lm[1:])
        assert oracle.verify(a, o, pm, lm)[0] == 0


def test_serve_large_requests_run_alone(ps5, ctx, oracle):
    """requests of more than 2^20 strings never share a batch (each is the
    first and only request of its claim), whatever the batch limit allows"""
    reqs = [ps5.gen_uniform(4000 + i, 6 << 20, 1, 8, b"ACGT") for i in range(3)]
    assert all(len(o) > (1 << 20) for _, o in reqs)
    say("serve large", [len(o) for _, o in reqs])
    srv = ps5.Server(0, hbm_budget=4 << 30, slots=1, max_batch_strings=16 << 20)
    try:
        ids = [srv.submit(a, o) for a, o in reqs]
        out = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    assert st["completed"] == len(reqs) and st["batches"] == len(reqs), st
    for (a, o), (p, l, ms) in zip(reqs, out):
        assert oracle.verify(a, o, p, l)[0] == 0


def test_serve_failure_isolation(ps5, ctx, oracle):
    """a request whose sort fails inside a batch (a fault planted from the
    host) fails alone: its batch is run again request by request, every other
    request's result equals its alone-sort, and every slot keeps serving.  An
    over-budget request runs alone and is exact too."""
    reqs = []
    for i in range(12):
        a, o = ps5.gen_uniform(500 + i, 200_000 + 50_000 * i, 0, 24, b"ACGT" if i % 2 else ps5.ALPHA_ASCII94)
        reqs.append((a, o))
    big = ps5.gen_prefixed(77, 24 << 20)          # larger than the whole budget below
    reqs.insert(6, big)
    alone = [ctx.sort(a, o) for a, o in reqs]
    bad_id = 4                                    # ids count from 1 in submission order
    srv = ps5.Server(0, hbm_budget=16 << 20, slots=2)
    try:
        srv.inject_fault(bad_id)
        ids = [srv.submit(a, o) for a, o in reqs]
        out = []
        for rid in ids:
            try:
                out.append(srv.wait(rid))
            except ps5.Ps5Error as e:
                say("failed", rid, e)
                assert rid == bad_id and "injected fault" in str(e)
                out.append(None)
        health = srv.health()
        ledger = srv.ledger()
    finally:
        srv.close()
    say("health", health, "ledger", ledger)
    assert health == {"live_slots": 2, "failed": 1}
    assert ledger == {"admitted": len(reqs), "in_use_bytes": 0}
    for k, ((a, o), (pa, la), res) in enumerate(zip(reqs, alone, out)):
        if ids[k] == bad_id:
            assert res is None
            continue
        pm, lm, _ = res
        assert np.array_equal(pa, pm) and np.array_equal(la[1:], lm[1:])
        assert oracle.verify(a, o, pm, lm)[0] == 0


# ---- LCP-aware merge of sorted runs ------------------------------------------

def _runs_of(oracle, arena, offs, bounds):
    """sort each index range with the brute-force oracle: global idx + run-local LCPs"""
    idx = np.zeros(len(offs), np.uint32)
    lcp = np.zeros(len(offs), np.uint32)
    for b, e in zip(bounds[:-1], bounds[1:]):
        if e > b:
            p, l = oracle.brute_sort(arena, offs[b:e])
            idx[b:e] = p + b
            lcp[b:e] = l
    return idx, lcp


@pytest.mark.parametrize("nruns,letters,hi", [(1, b"ab", 10), (2, b"ab", 12), (3, b"abc", 20), (5, b"ACGT", 40),
                                               (8, bytes(range(1, 256)), 6), (16, b"ab", 30)])
def test_merge_runs_equals_full_sort(ps5, ctx, oracle, nruns, letters, hi):
    """runs that are sorted index ranges merge to the full sort (ties in index order)"""
    n = 40000
    strs = random_strings(n, 0, hi, letters, nruns * 7 + hi)
    arena, offs = pack(strs)
    rng = np.random.default_rng(nruns)
    cuts = np.sort(rng.integers(0, n + 1, nruns - 1))
    bounds = np.concatenate([[0], cuts, [n]]).astype(np.uint64)
    idx, lcp = _runs_of(oracle, arena, offs, bounds)
    mi, ml = ctx.merge(arena, offs, bounds, idx, lcp)
    bp, bl = oracle.brute_sort(arena, offs)
    assert np.array_equal(mi, bp)
    assert np.array_equal(ml[1:], bl[1:])


def test_merge_deep_prefixes_and_empty_runs(ps5, ctx, oracle):
    """long shared prefixes (the LCP rules decide most comparisons) and empty runs"""
    base = b"x" * 200
    strs = [base[:int(i % 190)] + s for i, s in enumerate(random_strings(20000, 0, 30, b"ab", 3))]
    arena, offs = pack(strs)
    n = len(strs)
    bounds = np.array([0, 0, 5000, 5000, 12000, n, n], np.uint64)
    idx, lcp = _runs_of(oracle, arena, offs, bounds)
    mi, ml = ctx.merge(arena, offs, bounds, idx, lcp)
    bp, bl = oracle.brute_sort(arena, offs)
    assert np.array_equal(mi, bp)
    assert np.array_equal(ml[1:], bl[1:])


def test_merge_of_device_sorted_shards(ps5, ctx, oracle):
    """the multi-GPU shape: shards sorted by ps5_sort_device, then merged"""
    arena, offs = ps5.workload("ascii2gb", scale=1 / 512)
    n = len(offs)
    bounds = np.linspace(0, n, 9).astype(np.uint64)
    idx = np.zeros(n, np.uint32)
    lcp = np.zeros(n, np.uint32)
    for b, e in zip(bounds[:-1], bounds[1:]):
        p, l = ctx.sort(arena, offs[b:e])
        idx[b:e] = p + b
        lcp[b:e] = l
    mi, ml = ctx.merge(arena, offs, bounds, idx, lcp)
    assert oracle.verify(arena, offs, mi, ml)[0] == 0


def test_host_lcp_merge_strings(ps5):
    runs = []
    lcps = []
    for r in range(4):
        strs = sorted(random_strings(3000, 0, 12, b"abc", 40 + r))
        l = [0] + [next((h for h in range(min(len(a), len(b))) if a[h] != b[h]), min(len(a), len(b)))
                   for a, b in zip(strs, strs[1:])]
        runs.append(strs)
        lcps.append(l)
    out, lcp = ps5.lcp_merge_strings(runs, lcps)
    flat = [s for run in runs for s in run]
    assert out == sorted(flat)
    el = [next((h for h in range(min(len(a), len(b))) if a[h] != b[h]), min(len(a), len(b)))
          for a, b in zip(out, out[1:])]
    assert lcp[0] == 42 and list(lcp[1:]) == el


# ---------------------------------------------------------------- planning modes

PLAN_MODES = {"host": (0, 0), "device": (1, 0), "auto": (2, 0), "graph2": (0, 2), "graph8": (1, 8)}


def _plan_workload(ps5, kind):
    if kind == "random":          # one root, three rounds at most
        arena, offs = ps5.gen_uniform(0x5EED0101, 4 << 20, 1, 20, b"ACGT")
        return arena, offs.astype(np.uint64), None
    if kind == "highlcp":         # windows with long repeats: more rounds than a graph of 2 or 8 holds
        arena, offs = ps5.workload("highlcp", scale=0.02)
        return arena, offs.astype(np.uint64), None
    arena, offs, sf = ps5.gen_batch(0x5EED0102, 300, 1, 20000)
    return arena, offs, sf


def _relabel(arena, variant):
    """the same strings' shape with other characters (NULs stay NULs)"""
    if variant == 0:
        return arena
    lut = np.arange(256, dtype=np.uint8)
    lut[1:] = np.random.default_rng(variant).permutation(255).astype(np.uint8) + 1
    return lut[arena]


def _sort_in_modes(ps5, ctxs, arena, offs, sf, bufs):
    da, do, dp, dl = bufs
    n = len(offs)
    da.upload(arena)
    out = {}
    for m, c in ctxs.items():
        dp.upload(np.full(n, 0xFFFFFFFF, np.uint32))
        dl.upload(np.full(n, 0xFFFFFFFF, np.uint32))
        if sf is None:
            do.upload(offs.astype(np.uint32))
            c.sort_device(da, int(arena.size), do, n, dp, dl)
        else:
            do.upload(offs)
            c.sort_batch_device(da, int(arena.size), do, sf, dp, dl)
        out[m] = (dp.download(np.uint32, n), dl.download(np.uint32, n), c.work_stats())
    return out


@pytest.mark.parametrize("kind", ["random", "highlcp", "batch"])
def test_planning_modes_agree(ps5, oracle, kind):
    """levels planned on the host, on the device, and rounds captured in a
    HIP graph (2 rounds, then round by round; 8 rounds) give identical
    results; each graph is replayed for new contents of the same buffers, with
    the other modes' sorts and synchronisations in between"""
    arena, offs, sf = _plan_workload(ps5, kind)
    n = len(offs)
    say("planning modes", kind, n)
    bufs = (ps5.DeviceBuffer(arena.size), ps5.DeviceBuffer(n * 8), ps5.DeviceBuffer(n * 4),
            ps5.DeviceBuffer(n * 4))
    ctxs = {}
    try:
        for m, (planning, rounds) in PLAN_MODES.items():
            ctxs[m] = ps5.Context(0)
            ctxs[m].set_planning(planning)
            ctxs[m].set_graph_rounds(rounds)
            assert ctxs[m].get_graph_rounds() == rounds
        for variant in range(3):
            a = _relabel(arena, variant)
            out = _sort_in_modes(ps5, ctxs, a, offs, sf, bufs)
            perm, lcp, _ = out["host"]
            if sf is None:
                assert oracle.verify(a, offs, perm, lcp)[0] == 0
            else:
                _check_batch(oracle, a, offs, sf, perm, lcp)
            for m, (p, l, ws) in out.items():
                assert np.array_equal(p, perm), f"{m} variant {variant}: permutation differs"
                assert np.array_equal(l[1:], lcp[1:]), f"{m} variant {variant}: lcp differs"
            # automatic planning: the device's for a single sort of this size, the host's for a batch
            assert out["auto"][2]["syncs"] == out["host" if sf is not None else "device"][2]["syncs"], kind
            if kind == "random":      # every round inside the graph: one synchronisation
                assert out["graph8"][2]["syncs"] == 1, out["graph8"][2]
    finally:
        for c in ctxs.values():
            c.close()
        for b in bufs:
            b.free()


def test_graph_levels_fallback(ps5, oracle, monkeypatch):
    """a captured round that needs more sample-sort levels than the graph
    holds (forced: one level per round on 4.7M strings) is sorted again with
    host-planned levels and gives the host-planned result, on every replay"""
    arena, offs = ps5.gen_uniform(0x5EED0103, 40 << 20, 1, 16, ps5.ALPHA_DIGITS)
    n = len(offs)
    say("graph fallback", n)
    monkeypatch.setenv("PS5_GRAPH_LEVELS", "1")
    g = ps5.Context(0)
    monkeypatch.delenv("PS5_GRAPH_LEVELS")
    h = ps5.Context(0)
    g.set_graph_rounds(4)
    bufs = (ps5.DeviceBuffer(arena.size), ps5.DeviceBuffer(n * 4), ps5.DeviceBuffer(n * 4), ps5.DeviceBuffer(n * 4))
    try:
        for variant in range(2):
            a = _relabel(arena, variant)
            out = _sort_in_modes(ps5, {"host": h, "graph": g}, a, offs.astype(np.uint64), None, bufs)
            perm, lcp, _ = out["host"]
            assert oracle.verify(a, offs, perm, lcp)[0] == 0
            p, l, ws = out["graph"]
            assert np.array_equal(p, perm) and np.array_equal(l[1:], lcp[1:])
            assert ws["syncs"] > 1, ws       # the host-planned sort after the graph
    finally:
        g.close()
        h.close()
        for b in bufs:
            b.free()


def test_chained_levels_match(ps5, oracle, monkeypatch):
    """a level of at least 4M strings has its successor planned on the
    device without a read-back (chained); with chaining switched off the
    output is the same, and both are exact"""
    arena, offs = ps5.workload("ascii2gb", scale=1 / 12)
    say("chained levels", len(offs))
    on = ps5.Context(0)
    monkeypatch.setenv("PS5_CHAIN_LEVELS", "0")
    off = ps5.Context(0)
    monkeypatch.delenv("PS5_CHAIN_LEVELS")
    on.set_planning(0)     # chaining follows host-planned levels
    off.set_planning(0)
    try:
        p1, l1 = on.sort(arena, offs)
        w1 = on.work_stats()
        p0, l0 = off.sort(arena, offs)
        w0 = off.work_stats()
    finally:
        on.close()
        off.close()
    say("work", w1, w0)
    assert w1["syncs"] < w0["syncs"]
    assert np.array_equal(p1, p0) and np.array_equal(l1[1:], l0[1:])
    assert oracle.verify(arena, offs, p1, l1)[0] == 0


def test_graph_rounds_bounds(ps5):
    c = ps5.Context(0)
    try:
        with pytest.raises(ps5.Ps5Error):
            c.set_graph_rounds(65)
        c.set_graph_rounds(0)
        assert c.get_graph_rounds() == 0
    finally:
        c.close()


def test_planning_mode_values(ps5):
    """modes 0, 1 and 2 are kept; others are rejected and leave the mode as it was"""
    c = ps5.Context(0)
    try:
        env = os.environ.get("PS5_DEV_LEVELS")
        assert c.get_planning() == (min(max(int(env), 0), 2) if env else 0)
        for mode in (0, 1, 2):
            c.set_planning(mode)
            assert c.get_planning() == mode
        for bad in (-1, 3):
            with pytest.raises(ps5.Ps5Error):
                c.set_planning(bad)
            assert c.get_planning() == 2
    finally:
        c.close()
