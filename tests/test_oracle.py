"""CPU tests of the oracle (oracle/ps5_oracle.c), pinned by what the
mathematics fixes: Python's own bytes ordering, closed-form keys, lower-bound
classification, invariants, and mistakes planted in the oracle's output that
the verifier must reject."""
import bisect
import os

import numpy as np
import pytest

from strsets import exhaustive, expected, pack, random_strings, unpack

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---- keys ------------------------------------------------------------------

def test_get_key_closed_form(oracle):
    # the key is the big-endian integer of the first 8 characters, zero-padded
    assert oracle.get_key(b"abcdefghij", 0) == int.from_bytes(b"abcdefgh", "big")
    assert oracle.get_key(b"abcdefghij", 2) == int.from_bytes(b"cdefghij", "big")
    assert oracle.get_key(b"ab", 0) == int.from_bytes(b"ab" + b"\0" * 6, "big")
    assert oracle.get_key(b"", 0) == 0
    assert oracle.get_key(b"\xff" * 9, 0) == 2**64 - 1
    assert oracle.get_key(b"a\xffb", 0) == int.from_bytes(b"a\xffb" + b"\0" * 5, "big")


def test_key_order_matches_bytes_order(oracle):
    strs = random_strings(400, 0, 12, b"ab\x01\x80\xff", 3)
    for i in range(0, 399):
        a, b = strs[i], strs[i + 1]
        ka, kb = oracle.get_key(a, 0), oracle.get_key(b, 0)
        pa, pb = (a[:8] + b"\0" * 8)[:8], (b[:8] + b"\0" * 8)[:8]
        assert (ka > kb) - (ka < kb) == (pa > pb) - (pa < pb)


def test_lcp_key_helpers(oracle):
    lib = oracle.load()
    rng = np.random.default_rng(5)
    for _ in range(500):
        a = bytes(rng.integers(1, 4, 8).astype(np.uint8))
        b = bytes(rng.integers(1, 4, 8).astype(np.uint8))
        ka, kb = int.from_bytes(a, "big"), int.from_bytes(b, "big")
        h = 0
        while h < 8 and a[h] == b[h]:
            h += 1
        assert lib.ps5o_lcp_key_type(ka, kb) == h
    for L in range(9):
        k = int.from_bytes((b"x" * L + b"\0" * 8)[:8], "big")
        assert lib.ps5o_lcp_key_depth(k) == L


# ---- classifier --------------------------------------------------------------

def test_classifier_is_lower_bound(oracle):
    lib = oracle.load()
    rng = np.random.default_rng(11)
    for dup in (False, True):
        hi = 50 if dup else 2**62
        samples = np.sort(rng.integers(0, hi, 2046).astype(np.uint64))
        keys = rng.integers(0, hi, 20000).astype(np.uint64)
        keys[:2046] = samples
        bkt = np.zeros(len(keys), np.uint16)
        spl = np.zeros(1023, np.uint64)
        slcp = np.zeros(1024, np.uint8)
        lib.ps5o_classify_keys(oracle._p(samples), oracle._p(keys), len(keys), oracle._p(bkt),
                               oracle._p(spl), oracle._p(slcp))
        assert np.all(spl[1:] >= spl[:-1]), "in-order splitters must be sorted"
        s = [int(x) for x in spl]
        for k, b in zip(keys[:3000], bkt[:3000]):
            i = bisect.bisect_left(s, int(k))
            want = 2 * i + (1 if i < 1023 and s[i] == int(k) else 0)
            assert b == want
        # splitter lcp: common leading bytes with the previous splitter, 0x80 if NUL-terminated
        for i in range(1, 1023):
            x = s[i - 1] ^ s[i]
            h = (64 - x.bit_length()) // 8
            assert slcp[i] & 0x7F == h
            assert bool(slcp[i] & 0x80) == ((s[i] & 0xFF) == 0)


# ---- brute force and pS5 -----------------------------------------------------

CASES = [
    ("empty_set", []),
    ("one", [b"abc"]),
    ("all_empty", [b""] * 50),
    ("all_equal_long", [b"x" * 37] * 300),
    ("prefix_chain", [b"a" * i for i in range(60, -1, -1)]),
    ("exhaustive_ab4", exhaustive(b"ab", 4)),
    ("high_bytes", random_strings(3000, 0, 20, bytes([1, 2, 127, 128, 200, 255]), 1)),
    ("alnum", random_strings(5000, 16, 19, b"0123456789ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz", 2)),
    ("dna_long", random_strings(4000, 0, 120, b"AC", 3)),
]


@pytest.mark.parametrize("name,strs", CASES, ids=[c[0] for c in CASES])
def test_brute_matches_python_sorted(oracle, name, strs):
    arena, offs = pack(strs)
    perm, lcp = oracle.brute_sort(arena, offs)
    ep, el = expected(strs)
    assert np.array_equal(perm, ep)
    assert np.array_equal(lcp[1:], el[1:])


@pytest.mark.parametrize("name,strs", CASES, ids=[c[0] for c in CASES])
def test_ps5_matches_brute(oracle, name, strs):
    arena, offs = pack(strs)
    perm, lcp = oracle.ps5_sort(arena, offs, threads=4)
    bp, bl = oracle.brute_sort(arena, offs)
    assert np.array_equal(perm, bp)
    assert np.array_equal(lcp[1:], bl[1:])
    assert oracle.verify(arena, offs, perm, lcp)[0] == 0


@pytest.mark.parametrize("threads", [1, 2, 8])
def test_ps5_parallel_steps(ps5, oracle, threads):
    """n > 1 Mi forces parallel sample-sort steps and sequential S5 below them."""
    arena, offs = ps5.gen_uniform(99, 40 * 2**20, 0, 40, b"ACGT")
    perm, lcp = oracle.ps5_sort(arena, offs, threads=threads)
    assert oracle.verify(arena, offs, perm, lcp)[0] == 0
    bp, bl = oracle.brute_sort(arena, offs)
    assert np.array_equal(perm, bp) and np.array_equal(lcp[1:], bl[1:])


def test_ps5_high_lcp_windows(ps5, oracle):
    arena, offs = ps5.workload("highlcp", scale=0.01)
    perm, lcp = oracle.ps5_sort(arena, offs)
    assert oracle.verify(arena, offs, perm, lcp)[0] == 0
    assert lcp.max() == 100


# ---- the plain contestant's branches (no LCP array) --------------------------
# parallel_sample_sortBTCUI runs pS5 without an LCP array: the generic
# insertion sort (inssort.hpp:61-97) and the lcp == NULL branches.  That is the
# sort the CPU baseline times, so it is checked on its own.

@pytest.mark.parametrize("name,strs", CASES, ids=[c[0] for c in CASES])
def test_ps5_nolcp_matches_brute(oracle, name, strs):
    arena, offs = pack(strs)
    perm, lcp = oracle.ps5_sort(arena, offs, threads=4, with_lcp=False)
    assert lcp is None
    bp, _ = oracle.brute_sort(arena, offs)
    assert np.array_equal(perm, bp)
    assert oracle.verify(arena, offs, perm)[0] == 0


@pytest.mark.parametrize("threads", [1, 2, 8])
def test_ps5_nolcp_parallel_steps(ps5, oracle, threads):
    """n > 1 Mi: parallel steps and sequential S5 without LCPs."""
    arena, offs = ps5.gen_uniform(77, 36 * 2**20, 0, 36, b"ACGT")
    assert len(offs) > 2**20
    perm, _ = oracle.ps5_sort(arena, offs, threads=threads, with_lcp=False)
    bp, _ = oracle.brute_sort(arena, offs)
    assert np.array_equal(perm, bp)


@pytest.mark.parametrize("with_lcp", [False, True])
def test_timed_cpu_sort_output(ps5, oracle, with_lcp):
    """The timed CPU baseline hands back its own output, which is exact."""
    arena, offs = ps5.gen_uniform(0x5EED0002, 24 * 2**20, 1, 64, bytes(range(33, 127)))
    t, perm, lcp = oracle.time_ps5(arena, offs, threads=4, with_lcp=with_lcp, output=True)
    assert t > 0
    assert (lcp is not None) == with_lcp
    assert oracle.verify(arena, offs, perm, lcp)[0] == 0
    # a planted mistake in that output is caught
    p = perm.copy()
    p[10], p[11] = p[11], p[10]
    assert oracle.verify(arena, offs, p, lcp)[0] != 0


# ---- distinguishing prefix (checker.hpp:155-180) -------------------------------

DPREFIX_KATS = [
    # sorted strings -> (D, lcpsum), worked by hand from the definition: per
    # neighbour pair, d = LCP + 1; D += d, plus d - d_prev when d grows
    ([], (0, 0)),
    ([b"abc"], (0, 0)),
    ([b"", b"a"], (2, 0)),
    ([b"a", b"ab", b"b"], (5, 1)),
    ([b"x", b"x", b"x"], (6, 2)),
    ([b"abc", b"abd", b"b"], (7, 2)),
    ([b"ab", b"abc", b"abcd", b"b"], (12, 5)),
]


@pytest.mark.parametrize("strs,want", DPREFIX_KATS)
def test_dprefix_known_answers(oracle, strs, want):
    arena, offs = pack(strs)
    perm = np.arange(len(strs), dtype=np.uint32)
    assert oracle.dprefix(arena, offs, perm) == want


def test_dprefix_lcpsum_and_ties(oracle):
    """lcpsum is the sum of the exact LCP array; D does not depend on the
    order of identical strings; an unsorted order changes D."""
    strs = random_strings(3000, 0, 7, b"abc", 21)
    arena, offs = pack(strs)
    perm, lcp = oracle.brute_sort(arena, offs)
    D, ls = oracle.dprefix(arena, offs, perm)
    assert ls == int(lcp[1:].astype(np.uint64).sum())
    j = next(i for i in range(1, len(perm)) if strs[perm[i]] == strs[perm[i - 1]])
    p = perm.copy()
    p[j - 1], p[j] = p[j], p[j - 1]
    assert oracle.dprefix(arena, offs, p) == (D, ls)
    shuffled = np.random.default_rng(1).permutation(perm).astype(np.uint32)
    assert oracle.dprefix(arena, offs, shuffled)[1] < ls


# ---- the verifier rejects planted mistakes ----------------------------------

def test_verify_rejects_planted_mistakes(oracle):
    strs = random_strings(2000, 0, 6, b"abc", 8)
    arena, offs = pack(strs)
    perm, lcp = oracle.brute_sort(arena, offs)
    assert oracle.verify(arena, offs, perm, lcp)[0] == 0
    # swapped neighbours with different content
    i = next(i for i in range(1, len(perm)) if strs[perm[i]] != strs[perm[i - 1]])
    p = perm.copy(); p[i - 1], p[i] = p[i], p[i - 1]
    assert oracle.verify(arena, offs, p, lcp)[0] != 0
    # reversed tie order
    j = next(i for i in range(1, len(perm)) if strs[perm[i]] == strs[perm[i - 1]])
    p = perm.copy(); p[j - 1], p[j] = p[j], p[j - 1]
    assert oracle.verify(arena, offs, p, lcp)[0] == j + 1
    # off-by-one LCP
    l2 = lcp.copy(); l2[17] += 1
    assert oracle.verify(arena, offs, perm, l2)[0] == 18
    # duplicated index
    p = perm.copy(); p[5] = p[6]
    assert oracle.verify(arena, offs, p, lcp)[0] != 0


# ---- distinguishing characters (verify_lcp_cache semantics) -----------------

@pytest.mark.parametrize("seed,letters,hi", [(1, b"ab", 6), (2, b"ACGT", 12), (3, bytes(range(1, 256)), 4)])
def test_dist_chars_definition(oracle, seed, letters, hi):
    """cache[i] is the first byte of s_i not shared with s_{i-1}: it is NUL
    exactly for equal neighbours and otherwise exceeds the predecessor's byte
    at the same position (which a shifted position or the wrong string fails)"""
    strs = random_strings(3000, 0, hi, letters, seed)
    arena, offs = pack(strs)
    perm, lcp = oracle.brute_sort(arena, offs)
    cache = oracle.dist_chars(arena, offs, perm, lcp)
    for i in range(1, len(strs)):
        a, b = strs[perm[i - 1]] + b"\0", strs[perm[i]] + b"\0"
        h = int(lcp[i])
        assert cache[i] == b[h]
        if strs[perm[i - 1]] == strs[perm[i]]:
            assert cache[i] == 0
        else:
            assert cache[i] > a[h]
    # planted mistakes: a shifted position and the predecessor's character
    bad_shift = oracle.dist_chars(arena, offs, perm, np.maximum(lcp, 1) - 1)
    assert not np.array_equal(bad_shift[1:], cache[1:])
    bad_prev = oracle.dist_chars(arena, offs, np.concatenate([perm[:1], perm[:-1]]), lcp)
    assert not np.array_equal(bad_prev[1:], cache[1:])


# ---- golden fixtures --------------------------------------------------------

def golden_files():
    if not os.path.isdir(GOLDEN):
        return []
    return sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))


@pytest.mark.parametrize("fname", golden_files())
def test_golden_fixture(oracle, fname):
    z = np.load(os.path.join(GOLDEN, fname))
    arena, offs = z["arena"], z["offs"]
    strs = unpack(arena, offs)
    ep, el = expected(strs)
    assert np.array_equal(z["perm"], ep) and np.array_equal(z["lcp"][1:], el[1:])
    perm, lcp = oracle.ps5_sort(arena, offs, threads=2)
    assert np.array_equal(perm, z["perm"]) and np.array_equal(lcp[1:], z["lcp"][1:])
