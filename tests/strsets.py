"""Small string-set builders shared by the tests (host numpy only)."""
import itertools

import numpy as np


def pack(strings):
    """list of bytes -> (arena uint8 with 16 zero bytes padding, offs uint32)"""
    offs = np.zeros(len(strings), np.uint32)
    parts = []
    pos = 0
    for i, s in enumerate(strings):
        assert b"\0" not in s
        offs[i] = pos
        parts.append(s + b"\0")
        pos += len(s) + 1
    arena = np.frombuffer(b"".join(parts) + b"\0" * 16, np.uint8).copy()
    return arena, offs


def random_strings(n, lo, hi, alphabet, seed):
    rng = np.random.default_rng(seed)
    alpha = np.frombuffer(bytes(alphabet), np.uint8)
    lens = rng.integers(lo, hi + 1, n)
    return [alpha[rng.integers(0, len(alpha), L)].tobytes() for L in lens]


def exhaustive(alphabet, maxlen):
    out = [b""]
    for L in range(1, maxlen + 1):
        out += [bytes(t) for t in itertools.product(alphabet, repeat=L)]
    return out


def unpack(arena, offs):
    a = arena.tobytes()
    return [a[o:a.index(b"\0", o)] for o in offs]


def expected(strings):
    """(perm, lcp) by definition: stable sort of python bytes (unsigned order)."""
    perm = sorted(range(len(strings)), key=lambda i: strings[i])
    lcp = [0] * len(strings)
    for i in range(1, len(strings)):
        a, b = strings[perm[i - 1]], strings[perm[i]]
        h = 0
        while h < len(a) and h < len(b) and a[h] == b[h]:
            h += 1
        lcp[i] = h
    return np.array(perm, np.uint32), np.array(lcp, np.uint32)
