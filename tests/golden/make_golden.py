"""Writes the golden fixtures of tests/golden/*.npz.

Each fixture holds an input (arena of NUL-terminated strings, uint32 offsets)
and its expected output: the permutation that sorts the strings by unsigned
byte order with ties in input order, and the LCP array (lcp[0] unused, 0).
Expected outputs come from Python's own bytes ordering, independent of the
oracle and of the GPU code.

Inputs (no reference data files exist for this path):
  - generic_alnum_*: the string sets of the reference's
    tests/test-bingmann-generic.cpp:51-67,223-232 (LCGRandom(1234567),
    length 16 + (rng() >> 8) % 4, letters alnum[(rng() / 100) % 62]);
  - random10_*: psstest's random10 input (src/tools/input.hpp:425-468);
  - exhaustive_ab5, dup_heavy, prefix_chain, high_bytes, windows: edge cases.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MASK = (1 << 64) - 1


class Lcg:   # LCGRandom, src/tools/lcgrandom.hpp:27-40
    def __init__(self, seed):
        self.x = seed & MASK

    def __call__(self):
        self.x = (0x27BB2EE687B0B0FD * self.x + 0xB504F32D) & MASK
        return self.x


ALNUM = b"0123456789ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz"


def generic_alnum(n):
    rng = Lcg(1234567)
    out = []
    for _ in range(n):
        slen = 16 + (rng() >> 8) % 4
        out.append(bytes(ALNUM[(rng() // 100) % 62] for _ in range(slen)))
    return out


def random10(size):
    rng = Lcg(1234567)
    letters = b"0123456789"
    data = bytearray(size)
    slen = 0
    starts = []
    for i in range(size):
        if i == slen:
            starts.append(i)
            slen += rng() % 3 + 16
        if i + 1 == slen:
            data[i] = 0
        else:
            data[i] = letters[(rng() // 100) % 10]
    data[size - 1] = 0
    return [bytes(data[s:data.index(0, s)]) for s in starts]


def exhaustive(alpha, maxlen):
    import itertools
    out = [b""]
    for L in range(1, maxlen + 1):
        out += [bytes(t) for t in itertools.product(alpha, repeat=L)]
    return out


def pack(strings):
    offs = np.zeros(len(strings), np.uint32)
    parts, pos = [], 0
    for i, s in enumerate(strings):
        offs[i] = pos
        parts.append(s + b"\0")
        pos += len(s) + 1
    return np.frombuffer(b"".join(parts) + b"\0" * 16, np.uint8).copy(), offs


def expected(strings):
    perm = sorted(range(len(strings)), key=lambda i: strings[i])
    lcp = [0] * len(strings)
    for i in range(1, len(strings)):
        a, b = strings[perm[i - 1]], strings[perm[i]]
        h = 0
        while h < len(a) and h < len(b) and a[h] == b[h]:
            h += 1
        lcp[i] = h
    return np.array(perm, np.uint32), np.array(lcp, np.uint32)


def main():
    rs = np.random.default_rng(20240601)
    sets = {
        "generic_alnum_256": generic_alnum(256),
        "generic_alnum_4096": generic_alnum(4096),
        "random10_64k": random10(65536),
        "exhaustive_ab5": [s for s in rs.permutation(np.array(exhaustive(b"ab", 5), dtype=object))],
        "dup_heavy": [bytes(rs.choice(list(b"xy"), rs.integers(0, 4)).tolist()) for _ in range(3000)],
        "prefix_chain": [b"q" * L for L in rs.permutation(200)] + [b"q" * 50 + b"r" * L for L in range(20)],
        "high_bytes": [bytes(rs.choice([1, 2, 127, 128, 129, 200, 255], rs.integers(0, 12)).tolist())
                       for _ in range(3000)],
    }
    text = bytes(rs.choice(list(b"ACGT"), 5000).tolist())
    text = text[:2000] + text[:1000] + text[3000:]
    sets["windows"] = [text[o:o + 60] for o in rs.integers(0, len(text) - 60, 3000)]
    for name, strs in sets.items():
        strs = [bytes(s) for s in strs]
        arena, offs = pack(strs)
        perm, lcp = expected(strs)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), arena=arena, offs=offs, perm=perm, lcp=lcp)
        print(f"{name}: {len(strs)} strings")


if __name__ == "__main__":
    main()
