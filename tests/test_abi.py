"""CPU checks of the C ABI library: it loads without a GPU and exports every
function include/ps5_mi355x.h declares; the product has no CPU fallback."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ps5_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "sizeof", "defined"}
    return sorted(set(n for n in names if n not in skip and (n.startswith("ps5_") or n.startswith("mi355x_"))))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["mi355x_ps5_sort", "mi355x_ps5_sort_lcp", "ps5_sort_device", "ps5_ctx_create"]:
        assert must in names


def test_library_exports_every_declared_symbol(ps5):
    lib = ps5.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert sorted(ps5.EXPORTS) == declared_functions()


def test_exports_are_c_symbols(ps5):
    out = subprocess.run(["nm", "-D", "--defined-only", ps5.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    syms = set(line.split()[-1] for line in out.splitlines() if line.strip())
    for name in declared_functions():
        assert name in syms, f"{name} is not an unmangled export"


def test_generators_are_deterministic(ps5):
    a1, o1 = ps5.gen_uniform(1, 100000, 1, 20, b"abc")
    a2, o2 = ps5.gen_uniform(1, 100000, 1, 20, b"abc")
    assert (a1 == a2).all() and (o1 == o2).all()
    a3, _ = ps5.gen_uniform(2, 100000, 1, 20, b"abc")
    assert not (a1[:1000] == a3[:1000]).all()
    # every string within the length bounds and alphabet
    strs = bytes(a1).split(b"\0")[: len(o1)]
    assert all(1 <= len(s) <= 20 and set(s) <= set(b"abc") for s in strs)


def test_reference_random_layout(ps5):
    """psstest's random10: strings of 15..17 letters (16..18 with the NUL)"""
    arena, offs = ps5.gen_reference_random(1 << 16, b"0123456789")
    assert offs[0] == 0
    lens = [int(offs[i + 1] - offs[i] - 1) for i in range(len(offs) - 1)]
    assert set(lens) <= {15, 16, 17}
    assert arena[(1 << 16) - 1] == 0


def test_missing_library_fails_loudly(ps5, monkeypatch):
    monkeypatch.setattr(ps5, "_lib", None)
    monkeypatch.setattr(ps5, "LIB_PATH", "/nonexistent/libps5_mi355x.so")
    with pytest.raises(ps5.Ps5Error):
        ps5.load()


def test_reference_random_matches_fixture(ps5):
    """the C++ restatement of psstest's random10 equals the independent
    Python restatement stored in the golden fixture"""
    import numpy as np
    from strsets import unpack
    z = np.load(os.path.join(ROOT, "tests", "golden", "random10_64k.npz"))
    arena, offs = ps5.gen_reference_random(65536, b"0123456789")
    assert unpack(arena, offs) == unpack(z["arena"], z["offs"])


def test_batch_generator_layout(ps5):
    """sets lie back to back, each 8-byte aligned, NUL-padded, with counts in range"""
    arena, offs, sf = ps5.gen_batch(0x5EED0004, 64, 1000, 20000, 4, 48)
    assert sf[0] == 0 and np.all(np.diff(sf.astype(np.int64)) >= 1000) and np.all(np.diff(sf.astype(np.int64)) <= 20000)
    starts = offs[sf[:-1].astype(np.int64)]
    assert np.all(starts % 8 == 0) and starts[0] == 0
    for k in range(64):
        b, e = int(sf[k]), int(sf[k + 1])
        o = offs[b:e].astype(np.int64)
        lens = np.diff(np.append(o, o[-1] + 1 + int(np.argmax(arena[o[-1]:] == 0)))) - 1
        assert lens.min() >= 4 and lens.max() <= 48
        end = int(starts[k + 1]) if k + 1 < 64 else arena.size - 16
        tail = arena[o[-1] + lens[-1]:end]
        assert np.all(tail == 0) and end - (o[-1] + lens[-1] + 1) < 8
    again = ps5.gen_batch(0x5EED0004, 64, 1000, 20000, 4, 48)
    assert np.array_equal(again[0], arena) and np.array_equal(again[1], offs)
