"""CPU checks of the C ABI library: it loads without a GPU and exports every
function include/ps5_mi355x.h declares; the product has no CPU fallback."""
import os
import re
import subprocess

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
