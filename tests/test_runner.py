"""CPU checks of the psstest-compatible runner's input side (no GPU: only the
host contestant std/stable_sort_strcmp runs): plain and compressed line files
with the decompressed size in the name (src/tools/input.hpp:225-422), suffix
mode (src/psstest.cpp:502-518), the reference's random generators, and the
contestant registry."""
import bz2
import gzip
import lzma
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "parallel-string-sorting_amd", "psstest_gpu")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="psstest_gpu not built")


def run(*args):
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout, r.stderr


def results(out):
    rows = []
    for ln in out.splitlines():
        if ln.startswith("RESULT"):
            rows.append(dict(kv.split("=", 1) for kv in ln.split("\t")[1:]))
    return rows


def sample_text():
    rng = np.random.default_rng(8)
    lines = [bytes(rng.choice(list(b"abcdef"), int(rng.integers(0, 30))).tolist()) for _ in range(3000)]
    return b"\n".join(lines) + b"\n"


def test_list_has_every_contestant():
    rc, out, _ = run("-L")
    names = [ln.split()[0] for ln in out.splitlines() if ln.strip()]
    assert rc == 0
    assert names == sorted(names)
    for want in ["mi355x/parallel_sample_sortBTCU2", "mi355x/parallel_sample_sortBTCU2_lcp",
                 "mi355x/parallel_sample_sortBTCU2_lcp_cache", "mi355x/parallel_sample_sortBTCU2_out",
                 "std/stable_sort_strcmp"]:
        assert want in names


@pytest.mark.parametrize("kind", ["plain", "gz", "bz2", "xz"])
def test_line_files(tmp_path, kind):
    raw = sample_text()
    comp = {"plain": lambda b: b, "gz": gzip.compress, "bz2": bz2.compress, "xz": lzma.compress}[kind]
    name = f"lines.{len(raw)}" + ("" if kind == "plain" else "." + kind)
    path = tmp_path / name
    path.write_bytes(comp(raw))
    rc, out, err = run("-A", "std/stable_sort_strcmp", str(path))
    assert rc == 0, err
    [r] = results(out)
    assert r["status"] == "ok" and r["string_count"] == str(raw.count(b"\n"))
    assert r["char_count"] == str(len(raw))


def test_suffix_mode(tmp_path):
    raw = sample_text()[:5000]
    path = tmp_path / "text.txt"
    path.write_bytes(raw)
    rc, out, err = run("-A", "std/stable_sort_strcmp", "--suffix", str(path))
    assert rc == 0, err
    [r] = results(out)
    assert r["status"] == "ok" and r["string_count"] == str(len(raw))


def test_compressed_without_size_is_an_error(tmp_path):
    path = tmp_path / "lines.gz"
    path.write_bytes(gzip.compress(b"a\nb\n"))
    rc, out, err = run("-A", "std/stable_sort_strcmp", str(path))
    assert rc != 0 and "decompressed size" in err


def test_size_limit_and_reference_random():
    rc, out, err = run("-A", "std/stable_sort_strcmp", "-s", "64K", "random10")
    assert rc == 0, err
    [r] = results(out)
    assert r["status"] == "ok" and r["char_count"] == str(64 * 1024 + 9)
