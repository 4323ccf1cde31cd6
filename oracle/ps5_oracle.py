"""ctypes binding of the CPU oracle (oracle/libps5_oracle.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py load it, as the checker or as the timed CPU
baseline.  See ps5_oracle.c for the reference file:line of every function.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libps5_oracle.so")

_lib = None
_P = ctypes.c_void_p
_U64 = ctypes.c_uint64


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.ps5o_get_key.restype = _U64
    lib.ps5o_get_key.argtypes = [_P, _U64]
    lib.ps5o_lcp_key_type.restype = ctypes.c_uint32
    lib.ps5o_lcp_key_type.argtypes = [_U64, _U64]
    lib.ps5o_lcp_key_depth.restype = ctypes.c_uint32
    lib.ps5o_lcp_key_depth.argtypes = [_U64]
    lib.ps5o_brute_sort.restype = None
    lib.ps5o_brute_sort.argtypes = [_P, _P, _U64, _P, _P]
    lib.ps5o_dist_chars.restype = _U64
    lib.ps5o_dist_chars.argtypes = [_P, _P, _U64, _P, _P, _P, ctypes.c_int]
    lib.ps5o_verify.restype = _U64
    lib.ps5o_verify.argtypes = [_P, _P, _U64, _P, _P, ctypes.c_char_p, _U64]
    lib.ps5o_canonicalize.restype = None
    lib.ps5o_canonicalize.argtypes = [_P, _P, _U64, _P]
    lib.ps5o_dprefix.restype = _U64
    lib.ps5o_dprefix.argtypes = [_P, _P, _U64, _P, ctypes.POINTER(_U64)]
    lib.ps5o_classify_keys.restype = None
    lib.ps5o_classify_keys.argtypes = [_P, _P, _U64, _P, _P, _P]
    lib.ps5o_ps5_sort_ptrs.restype = None
    lib.ps5o_ps5_sort_ptrs.argtypes = [_P, _U64, _P, ctypes.c_int, _U64]
    lib.ps5o_ps5_sort.restype = None
    lib.ps5o_ps5_sort.argtypes = [_P, _P, _U64, _P, _P, ctypes.c_int, _U64]
    lib.ps5o_time_ps5.restype = ctypes.c_double
    lib.ps5o_time_ps5.argtypes = [_P, _P, _U64, ctypes.c_int, ctypes.c_int, _P, _P]
    lib.ps5o_max_threads.restype = ctypes.c_int
    lib.ps5o_max_threads.argtypes = []
    _lib = lib
    return lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def offs64(offs):
    return np.ascontiguousarray(offs, dtype=np.uint64)


def brute_sort(arena, offs):
    """Stable sort by the definition: (perm, lcp)."""
    n = len(offs)
    o = offs64(offs)
    perm = np.zeros(max(1, n), np.uint32)
    lcp = np.zeros(max(1, n), np.uint32)
    load().ps5o_brute_sort(_p(arena), _p(o), n, _p(perm), _p(lcp))
    return perm[:n], lcp[:n]


def ps5_sort(arena, offs, threads=0, seed=7, with_lcp=True):
    """The pS5 restatement, ties canonicalised: (perm, lcp).  with_lcp=False
    runs the plain contestant (no LCP array; lcp is then None)."""
    n = len(offs)
    o = offs64(offs)
    perm = np.zeros(max(1, n), np.uint32)
    lcp = np.zeros(max(1, n), np.uint32) if with_lcp else None
    load().ps5o_ps5_sort(_p(arena), _p(o), n, _p(perm), _p(lcp) if with_lcp else None, threads, seed)
    return perm[:n], (lcp[:n] if with_lcp else None)


def verify(arena, offs, perm, lcp=None):
    """0 if (perm, lcp) is the unique correct output, else 1 + first bad position."""
    n = len(offs)
    o = offs64(offs)
    perm = np.ascontiguousarray(perm, np.uint32)
    msg = ctypes.create_string_buffer(512)
    lp = _p(np.ascontiguousarray(lcp, np.uint32)) if lcp is not None else None
    if lcp is not None:
        lcp = np.ascontiguousarray(lcp, np.uint32)
        lp = _p(lcp)
    r = load().ps5o_verify(_p(arena), _p(o), n, _p(perm), lp, msg, 512)
    return int(r), msg.value.decode()


def dist_chars(arena, offs, perm, lcp):
    """Distinguishing characters cache[i] = s_perm[i][lcp[i]], i >= 1 (cache[0] = 0)."""
    n = len(offs)
    o = offs64(offs)
    perm = np.ascontiguousarray(perm, np.uint32)
    lcp = np.ascontiguousarray(lcp, np.uint32)
    cache = np.zeros(max(1, n), np.uint8)
    load().ps5o_dist_chars(_p(arena), _p(o), n, _p(perm), _p(lcp), _p(cache), 0)
    return cache[:n]


def time_ps5(arena, offs, threads=0, with_lcp=False, output=False):
    """Seconds of one timed pS5 sort (the sort call only).  With output=True
    returns (seconds, perm, lcp) with the timed sort's own canonical output
    (lcp None without with_lcp)."""
    n = len(offs)
    o = offs64(offs)
    perm = np.zeros(max(1, n), np.uint32) if output else None
    lcp = np.zeros(max(1, n), np.uint32) if output and with_lcp else None
    t = load().ps5o_time_ps5(_p(arena), _p(o), n, threads, 1 if with_lcp else 0,
                             _p(perm) if perm is not None else None, _p(lcp) if lcp is not None else None)
    if not output:
        return t
    return t, perm[:n], (lcp[:n] if lcp is not None else None)


def dprefix(arena, offs, perm):
    """(D, lcpsum) of a sorted order: calc_distinguishing_prefix,
    src/tools/checker.hpp:155-180."""
    o = offs64(offs)
    perm = np.ascontiguousarray(perm, np.uint32)
    ls = _U64(0)
    d = load().ps5o_dprefix(_p(arena), _p(o), len(offs), _p(perm), ctypes.byref(ls))
    return int(d), int(ls.value)


def get_key(buf, depth):
    b = np.frombuffer(bytes(buf) + b"\0" * 16, np.uint8).copy()
    return int(load().ps5o_get_key(_p(b), depth))


def max_threads():
    return int(load().ps5o_max_threads())
