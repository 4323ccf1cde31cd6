"""ctypes binding of libps5_mi355x.so, the C ABI declared in include/ps5_mi355x.h.

The library is the product: HIP kernels for gfx950 plus their host
orchestration.  This module only marshals numpy arrays to it.  It does not
import torch (torch ships its own HIP runtime with the same soname) and has no
CPU fallback: when the library is missing every call raises Ps5Error.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libps5_mi355x.so")

PS5_OK = 0

# every symbol include/ps5_mi355x.h declares
EXPORTS = [
    "ps5_last_error",
    "mi355x_ps5_sort", "mi355x_ps5_sort_lcp", "mi355x_ps5_sort_out", "mi355x_ps5_sort_out_base",
    "mi355x_ps5_sort_lcp_cache",
    "ps5_sort_strings", "ps5_sort_strings_cache", "ps5_dist_chars_device",
    "ps5_ctx_create", "ps5_ctx_destroy", "ps5_ctx_reserve", "ps5_ctx_stream",
    "ps5_sort_device", "ps5_sort_device_noalloc", "ps5_noalloc_max_strings", "ps5_ctx_set_profile",
    "ps5_ctx_set_planning", "ps5_ctx_get_planning", "ps5_ctx_set_graph_rounds",
    "ps5_ctx_get_graph_rounds", "ps5_ctx_kernel_count", "ps5_ctx_kernel_stat",
    "ps5_ctx_work_stats",
    "ps5_dev_malloc", "ps5_dev_free", "ps5_memcpy_h2d", "ps5_memcpy_d2h", "ps5_device_synchronize",
    "ps5_dev_copy",
    "ps5_sort_device_timed",
    "ps5_gen_uniform", "ps5_gen_windows", "ps5_gen_reference_random", "ps5_gen_prefixed",
    "ps5_sort_batch_device", "ps5_gen_batch_plan", "ps5_gen_batch_fill",
    "ps5_server_create", "ps5_server_destroy", "ps5_server_submit", "ps5_server_wait", "ps5_server_stats",
    "ps5_server_health", "ps5_server_ledger", "ps5_server_inject_fault",
    "ps5_ctx_set_tuning", "ps5_ctx_get_tuning", "ps5_ctx_set_shape", "ps5_ctx_get_shape",
    "ps5_autotune_device", "ps5_autotune_batch_device",
    "ps5_merge_device", "ps5_lcp_merge_strings",
]


class Ps5Error(RuntimeError):
    pass


_lib = None
_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_U32 = ctypes.c_uint32
_I = ctypes.c_int


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


def load():
    """Load the library (raises Ps5Error if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise Ps5Error(f"{LIB_PATH} is missing: build it first (__graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    _sig(lib, "ps5_last_error", ctypes.c_char_p, [])
    _sig(lib, "mi355x_ps5_sort", None, [_P, ctypes.c_size_t])
    _sig(lib, "mi355x_ps5_sort_lcp", None, [_P, _P, ctypes.c_size_t])
    _sig(lib, "mi355x_ps5_sort_out", None, [_P, ctypes.c_size_t])
    _sig(lib, "mi355x_ps5_sort_out_base", None, [_P, _P, ctypes.c_size_t])
    _sig(lib, "ps5_sort_strings", _I, [_P, ctypes.c_size_t, _P, _P])
    _sig(lib, "mi355x_ps5_sort_lcp_cache", None, [_P, _P, _P, ctypes.c_size_t])
    _sig(lib, "ps5_sort_strings_cache", _I, [_P, ctypes.c_size_t, _P, _P, _P])
    _sig(lib, "ps5_dist_chars_device", _I, [_P, _P, _U64, _P, _U64, _P, _P, _P])
    _sig(lib, "ps5_ctx_create", _I, [ctypes.POINTER(_P), _I])
    _sig(lib, "ps5_ctx_destroy", None, [_P])
    _sig(lib, "ps5_ctx_reserve", _I, [_P, _U64])
    _sig(lib, "ps5_ctx_stream", _P, [_P])
    _sig(lib, "ps5_sort_device", _I, [_P, _P, _U64, _P, _U64, _P, _P])
    _sig(lib, "ps5_sort_device_noalloc", _I, [_P, _P, _U64, _P, _U64, _P, _P])
    _sig(lib, "ps5_noalloc_max_strings", _U64, [])
    _sig(lib, "ps5_sort_device_timed", _I, [_P, _P, _U64, _P, _U64, _P, _P, ctypes.POINTER(ctypes.c_double)])
    _sig(lib, "ps5_ctx_set_profile", _I, [_P, _I])
    _sig(lib, "ps5_ctx_set_planning", _I, [_P, _I])
    _sig(lib, "ps5_ctx_get_planning", _I, [_P])
    _sig(lib, "ps5_ctx_set_graph_rounds", _I, [_P, ctypes.c_uint32])
    _sig(lib, "ps5_ctx_get_graph_rounds", _I, [_P])
    _sig(lib, "ps5_ctx_kernel_count", _I, [_P])
    _sig(lib, "ps5_ctx_kernel_stat", _I, [_P, _I, ctypes.POINTER(ctypes.c_char_p),
                                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_U64),
                                          ctypes.POINTER(_U64)])
    _sig(lib, "ps5_ctx_work_stats", _I, [_P] + [ctypes.POINTER(_U64)] * 5)
    _sig(lib, "ps5_dev_malloc", _I, [ctypes.POINTER(_P), _U64])
    _sig(lib, "ps5_dev_free", _I, [_P])
    _sig(lib, "ps5_memcpy_h2d", _I, [_P, _P, _U64])
    _sig(lib, "ps5_memcpy_d2h", _I, [_P, _P, _U64])
    _sig(lib, "ps5_device_synchronize", _I, [])
    _sig(lib, "ps5_dev_copy", _I, [_P, _P, _U64])
    _sig(lib, "ps5_gen_uniform", _U64, [_U64, _U64, _U32, _U32, _P, _U32, _P, _P, _U64,
                                        ctypes.POINTER(_U64)])
    _sig(lib, "ps5_gen_windows", _I, [_U64, _U64, _U32, _U32, ctypes.c_double, _U64, _U32, _P, _P])
    _sig(lib, "ps5_gen_reference_random", _U64, [_U64, _P, _U32, _P, _P, _U64])
    _sig(lib, "ps5_gen_prefixed", _U64, [_U64, _U64, _U32, ctypes.c_double, _P, _P, _U64,
                                         ctypes.POINTER(_U64)])
    _sig(lib, "ps5_sort_batch_device", _I, [_P, _P, _U64, _P, _P, _U32, _P, _P])
    _sig(lib, "ps5_gen_batch_plan", _I, [_U64, _U32, _U64, _U64, _U32, _U32, _P, _P])
    _sig(lib, "ps5_gen_batch_fill", _I, [_U64, _U32, _U64, _U64, _U32, _U32, _P, _P])
    _sig(lib, "ps5_server_create", _I, [ctypes.POINTER(_P), _I, _U64, _U32, _U64])
    _sig(lib, "ps5_server_destroy", None, [_P])
    _sig(lib, "ps5_server_submit", _U64, [_P, _P, _U64, _P, _U64, _P, _P])
    _sig(lib, "ps5_server_wait", _I, [_P, _U64, ctypes.POINTER(ctypes.c_double)])
    _sig(lib, "ps5_server_health", _I, [_P, ctypes.POINTER(_U32), ctypes.POINTER(_U64)])
    _sig(lib, "ps5_server_inject_fault", _I, [_P, _U64])
    _sig(lib, "ps5_server_ledger", _I, [_P, ctypes.POINTER(_U64), ctypes.POINTER(_U64)])
    _sig(lib, "ps5_server_stats", _I, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(_U64), ctypes.POINTER(_U64), ctypes.POINTER(ctypes.c_double)])
    _sig(lib, "ps5_ctx_set_tuning", _I, [_P, _U32, _U32, _U32])
    _sig(lib, "ps5_ctx_get_tuning", _I, [_P] + [ctypes.POINTER(_U32)] * 3)
    _sig(lib, "ps5_ctx_set_shape", _I, [_P, _U32, _U32, _U32, _U32])
    _sig(lib, "ps5_ctx_get_shape", _I, [_P] + [ctypes.POINTER(_U32)] * 4)
    _sig(lib, "ps5_autotune_device", _I, [_P, _P, _U64, _P, _U64, _P, _P] + [ctypes.POINTER(_U32)] * 3
         + [ctypes.POINTER(ctypes.c_double)])
    _sig(lib, "ps5_autotune_batch_device", _I, [_P, _P, _U64, _P, _P, _U32, _P, _P] + [ctypes.POINTER(_U32)] * 3
         + [ctypes.POINTER(ctypes.c_double)])
    _sig(lib, "ps5_merge_device", _I, [_P, _P, _U64, _P, _P, _U32, _P, _P, _P, _P])
    _sig(lib, "ps5_lcp_merge_strings", _I, [_P, _P, _P, ctypes.c_uint, _P, _P])
    _lib = lib
    return lib


def _check(rc, what):
    if rc != PS5_OK:
        msg = load().ps5_last_error().decode(errors="replace")
        raise Ps5Error(f"{what} failed with status {rc}: {msg}")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------- inputs

ALPHA_DIGITS = b"0123456789"
ALPHA_ACGT = b"ACGT"
ALPHA_ASCII94 = bytes(range(33, 127))
ALPHA_ALNUM = b"0123456789ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz"


def gen_uniform(seed, arena_bytes, len_lo, len_hi, alphabet):
    """Strings with uniform lengths in [len_lo, len_hi] over `alphabet`,
    appended until `arena_bytes` (NULs included) is full.  Returns
    (arena uint8 with 16 zero bytes of padding, offs uint32)."""
    lib = load()
    alpha = np.frombuffer(bytes(alphabet), np.uint8).copy()
    arena = np.zeros(arena_bytes + 16, np.uint8)
    cap = arena_bytes // (len_lo + 1) + 1
    offs = np.zeros(cap, np.uint32)
    used = _U64(0)
    n = lib.ps5_gen_uniform(seed, arena_bytes, len_lo, len_hi, _ptr(alpha), len(alpha),
                            _ptr(arena), _ptr(offs), cap, ctypes.byref(used))
    return arena[: used.value + 16], offs[:n].copy()


def gen_windows(seed, text_len, nseg, seg_len, repeat_frac, n, win_len):
    lib = load()
    arena = np.zeros(n * (win_len + 1) + 16, np.uint8)
    offs = np.zeros(n, np.uint32)
    _check(lib.ps5_gen_windows(seed, text_len, nseg, seg_len, repeat_frac, n, win_len,
                               _ptr(arena), _ptr(offs)), "ps5_gen_windows")
    return arena, offs


def gen_reference_random(size, letters):
    """The reference's random2/4/10/62/255 inputs (src/tools/input.hpp:425-468)."""
    lib = load()
    let = np.frombuffer(bytes(letters), np.uint8).copy()
    arena = np.zeros(size + 9 + 16, np.uint8)
    cap = size // 16 + 2
    offs = np.zeros(cap, np.uint32)
    n = lib.ps5_gen_reference_random(size, _ptr(let), len(let), _ptr(arena), _ptr(offs), cap)
    return arena, offs[:n].copy()


def gen_prefixed(seed, arena_bytes, nprefix=1000, zipf_s=1.1):
    lib = load()
    arena = np.zeros(arena_bytes + 16, np.uint8)
    cap = arena_bytes // 19 + 1
    offs = np.zeros(cap, np.uint32)
    used = _U64(0)
    n = lib.ps5_gen_prefixed(seed, arena_bytes, nprefix, zipf_s, _ptr(arena), _ptr(offs), cap,
                             ctypes.byref(used))
    return arena, offs[:n].copy()


def gen_batch(seed, nsets, count_lo, count_hi, len_lo=4, len_hi=48):
    """Batch of independent sets (ps5_gen_batch_plan/fill): returns
    (arena uint8 with 16 bytes padding, offs uint64, set_first uint64[nsets+1])."""
    lib = load()
    counts = np.zeros(nsets, np.uint64)
    nbytes = np.zeros(nsets, np.uint64)
    _check(lib.ps5_gen_batch_plan(seed, nsets, count_lo, count_hi, len_lo, len_hi, _ptr(counts), _ptr(nbytes)),
           "ps5_gen_batch_plan")
    set_first = np.zeros(nsets + 1, np.uint64)
    set_first[1:] = np.cumsum(counts)
    arena = np.zeros(int(nbytes.sum()) + 16, np.uint8)
    offs = np.zeros(int(set_first[-1]), np.uint64)
    _check(lib.ps5_gen_batch_fill(seed, nsets, count_lo, count_hi, len_lo, len_hi, _ptr(arena), _ptr(offs)),
           "ps5_gen_batch_fill")
    return arena, offs, set_first


def workload(name, scale=1.0):
    """Named workloads; `scale` shrinks the byte/string budget for tests."""
    if name == "random10mb":        # 10 MiB, L in [8,32], '0'..'9'
        return gen_uniform(0x5EED0001, int(10 * 2**20 * scale), 8, 32, ALPHA_DIGITS)
    if name == "ascii2gb":          # 2 GiB, L in [1,64], bytes 33..126
        return gen_uniform(0x5EED0002, int(2 * 2**30 * scale), 1, 64, ALPHA_ASCII94)
    if name == "highlcp":           # 20M windows of 100 over a 2^29 ACGT text, 5% repeats
        n = max(1, int(20_000_000 * scale))
        text = max(4096, int(2**29 * scale))
        return gen_windows(0x5EED0003, text, 10, 2048, 0.05, n, 100)
    raise ValueError(name)


# ---------------------------------------------------------------- device API

class DeviceBuffer:
    """Device memory owned through the library (no HIP headers needed)."""

    def __init__(self, nbytes):
        lib = load()
        p = _P()
        _check(lib.ps5_dev_malloc(ctypes.byref(p), max(1, int(nbytes))), "ps5_dev_malloc")
        self.ptr = p
        self.nbytes = int(nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(load().ps5_memcpy_h2d(self.ptr, _ptr(arr), arr.nbytes), "ps5_memcpy_h2d")

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        _check(load().ps5_memcpy_d2h(_ptr(out), self.ptr, out.nbytes), "ps5_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            load().ps5_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """A sort context (stream + work buffers) on one device."""

    def __init__(self, device=0):
        lib = load()
        p = _P()
        _check(lib.ps5_ctx_create(ctypes.byref(p), device), "ps5_ctx_create")
        self.h = p

    def close(self):
        if self.h:
            load().ps5_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sort_device(self, d_arena, arena_bytes, d_offs, n, d_perm, d_lcp=None, timed=False):
        lib = load()
        lcp = d_lcp.ptr if d_lcp is not None else None
        if timed:
            ms = ctypes.c_double(0)
            _check(lib.ps5_sort_device_timed(self.h, d_arena.ptr, arena_bytes, d_offs.ptr, n, d_perm.ptr, lcp,
                                             ctypes.byref(ms)), "ps5_sort_device_timed")
            return ms.value
        _check(lib.ps5_sort_device(self.h, d_arena.ptr, arena_bytes, d_offs.ptr, n, d_perm.ptr, lcp),
               "ps5_sort_device")
        return None

    def sort(self, arena, offs, with_lcp=True):
        """Upload, sort on the device, download: returns (perm, lcp)."""
        n = len(offs)
        arena_bytes = int(arena.size)
        # the last byte of the arena must be a NUL (padding is zero)
        da = DeviceBuffer(arena_bytes)
        do = DeviceBuffer(max(4, n * 4))
        dp = DeviceBuffer(max(4, n * 4))
        dl = DeviceBuffer(max(4, n * 4)) if with_lcp else None
        try:
            da.upload(arena)
            do.upload(np.ascontiguousarray(offs, np.uint32))
            self.sort_device(da, arena_bytes, do, n, dp, dl)
            perm = dp.download(np.uint32, n)
            lcp = dl.download(np.uint32, n) if with_lcp else None
        finally:
            for b in (da, do, dp, dl):
                if b is not None:
                    b.free()
        return perm, lcp

    def dist_chars_device(self, d_arena, arena_bytes, d_offs, n, d_perm, d_lcp, d_cache):
        _check(load().ps5_dist_chars_device(self.h, d_arena.ptr, arena_bytes, d_offs.ptr, n, d_perm.ptr, d_lcp.ptr,
                                            d_cache.ptr), "ps5_dist_chars_device")

    def sort_cache(self, arena, offs):
        """Upload, sort, distinguishing characters, download: (perm, lcp, cache)."""
        n = len(offs)
        bufs = [DeviceBuffer(arena.size), DeviceBuffer(max(4, n * 4)), DeviceBuffer(max(4, n * 4)),
                DeviceBuffer(max(4, n * 4)), DeviceBuffer(max(1, n))]
        da, do, dp, dl, dc = bufs
        try:
            da.upload(arena)
            do.upload(np.ascontiguousarray(offs, np.uint32))
            dc.upload(np.full(max(1, n), 0xEE, np.uint8))
            self.sort_device(da, int(arena.size), do, n, dp, dl)
            self.dist_chars_device(da, int(arena.size), do, n, dp, dl, dc)
            return dp.download(np.uint32, n), dl.download(np.uint32, n), dc.download(np.uint8, n)
        finally:
            for b in bufs:
                b.free()

    def merge_device(self, d_arena, arena_bytes, d_offs, run_first, d_idx, d_lcp, d_idx_out, d_lcp_out):
        rf = np.ascontiguousarray(run_first, np.uint64)
        _check(load().ps5_merge_device(self.h, d_arena.ptr, arena_bytes, d_offs.ptr, _ptr(rf), len(rf) - 1,
                                       d_idx.ptr, d_lcp.ptr, d_idx_out.ptr, d_lcp_out.ptr), "ps5_merge_device")

    def merge(self, arena, offs, run_first, idx, lcp):
        """Merge sorted runs (idx/lcp arrays, run r at [run_first[r], run_first[r+1])): (idx, lcp)."""
        n = int(run_first[-1])
        bufs = [DeviceBuffer(arena.size)] + [DeviceBuffer(max(4, 4 * n)) for _ in range(5)]
        da, do, di, dl, dio, dlo = bufs
        try:
            da.upload(arena)
            do.upload(np.ascontiguousarray(offs, np.uint32))
            di.upload(np.ascontiguousarray(idx, np.uint32))
            dl.upload(np.ascontiguousarray(lcp, np.uint32))
            self.merge_device(da, int(arena.size), do, run_first, di, dl, dio, dlo)
            return dio.download(np.uint32, n), dlo.download(np.uint32, n)
        finally:
            for b in bufs:
                b.free()

    def sort_batch_device(self, d_arena, arena_bytes, d_offs64, set_first, d_perm, d_lcp=None):
        sf = np.ascontiguousarray(set_first, np.uint64)
        _check(load().ps5_sort_batch_device(self.h, d_arena.ptr, arena_bytes, d_offs64.ptr, _ptr(sf), len(sf) - 1,
                                            d_perm.ptr, d_lcp.ptr if d_lcp is not None else None),
               "ps5_sort_batch_device")

    def sort_batch(self, arena, offs64, set_first, with_lcp=True):
        """Upload a batch, sort every set, download: (perm, lcp) with per-set local indices."""
        n = int(set_first[-1])
        da = DeviceBuffer(arena.size)
        do = DeviceBuffer(max(8, n * 8))
        dp = DeviceBuffer(max(4, n * 4))
        dl = DeviceBuffer(max(4, n * 4)) if with_lcp else None
        try:
            da.upload(arena)
            do.upload(np.ascontiguousarray(offs64, np.uint64))
            self.sort_batch_device(da, int(arena.size), do, set_first, dp, dl)
            perm = dp.download(np.uint32, n)
            lcp = dl.download(np.uint32, n) if with_lcp else None
        finally:
            for b in (da, do, dp, dl):
                if b is not None:
                    b.free()
        return perm, lcp

    def set_tuning(self, leaf_target=768, max_treebits=10, max_oversample=32):
        _check(load().ps5_ctx_set_tuning(self.h, leaf_target, max_treebits, max_oversample), "ps5_ctx_set_tuning")

    def get_tuning(self):
        v = [_U32() for _ in range(3)]
        _check(load().ps5_ctx_get_tuning(self.h, *[ctypes.byref(x) for x in v]), "ps5_ctx_get_tuning")
        return tuple(x.value for x in v)

    def set_shape(self, chunk_max=9216, chunk_max_root=12032, min_tiles=512, leaf_keys=8):
        _check(load().ps5_ctx_set_shape(self.h, chunk_max, chunk_max_root, min_tiles, leaf_keys), "ps5_ctx_set_shape")

    def get_shape(self):
        """(chunk_max, chunk_max_root, min_tiles, leaf_keys)"""
        v = [_U32() for _ in range(4)]
        _check(load().ps5_ctx_get_shape(self.h, *[ctypes.byref(x) for x in v]), "ps5_ctx_get_shape")
        return tuple(x.value for x in v)

    def autotune_device(self, d_arena, arena_bytes, d_offs, n, d_perm, d_lcp=None):
        """Tune on this input; returns ((leaf_target, max_treebits, max_oversample), best ms)."""
        v = [_U32() for _ in range(3)]
        ms = ctypes.c_double()
        _check(load().ps5_autotune_device(self.h, d_arena.ptr, arena_bytes, d_offs.ptr, n, d_perm.ptr,
                                          d_lcp.ptr if d_lcp is not None else None,
                                          *[ctypes.byref(x) for x in v], ctypes.byref(ms)), "ps5_autotune_device")
        return tuple(x.value for x in v), ms.value

    def autotune_batch_device(self, d_arena, arena_bytes, d_offs64, set_first, d_perm, d_lcp=None):
        sf = np.ascontiguousarray(set_first, np.uint64)
        v = [_U32() for _ in range(3)]
        ms = ctypes.c_double()
        _check(load().ps5_autotune_batch_device(self.h, d_arena.ptr, arena_bytes, d_offs64.ptr, _ptr(sf), len(sf) - 1,
                                                d_perm.ptr, d_lcp.ptr if d_lcp is not None else None,
                                                *[ctypes.byref(x) for x in v], ctypes.byref(ms)),
               "ps5_autotune_batch_device")
        return tuple(x.value for x in v), ms.value

    def set_planning(self, mode):
        """0 (the default): levels planned on the host; 1: on the device (no synchronisation inside a
        round); 2: on the device for a single sort of fewer than 4M strings, else on the host."""
        _check(load().ps5_ctx_set_planning(self.h, int(mode)), "ps5_ctx_set_planning")

    def get_planning(self):
        return int(load().ps5_ctx_get_planning(self.h))

    def set_graph_rounds(self, rounds):
        """Capture a sort's first `rounds` rounds as one HIP graph, replayed while the shape repeats (0: off)."""
        _check(load().ps5_ctx_set_graph_rounds(self.h, int(rounds)), "ps5_ctx_set_graph_rounds")

    def get_graph_rounds(self):
        return int(load().ps5_ctx_get_graph_rounds(self.h))

    def set_profile(self, on):
        _check(load().ps5_ctx_set_profile(self.h, 1 if on else 0), "ps5_ctx_set_profile")

    def kernel_stats(self):
        lib = load()
        out = {}
        for i in range(lib.ps5_ctx_kernel_count(self.h)):
            name = ctypes.c_char_p()
            ms = ctypes.c_double()
            la = _U64()
            by = _U64()
            _check(lib.ps5_ctx_kernel_stat(self.h, i, ctypes.byref(name), ctypes.byref(ms), ctypes.byref(la),
                                           ctypes.byref(by)), "ps5_ctx_kernel_stat")
            out[name.value.decode()] = {"ms": ms.value, "launches": la.value, "bytes": by.value}
        return out

    def work_stats(self):
        vals = [_U64() for _ in range(5)]
        _check(load().ps5_ctx_work_stats(self.h, *[ctypes.byref(v) for v in vals]), "ps5_ctx_work_stats")
        return dict(zip(["rounds", "levels", "segments", "leaves", "syncs"], [v.value for v in vals]))


class Server:
    """Continuous-batching sort service (ps5_server_*)."""

    def __init__(self, device=0, hbm_budget=8 << 30, slots=4, max_batch_strings=0):
        p = _P()
        _check(load().ps5_server_create(ctypes.byref(p), device, hbm_budget, slots, max_batch_strings),
               "ps5_server_create")
        self.h = p
        self._keep = {}

    def submit(self, arena, offs, with_lcp=True):
        n = len(offs)
        offs = np.ascontiguousarray(offs, np.uint32)
        perm = np.zeros(max(1, n), np.uint32)
        lcp = np.zeros(max(1, n), np.uint32) if with_lcp else None
        rid = load().ps5_server_submit(self.h, _ptr(arena), int(arena.size), _ptr(offs), n, _ptr(perm),
                                       _ptr(lcp) if with_lcp else None)
        if rid == 0:
            raise Ps5Error(f"ps5_server_submit: {load().ps5_last_error().decode()}")
        self._keep[rid] = (arena, offs, perm, lcp, n)
        return rid

    def wait(self, rid):
        ms = ctypes.c_double()
        _check(load().ps5_server_wait(self.h, rid, ctypes.byref(ms)), "ps5_server_wait")
        arena, offs, perm, lcp, n = self._keep.pop(rid)
        return perm[:n], (lcp[:n] if lcp is not None else None), ms.value

    def health(self):
        live, failed = _U32(), _U64()
        _check(load().ps5_server_health(self.h, ctypes.byref(live), ctypes.byref(failed)), "ps5_server_health")
        return {"live_slots": live.value, "failed": failed.value}

    def ledger(self):
        """the device-side admission ledger: requests admitted, bytes in use"""
        adm, used = _U64(), _U64()
        _check(load().ps5_server_ledger(self.h, ctypes.byref(adm), ctypes.byref(used)), "ps5_server_ledger")
        return {"admitted": adm.value, "in_use_bytes": used.value}

    def inject_fault(self, rid):
        _check(load().ps5_server_inject_fault(self.h, rid), "ps5_server_inject_fault")

    def stats(self):
        p50, p99, tp = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        done, batches = _U64(), _U64()
        _check(load().ps5_server_stats(self.h, ctypes.byref(p50), ctypes.byref(p99), ctypes.byref(done),
                                       ctypes.byref(batches), ctypes.byref(tp)), "ps5_server_stats")
        return {"p50_ms": p50.value, "p99_ms": p99.value, "completed": done.value, "batches": batches.value,
                "throughput_gbs": tp.value}

    def close(self):
        if self.h:
            load().ps5_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sort_strings(strings, with_lcp=False, with_cache=False):
    """Sort a list of bytes objects through the Contest-shaped host entry
    (ps5_sort_strings_cache): returns (sorted list, order, lcp or None[, cache])."""
    lib = load()
    n = len(strings)
    bufs = [ctypes.create_string_buffer(s, len(s) + 1) for s in strings]
    arr = (ctypes.c_void_p * max(1, n))(*[ctypes.addressof(b) for b in bufs])
    with_lcp = with_lcp or with_cache
    lcp = np.full(max(1, n), 42, np.uint64) if with_lcp else None
    cache = np.full(max(1, n), 0xEE, np.uint8) if with_cache else None
    _check(lib.ps5_sort_strings_cache(arr, n, _ptr(lcp) if with_lcp else None,
                                      _ptr(cache) if with_cache else None, None), "ps5_sort_strings_cache")
    addr = {ctypes.addressof(b): i for i, b in enumerate(bufs)}
    order = [addr[arr[i]] for i in range(n)]
    out = [strings[i] for i in order], order, (lcp[:n] if with_lcp else None)
    return out + (cache[:n],) if with_cache else out


def sort_suffixes(text, with_lcp=True):
    """Sort every suffix of `text` (bytes) through the Contest-shaped host
    entry, with the pointers into one buffer as psstest --suffix lays them out:
    returns (order of suffix start positions, lcp or None)."""
    lib = load()
    n = len(text)
    buf = ctypes.create_string_buffer(bytes(text), n + 16)
    base = ctypes.addressof(buf)
    arr = (ctypes.c_void_p * max(1, n))(*[base + i for i in range(n)])
    lcp = np.full(max(1, n), 42, np.uint64) if with_lcp else None
    _check(lib.ps5_sort_strings(arr, n, _ptr(lcp) if with_lcp else None, None), "ps5_sort_strings")
    order = np.array([arr[i] - base for i in range(n)], np.int64)
    return order, (lcp[:n] if with_lcp else None)


def lcp_merge_strings(runs, lcps):
    """Host-form merge (ps5_lcp_merge_strings) of sorted runs of bytes objects
    with their LCP lists: returns (merged list, lcp array)."""
    lib = load()
    nr = len(runs)
    bufs = [[ctypes.create_string_buffer(x, len(x) + 1) for x in run] for run in runs]
    arrs = [(ctypes.c_void_p * max(1, len(r)))(*[ctypes.addressof(b) for b in r]) for r in bufs]
    lcp_arrs = [np.ascontiguousarray(np.asarray(l if len(l) else [0], np.uint64)) for l in lcps]
    run_ptrs = (ctypes.c_void_p * max(1, nr))(*[ctypes.addressof(a) for a in arrs])
    lcp_ptrs = (ctypes.c_void_p * max(1, nr))(*[a.ctypes.data for a in lcp_arrs])
    sizes = np.array([len(r) for r in runs], np.uint64)
    n = int(sizes.sum())
    out = (ctypes.c_void_p * max(1, n))()
    out_lcp = np.full(max(1, n), 42, np.uint64)
    _check(lib.ps5_lcp_merge_strings(run_ptrs, lcp_ptrs, _ptr(sizes), nr, out, _ptr(out_lcp)),
           "ps5_lcp_merge_strings")
    addr = {}
    for r, run in enumerate(bufs):
        for i, b in enumerate(run):
            addr[ctypes.addressof(b)] = runs[r][i]
    return [addr[out[i]] for i in range(n)], out_lcp[:n]
