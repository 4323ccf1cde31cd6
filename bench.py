#!/usr/bin/env python3
"""Benchmark of the MI355X pS5 string sort (one GPU).

A step is one complete sort (permutation + LCP array) of the 2 GiB workload:
strings with length uniform in [1,64] over bytes 33..126, seeded, arena and
offsets already resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ascii2gb|random10mb|highlcp]

Two further comparisons print their own JSON line when asked for:
  --mode batch   independent sorts in one batched pass vs one sort call per set
                 (default: 4096 sets, log-uniform 1k..2M strings, length U[4,48],
                 alphabets ACGT / digits / bytes 33..126, seed 0x5EED0004)
  --mode serve   Poisson arrivals served by the continuous-batching server vs a
                 FIFO of plain sorts (log-uniform 10 KB..256 MB requests of
                 prefix-heavy strings, seed 0x5EED0005)
  --mode merge   LCP merge of --shards sorted index ranges of --workload
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
METRIC = "sorted chars/s (GB/s) & % HBM roofline; batched vs plain sorts/s, p99 latency"


def load_ps5():
    spec = importlib.util.spec_from_file_location("ps5", os.path.join(ROOT, "parallel-string-sorting_amd", "ps5.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


WORKLOADS = {
    "ascii2gb": "2 GiB of strings, length uniform in [1,64], bytes 33..126 i.i.d., seed 0x5EED0002",
    "random10mb": "10 MiB of strings, length uniform in [8,32], '0'..'9' i.i.d., seed 0x5EED0001",
    "highlcp": "20M windows of 100 chars of a 2^29-base ACGT text with 5% 2 kb repeats, seed 0x5EED0003",
}


def host_info():
    """The host the CPU baseline runs on: CPU model, logical CPUs, and the
    cores this process may use (affinity mask, OMP_NUM_THREADS)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    usable = max(1, min(aff, int(omp))) if omp and omp.isdigit() else aff
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "affinity_cpus": aff,
            "omp_num_threads": omp, "threads_used": usable}


def cpu_baseline(ps5, workload, arena, offs, budget_s):
    """The oracle's pS5 restatement timed on the host cores this process may
    use (every CPU of its affinity mask, capped by OMP_NUM_THREADS: the GPU
    box grants 16 cores per GPU), on the same workload as the GPU step.
    Primary figure: with the LCP array (parallel_sample_sort_lcp_base), the
    output the GPU step produces; the plain contestant (no LCP) beside it.
    Each timed sort's own output is verified exactly afterwards."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ps5_oracle
    hi = host_info()
    threads = hi["threads_used"]
    n = len(offs)
    # a prefix of the workload when one run of the whole would take more than
    # a quarter of the budget (strings are laid out in input order)
    k = n
    t, perm, lcp = ps5_oracle.time_ps5(arena, offs, threads=threads, with_lcp=True, output=True)
    while t * 4 > budget_s and k > 1024:
        k = max(1024, int(k * (budget_s / 4) / t))
        t, perm, lcp = ps5_oracle.time_ps5(arena, offs[:k], threads=threads, with_lcp=True, output=True)
    sub = offs[:k]
    chars = int(offs[k]) if k < n else int(arena.size) - 16
    runs = [t]
    verified = ps5_oracle.verify(arena, sub, perm, lcp)[0] == 0
    while sum(runs) < budget_s / 2 and len(runs) < 5:
        t, perm, lcp = ps5_oracle.time_ps5(arena, sub, threads=threads, with_lcp=True, output=True)
        runs.append(t)
        verified = verified and ps5_oracle.verify(arena, sub, perm, lcp)[0] == 0
    nolcp = []
    while sum(nolcp) < budget_s / 4 and len(nolcp) < 3:
        t, perm, _ = ps5_oracle.time_ps5(arena, sub, threads=threads, with_lcp=False, output=True)
        nolcp.append(t)
        verified = verified and ps5_oracle.verify(arena, sub, perm)[0] == 0
    med = float(np.median(runs))
    whole = "the whole workload" if k == n else f"a prefix of {k} of its {n} strings"
    return {"value": round(chars / med / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"{workload}: {whole} ({chars} bytes), median of {len(runs)} runs of the oracle's pS5 "
                      f"restatement with the LCP array (treebits 10, oversample 2), {threads} OpenMP threads",
            "ms": round(med * 1e3, 3), "value_without_lcp": round(chars / float(np.median(nolcp)) / 1e9, 4),
            "output_verified": bool(verified), "host": hi}


def shape_of(ctx):
    """the context's tile shapes and leaf key width"""
    cm, cr, mt, lk = ctx.get_shape()
    return {"chunk_max": cm, "chunk_max_root": cr, "min_tiles": mt, "leaf_keys": lk}


def bench_batch(ps5, args):
    """Batched pass over many independent sets vs one device sort per set."""
    lib = ps5.load()
    t0 = time.time()
    arena, offs, sf = ps5.gen_batch(0x5EED0004, args.sets, args.count_lo, args.count_hi, 4, 48)
    n = int(sf[-1])
    print(f"[bench] batch: {args.sets} sets, {n} strings, {arena.size} bytes generated in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    ctx = ps5.Context(0)
    assert lib.ps5_ctx_reserve(ctx.h, n) == 0
    d_arena = ps5.DeviceBuffer(arena.size)
    d_offs = ps5.DeviceBuffer(n * 8)
    d_perm = ps5.DeviceBuffer(n * 4)
    d_lcp = ps5.DeviceBuffer(n * 4)
    d_arena.upload(arena)
    d_offs.upload(offs)
    tuning = ctx.get_tuning()
    autotune = args.autotune is not False        # on unless --no-autotune
    if autotune:
        t0 = time.time()
        tuning, tms = ctx.autotune_batch_device(d_arena, int(arena.size), d_offs, sf, d_perm, d_lcp)
        print(f"[bench] autotune: leaf_target/max_treebits/max_oversample = {tuning}, {tms:.3f} ms "
              f"({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        ctx.sort_batch_device(d_arena, int(arena.size), d_offs, sf, d_perm, d_lcp)
    lib.ps5_device_synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        ctx.sort_batch_device(d_arena, int(arena.size), d_offs, sf, d_perm, d_lcp)
    lib.ps5_device_synchronize()
    batched_s = (time.perf_counter() - t) / args.steps
    perm = d_perm.download(np.uint32, n)
    lcp = d_lcp.download(np.uint32, n)
    # the same batch as one captured HIP graph per call: the first rounds
    # planned on the device and replayed (the first call captures)
    ctx.set_planning(1)
    ctx.set_graph_rounds(args.graph_rounds)
    ctx.sort_batch_device(d_arena, int(arena.size), d_offs, sf, d_perm, d_lcp)
    lib.ps5_device_synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        ctx.sort_batch_device(d_arena, int(arena.size), d_offs, sf, d_perm, d_lcp)
    lib.ps5_device_synchronize()
    graph_s = (time.perf_counter() - t) / args.steps
    graph_work = ctx.work_stats()
    graph_same = bool(np.array_equal(d_perm.download(np.uint32, n), perm))
    gl = d_lcp.download(np.uint32, n)
    for k in range(args.sets):
        b, e = int(sf[k]), int(sf[k + 1])
        if e - b > 1:
            graph_same &= bool(np.array_equal(gl[b + 1:e], lcp[b + 1:e]))
    del gl
    ctx.set_graph_rounds(0)
    ctx.set_planning(0)     # back to the default (host planning) for the plain per-set sorts
    # plain path: one ps5_sort_device call per set (uint32 offsets local to the set)
    starts = np.array([int(offs[int(sf[k])]) for k in range(args.sets)], np.uint64)
    ends = np.append(starts[1:], np.uint64(arena.size - 16))      # sets are laid out back to back
    local = np.empty(n, np.uint32)
    for k in range(args.sets):
        b, e = int(sf[k]), int(sf[k + 1])
        local[b:e] = (offs[b:e] - starts[k]).astype(np.uint32)
    d_local = ps5.DeviceBuffer(max(4, n * 4))
    d_local.upload(local)
    pctx = ps5.Context(0)
    d_pp = ps5.DeviceBuffer(n * 4)
    d_pl = ps5.DeviceBuffer(n * 4)
    import ctypes
    P = ctypes.c_void_p
    ab, ob, pb, lb = d_arena.ptr.value, d_local.ptr.value, d_pp.ptr.value, d_pl.ptr.value

    def plain_pass():
        for k in range(args.sets):
            b, e = int(sf[k]), int(sf[k + 1])
            if e == b:
                continue
            rc = lib.ps5_sort_device(pctx.h, P(ab + int(starts[k])), int(ends[k] - starts[k]), P(ob + 4 * b),
                                     e - b, P(pb + 4 * b), P(lb + 4 * b))
            ps5._check(rc, "ps5_sort_device")
    plain_pass()
    lib.ps5_device_synchronize()
    t = time.perf_counter()
    plain_pass()
    lib.ps5_device_synchronize()
    plain_s = time.perf_counter() - t
    pperm = d_pp.download(np.uint32, n)
    plcp = d_pl.download(np.uint32, n)
    same = bool(np.array_equal(perm, pperm))
    for k in range(args.sets):
        b, e = int(sf[k]), int(sf[k + 1])
        if e - b > 1:
            same &= bool(np.array_equal(lcp[b + 1:e], plcp[b + 1:e]))
    # oracle check of every 16th set
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ps5_oracle
    bad = 0
    for k in range(0, args.sets, 16):
        b, e = int(sf[k]), int(sf[k + 1])
        if e > b and ps5_oracle.verify(arena, offs[b:e], perm[b:e], lcp[b:e])[0] != 0:
            bad += 1
    out = {"metric": "batched vs plain sorts/s", "sets": args.sets, "strings": n, "arena_bytes": int(arena.size),
           "batched_sorts_per_s": round(args.sets / batched_s, 2), "batched_ms": round(batched_s * 1e3, 3),
           "plain_sorts_per_s": round(args.sets / plain_s, 2), "plain_ms": round(plain_s * 1e3, 3),
           "speedup": round(plain_s / batched_s, 2), "batched_chars_gbs": round(arena.size / batched_s / 1e9, 2),
           "identical_to_plain": same, "oracle_checked_sets": (args.sets + 15) // 16, "oracle_failures": bad,
           "graph": {"rounds_captured": args.graph_rounds, "ms": round(graph_s * 1e3, 3),
                     "sorts_per_s": round(args.sets / graph_s, 2), "host_syncs_per_sort": graph_work["syncs"],
                     "identical_to_batched": graph_same,
                     "note": "levels planned on the device, the first rounds replayed as one captured HIP graph"},
           "tuning": {"leaf_target": tuning[0], "max_treebits": tuning[1], "max_oversample": tuning[2],
                      **shape_of(ctx), "autotuned": autotune},
           "config": {"workload": f"{args.sets} sets, strings log-uniform [{args.count_lo}, {args.count_hi}], "
                                  "length U[4,48], alphabet ACGT/digits/bytes 33..126, seed 0x5EED0004"}}
    return out


def calibration_copy(ps5):
    """Two plain 1 GiB device copies (k_copy16): the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    cb = 1 << 30
    src = ps5.DeviceBuffer(cb)
    dst = ps5.DeviceBuffer(cb)
    for _ in range(2):
        ps5._check(ps5.load().ps5_dev_copy(dst.ptr, src.ptr, cb), "ps5_dev_copy")
    src.free()
    dst.free()
    print(f"[bench] calibration copy: k_copy16 moved {cb} bytes read + {cb} bytes written per launch",
          file=sys.stderr, flush=True)


def bench_merge(ps5, args):
    """LCP merge of K sorted shards (contiguous index ranges) of a workload."""
    import ctypes
    lib = ps5.load()
    arena, offs = ps5.workload(args.workload)
    n = len(offs)
    K = args.shards
    bounds = np.linspace(0, n, K + 1).astype(np.uint64)
    ctx = ps5.Context(0)
    assert lib.ps5_ctx_reserve(ctx.h, n) == 0
    d_arena, d_offs = ps5.DeviceBuffer(arena.size), ps5.DeviceBuffer(4 * n)
    d_idx, d_lcp = ps5.DeviceBuffer(4 * n), ps5.DeviceBuffer(4 * n)
    d_io, d_lo = ps5.DeviceBuffer(4 * n), ps5.DeviceBuffer(4 * n)
    d_arena.upload(arena)
    d_offs.upload(offs)
    if args.calibrate:
        calibration_copy(ps5)
    P = ctypes.c_void_p
    t0 = time.perf_counter()
    for b, e in zip(bounds[:-1], bounds[1:]):      # each shard sorted on its own (not timed below)
        b, e = int(b), int(e)
        ps5._check(lib.ps5_sort_device(ctx.h, d_arena.ptr, int(arena.size), P(d_offs.ptr.value + 4 * b), e - b,
                                       P(d_idx.ptr.value + 4 * b), P(d_lcp.ptr.value + 4 * b)), "ps5_sort_device")
    lib.ps5_device_synchronize()
    shard_ms = (time.perf_counter() - t0) * 1e3
    idx = d_idx.download(np.uint32, n)
    for b, e in zip(bounds[:-1], bounds[1:]):      # shard-local indices -> global
        idx[int(b):int(e)] += np.uint32(int(b))
    d_idx.upload(idx)
    for _ in range(args.warmup):
        ctx.merge_device(d_arena, int(arena.size), d_offs, bounds, d_idx, d_lcp, d_io, d_lo)
    lib.ps5_device_synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        ctx.merge_device(d_arena, int(arena.size), d_offs, bounds, d_idx, d_lcp, d_io, d_lo)
    lib.ps5_device_synchronize()
    merge_ms = (time.perf_counter() - t) * 1e3 / args.steps
    mi, ml = d_io.download(np.uint32, n), d_lo.download(np.uint32, n)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ps5_oracle
    vr, vmsg = ps5_oracle.verify(arena, offs, mi, ml)
    chars = int(arena.size) - 16
    out = {"metric": "LCP merge of sorted shards, merged chars/s (GB/s)", "value": round(chars / merge_ms / 1e6, 3),
           "unit": "GB/s", "merge_ms": round(merge_ms, 3), "shards": K, "shard_sorts_ms": round(shard_ms, 3),
           "rounds": int(np.ceil(np.log2(K))) if K > 1 else 0, "verified": vr == 0,
           "config": {"workload": WORKLOADS[args.workload], "strings": n, "arena_bytes": chars,
                      "shards": "contiguous index ranges, each sorted by ps5_sort_device"}}
    if vr != 0:
        print(f"[bench] MERGE OUTPUT WRONG: {vmsg}", file=sys.stderr)
    print(json.dumps(out), flush=True)
    return 0 if vr == 0 else 1


def bench_serve(ps5, args):
    """Poisson arrivals: continuous-batching server vs FIFO of plain sorts."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(0x5EED0005)
    sizes = np.exp(rng.uniform(np.log(args.req_lo), np.log(args.req_hi), args.requests)).astype(np.int64)
    t0 = time.time()
    with ThreadPoolExecutor(16) as ex:
        reqs = list(ex.map(lambda i: ps5.gen_prefixed(0x5EED0005 * 1000 + i, int(sizes[i])), range(args.requests)))
    total = sum(int(a.size) for a, _ in reqs)
    print(f"[bench] serve: {args.requests} requests, {total} bytes generated in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    ctx = ps5.Context(0)
    max_b = max(int(a.size) for a, _ in reqs)
    max_n = max(len(o) for _, o in reqs)
    d_a, d_o = ps5.DeviceBuffer(max_b), ps5.DeviceBuffer(4 * max_n)
    d_p, d_l = ps5.DeviceBuffer(4 * max_n), ps5.DeviceBuffer(4 * max_n)

    def plain_sort(a, o):
        """the unbatched path: upload, one device sort, download (buffers reused)"""
        d_a.upload(a)
        d_o.upload(o)
        ctx.sort_device(d_a, int(a.size), d_o, len(o), d_p, d_l)
        return d_p.download(np.uint32, len(o)), d_l.download(np.uint32, len(o))

    for a, o in reqs[:3]:
        plain_sort(a, o)
    alone, alone_ms = [], []
    t = time.perf_counter()
    for a, o in reqs:
        t1 = time.perf_counter()
        alone.append(plain_sort(a, o))
        alone_ms.append((time.perf_counter() - t1) * 1e3)
    plain_busy = time.perf_counter() - t
    rate = args.load * args.requests / plain_busy           # requests per second offered
    gaps = rng.exponential(1.0 / rate, args.requests)
    arrivals = np.cumsum(gaps)

    def run_fifo():
        lat = []
        start = time.perf_counter()
        for i, (a, o) in enumerate(reqs):
            now = time.perf_counter() - start
            if now < arrivals[i]:
                time.sleep(arrivals[i] - now)
            plain_sort(a, o)
            lat.append((time.perf_counter() - start - arrivals[i]) * 1e3)
        return lat, time.perf_counter() - start

    def run_served():
        srv = ps5.Server(0, hbm_budget=args.budget_gb << 30, slots=args.slots, max_batch_strings=args.max_batch)
        ids = []
        start = time.perf_counter()
        for i, (a, o) in enumerate(reqs):
            now = time.perf_counter() - start
            if now < arrivals[i]:
                time.sleep(arrivals[i] - now)
            ids.append(srv.submit(a, o))
        res = [srv.wait(i) for i in ids]
        wall = time.perf_counter() - start
        st = srv.stats()
        srv.close()
        return res, st, wall

    fifo_lat, fifo_wall = run_fifo()
    res, st, served_wall = run_served()
    same = all(np.array_equal(p1, p2) and np.array_equal(l1[1:], l2[1:])
               for (p1, l1), (p2, l2, _) in zip(alone, res))
    served_lat = [ms for _, _, ms in res]
    big = sorted(range(args.requests), key=lambda i: -int(reqs[i][0].size))[:5]
    largest = [{"bytes": int(reqs[i][0].size), "alone_ms": round(alone_ms[i], 3), "served_ms": round(served_lat[i], 3),
                "fifo_ms": round(fifo_lat[i], 3)} for i in big]
    out = {"metric": "served vs plain p50/p99 latency (ms) and throughput", "requests": args.requests,
           "offered_rate_per_s": round(rate, 2), "offered_load": args.load, "bytes": total,
           "served": {"p50_ms": round(float(np.percentile(served_lat, 50)), 3),
                      "p99_ms": round(float(np.percentile(served_lat, 99)), 3),
                      "throughput_gbs": round(total / served_wall / 1e9, 3), "batches": st["batches"],
                      "slots": args.slots, "hbm_budget_gb": args.budget_gb, "max_batch_strings": args.max_batch},
           "plain_fifo": {"p50_ms": round(float(np.percentile(fifo_lat, 50)), 3),
                          "p99_ms": round(float(np.percentile(fifo_lat, 99)), 3),
                          "throughput_gbs": round(total / fifo_wall / 1e9, 3)},
           "largest_requests": largest,
           "identical_to_alone": bool(same),
           "config": {"workload": f"{args.requests} requests, size log-uniform [{args.req_lo}, {args.req_hi}] bytes, "
                                  "one of 1000 prefixes (Zipf 1.1) + 10..60 random chars, Poisson arrivals at "
                                  f"{args.load} of the plain path's capacity, seed 0x5EED0005",
                      "latency": "submit (host arrays) to results in host memory, incl. PCIe both ways"}}
    return out


def _kernel_base(name):
    """'void ps5::(anonymous namespace)::k_block<2>(...)' -> 'k_block'"""
    name = name.strip('"')
    name = name.split("(", 1)[0] if not name.startswith("void ") else name[5:].split("(", 1)[0]
    name = name.split("<", 1)[0]
    return name.rsplit("::", 1)[-1].strip()


def rocprof_summary(d):
    """Per-kernel table from a rocprofv3 kernel trace (--stats) and two PMC
    passes (FETCH_SIZE, WRITE_SIZE in KiB per dispatch), the counters scaled
    so that k_copy16 (1 GiB read + 1 GiB written) reads 1 GiB each."""
    import collections
    import csv
    import glob

    def one(pattern):
        hits = sorted(glob.glob(os.path.join(d, pattern), recursive=True))
        if not hits:
            raise SystemExit(f"no {pattern} under {d}")
        return hits[0]

    stats = collections.defaultdict(lambda: [0, 0.0])
    for r in csv.DictReader(open(one("**/*kernel_stats.csv"))):
        k = _kernel_base(r["Name"])
        stats[k][0] += int(r["Calls"])
        stats[k][1] += float(r["TotalDurationNs"])

    def counters(pattern, cname):
        per = collections.defaultdict(list)
        for r in csv.DictReader(open(one(pattern))):
            if r["Counter_Name"] == cname:
                per[_kernel_base(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        return per
    fe = counters("**/fetch*counter_collection.csv", "FETCH_SIZE")
    wr = counters("**/write*counter_collection.csv", "WRITE_SIZE")
    gib_kib = float(1 << 20)
    # the 1 GiB calibration copies are the largest k_copy16 dispatches
    ff = gib_kib / max(fe["k_copy16"])
    fw = gib_kib / max(wr["k_copy16"])
    lines = [f"calibration: k_copy16 moves 1 GiB in + 1 GiB out; FETCH_SIZE x{ff:.3f}, WRITE_SIZE x{fw:.3f}",
             "",
             "| kernel | dispatches | avg us | total ms | fetch MB/dispatch | write MB/dispatch | measured GB/s |",
             "|---|---|---|---|---|---|---|"]
    for k, (calls, tot_ns) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        avg_us = tot_ns / calls / 1e3
        f = float(np.mean(fe[k])) * ff * 1024 / 1e6 if k in fe else 0.0
        w = float(np.mean(wr[k])) * fw * 1024 / 1e6 if k in wr else 0.0
        gbs = (f + w) * 1e6 / (avg_us * 1e-6) / 1e9 if avg_us > 0 else 0.0
        lines.append(f"| {k} | {calls} | {avg_us:.1f} | {tot_ns / 1e6:.3f} | {f:.1f} | {w:.1f} | {gbs:.1f} |")
    # SQ counters (one pass of its own), per dispatch, averaged
    sq = sorted(glob.glob(os.path.join(d, "**", "sq*counter_collection.csv"), recursive=True))
    if sq:
        per = collections.defaultdict(lambda: collections.defaultdict(list))
        for r in csv.DictReader(open(sq[0])):
            per[_kernel_base(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
        names = ["SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT", "SQ_WAIT_ANY",
                 "SQ_WAIT_INST_LDS", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES"]
        lines += ["", "SQ counters per dispatch (mean; WAIT_* as a fraction of WAVE_CYCLES)", "",
                  "| kernel | VALU | SALU | LDS | LDS bank conflicts | wait any | wait LDS issue |",
                  "|---|---|---|---|---|---|---|"]
        for k, _ in sorted(stats.items(), key=lambda kv: -kv[1][1]):
            if k not in per or not all(n in per[k] for n in names):
                continue
            m = {n: float(np.mean(per[k][n])) for n in names}
            wc = max(m["SQ_WAVE_CYCLES"], 1.0)
            lines.append(f"| {k} | {m['SQ_INSTS_VALU'] / 1e6:.1f}M | {m['SQ_INSTS_SALU'] / 1e6:.1f}M | "
                         f"{m['SQ_INSTS_LDS'] / 1e6:.1f}M | {m['SQ_LDS_BANK_CONFLICT'] / 1e6:.1f}M | "
                         f"{m['SQ_WAIT_ANY'] / wc:.2f} | {m['SQ_WAIT_INST_LDS'] / wc:.2f} |")
    # dispatch timeline of the last complete sort: a sort starts with k_setup,
    # or with k_root_fetch (a single root planned on the host)
    tr = sorted(glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True))
    if tr:
        rows = sorted(csv.DictReader(open(tr[0])), key=lambda r: int(r["Start_Timestamp"]))
        starts = [i for i, r in enumerate(rows) if _kernel_base(r["Kernel_Name"]) in ("k_setup", "k_root_fetch")]
        if len(starts) >= 2:
            seq = rows[starts[-2]:starts[-1]]
            t0 = int(seq[0]["Start_Timestamp"])
            lines += ["", "Dispatch timeline of one sort (us from its first dispatch)", "",
                      "| kernel | start | duration | grid |", "|---|---|---|---|"]
            for r in seq:
                name = r["Kernel_Name"].strip('"')
                short = _kernel_base(name) + ("<" + name.split("<", 1)[1].split(">", 1)[0] + ">" if "<" in name else "")
                a, b = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
                lines.append(f"| {short} | {(a - t0) / 1e3:.1f} | {(b - a) / 1e3:.1f} | {r['Grid_Size_X']} |")
            lines.append(f"| (end) | {(int(seq[-1]['End_Timestamp']) - t0) / 1e3:.1f} | | |")
    print("\n".join(lines))


def profiled_traffic(kernel, workload):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 summary
    of this workload (profiles/<workload>_kernels.md), or None."""
    path = os.path.join(ROOT, "profiles", f"{workload}_kernels.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 7 and cells[0] == kernel:
            return int((float(cells[4]) + float(cells[5])) * 1e6)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rocprof-summary", metavar="DIR",
                    help="print the per-kernel table of a rocprofv3 trace + FETCH_SIZE + WRITE_SIZE run and exit")
    ap.add_argument("--mode", default="plain", choices=["plain", "batch", "serve", "merge"])
    ap.add_argument("--shards", type=int, default=8, help="sorted shards merged in --mode merge")
    ap.add_argument("--sets", type=int, default=4096)
    ap.add_argument("--count-lo", type=int, default=1000)
    ap.add_argument("--count-hi", type=int, default=2_000_000)
    ap.add_argument("--graph-rounds", type=int, default=7,
                    help="rounds of the batch captured in one HIP graph (strings of at most 48 characters: 7)")
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--req-lo", type=int, default=10_000)
    ap.add_argument("--req-hi", type=int, default=256 << 20)
    ap.add_argument("--load", type=float, default=0.7)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--budget-gb", type=int, default=64)
    ap.add_argument("--max-batch", type=int, default=16 << 20, help="strings per served batch (buffers pre-allocated)")
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=None,
                    help="tune the sort's shape on the workload before timing (default: on for the batch, "
                         "off for single sorts)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="ascii2gb", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--extras", action=argparse.BooleanOptionalAction, default=True,
                    help="also measure the 10 MiB and high-LCP sorts, the batch and the serving comparison "
                         "(default on; --no-extras for the headline alone)")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--profile", action="store_true", help="print the per-kernel breakdown to stderr")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a plain 1 GiB device copy (k_copy16) to scale the HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    if args.rocprof_summary:
        return rocprof_summary(args.rocprof_summary)
    ps5 = load_ps5()
    if args.mode == "batch":
        print(json.dumps(bench_batch(ps5, args)), flush=True)
        return 0
    if args.mode == "serve":
        print(json.dumps(bench_serve(ps5, args)), flush=True)
        return 0
    if args.mode == "merge":
        return bench_merge(ps5, args)
    t_begin = time.time()
    head = measure_plain(ps5, args, args.workload, args.steps, args.warmup, kernels=True)
    out = {
        "metric": METRIC,
        "value": head["gbs"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": head["config"],
        "roofline": head["roofline"],
        "cpu_baseline": head.get("cpu_baseline"),
        "output_verified": head["verified"],
    }
    if not head["verified"]:
        print(json.dumps(out), flush=True)
        return 1
    ok = True
    if args.extras:
        # the other single-GPU configurations, each with its CPU baseline
        configs = {}
        for w in [x for x in ("random10mb", "highlcp") if x != args.workload]:
            r = measure_plain(ps5, args, w, max(args.steps, 10), args.warmup, kernels=False)
            ok = ok and r["verified"]
            configs[w] = {"value": r["gbs"], "unit": "GB/s", "ms_per_step": r["ms_per_step"],
                          "device_ms": r["device_ms"], "output_verified": r["verified"],
                          "config": r["config"], "work": r["work"], "cpu_baseline": r.get("cpu_baseline")}
            print(f"[bench] {w}: {configs[w]}", file=sys.stderr, flush=True)
        out["configs"] = configs
        # batched vs plain sorts/s, then served vs plain p50/p99
        for name, fn in (("batch", bench_batch), ("serve", bench_serve)):
            t0 = time.time()
            try:
                out[name] = fn(ps5, args)
            except Exception as e:       # report, keep the headline line
                out[name] = {"error": f"{type(e).__name__}: {e}"}
                ok = False
            print(f"[bench] {name} ({time.time() - t0:.1f}s): {out[name]}", file=sys.stderr, flush=True)
    out["bench_wall_s"] = round(time.time() - t_begin, 1)
    print(json.dumps(out), flush=True)
    return 0 if ok else 1


def measure_plain(ps5, args, workload, steps, warmup, kernels):
    """One device-resident sort per step of `workload`: times `steps` sorts
    after `warmup`, verifies the timed steps' own output exactly, then
    (kernels=True) one extra profiled sort for the per-kernel roofline, and
    the CPU baseline on the same input."""
    t0 = time.time()
    arena, offs = ps5.workload(workload)
    n = len(offs)
    chars = int(arena.size) - 16          # arena bytes including NULs (16 bytes of padding excluded)
    print(f"[bench] generated {workload}: {n} strings, {chars} bytes in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    lib = ps5.load()
    ctx = ps5.Context(0)
    assert lib.ps5_ctx_reserve(ctx.h, n) == 0
    d_arena = ps5.DeviceBuffer(arena.size)
    d_offs = ps5.DeviceBuffer(n * 4)
    d_perm = ps5.DeviceBuffer(n * 4)
    d_lcp = ps5.DeviceBuffer(n * 4)
    d_arena.upload(arena)
    d_offs.upload(offs)
    alen = int(arena.size)
    if args.calibrate and kernels:
        calibration_copy(ps5)
    tuning = ctx.get_tuning()
    if args.autotune:
        t0 = time.time()
        tuning, tms = ctx.autotune_device(d_arena, alen, d_offs, n, d_perm, d_lcp)
        print(f"[bench] autotune: leaf_target/max_treebits/max_oversample = {tuning}, {tms:.3f} ms "
              f"({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    for _ in range(warmup):
        ctx.sort_device(d_arena, alen, d_offs, n, d_perm, d_lcp)
    step_ms = []
    lib.ps5_device_synchronize()
    # poison the outputs: what is verified below was written by the timed
    # steps (uploaded from the host, so that no copy kernel joins a profile)
    poison = np.full(n, 0xFFFFFFFF, np.uint32)
    d_perm.upload(poison)
    d_lcp.upload(poison)
    del poison
    lib.ps5_device_synchronize()
    t_start = time.perf_counter()
    for _ in range(steps):
        step_ms.append(ctx.sort_device(d_arena, alen, d_offs, n, d_perm, d_lcp, timed=True))
    lib.ps5_device_synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = wall * 1e3 / steps
    work = ctx.work_stats()
    print(f"[bench] {workload} {steps} steps: {ms_per_step:.3f} ms/step wall, device {np.median(step_ms):.3f} ms "
          f"median, work {work}", file=sys.stderr, flush=True)
    # exact check of the last timed step's output (the oracle's verifier:
    # permutation, order, tie order, every LCP)
    perm = d_perm.download(np.uint32, n)
    lcp = d_lcp.download(np.uint32, n)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ps5_oracle
    vr, vmsg = ps5_oracle.verify(arena, offs, perm, lcp)
    if vr != 0:
        print(f"[bench] OUTPUT WRONG ({workload}): {vmsg}", file=sys.stderr, flush=True)
    else:
        print(f"[bench] {workload}: timed output verified (permutation, order, tie order, LCP)",
              file=sys.stderr, flush=True)
    res = {"n": n, "chars": chars, "ms_per_step": round(ms_per_step, 4),
           "gbs": round(chars / (wall / steps) / 1e9, 3), "device_ms": round(float(np.median(step_ms)), 4),
           "verified": vr == 0, "work": work,
           "config": {"workload": WORKLOADS[workload], "strings": n, "arena_bytes": chars,
                      "output": "permutation + LCP array", "device_resident": True,
                      "tuning": {"leaf_target": tuning[0], "max_treebits": tuning[1], "max_oversample": tuning[2],
                                 **shape_of(ctx), "autotuned": bool(args.autotune)}}}
    if kernels:
        res["roofline"] = kernel_roofline_of(ctx, args, d_arena, alen, d_offs, n, d_perm, d_lcp, workload,
                                             float(np.median(step_ms)))
    ctx.close()
    del d_arena, d_offs, d_perm, d_lcp
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(ps5, workload, arena, offs, args.cpu_budget)
        print(f"[bench] cpu baseline {workload}: {res['cpu_baseline']}", file=sys.stderr, flush=True)
    return res


def kernel_roofline_of(ctx, args, d_arena, alen, d_offs, n, d_perm, d_lcp, workload, sort_ms):
    """Per-kernel breakdown of one extra sort, each launch bracketed by HIP
    events on the context's stream: the dominant kernel's algorithmic bytes
    per launch over its average launch time, against the HBM peak."""
    ctx.set_profile(True)
    ctx.sort_device(d_arena, alen, d_offs, n, d_perm, d_lcp)
    ks = ctx.kernel_stats()
    ctx.set_profile(False)
    ranked = sorted((k for k in ks if ks[k]["launches"]), key=lambda k: -ks[k]["ms"])
    dom = ranked[0]
    d = ks[dom]
    avg_ms = d["ms"] / d["launches"]
    bytes_per_launch = d["bytes"] / d["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9

    def kernel_roofline(k):     # the same figures for another kernel
        v = ks[k]
        a_ms = v["ms"] / v["launches"]
        a = v["bytes"] / v["launches"] / (a_ms * 1e-3) / 1e9
        return {"kernel": k, "achieved": round(a, 2), "frac": round(a / HBM_PEAK_GBS, 4),
                "traffic": profiled_traffic(k, workload), "avg_launch_ms": round(a_ms, 4),
                "alg_bytes_per_launch": int(v["bytes"] / v["launches"])}
    sort_bytes = sum(v["bytes"] for v in ks.values())       # every pass of one sort
    if args.profile:
        tot = sum(v["ms"] for v in ks.values())
        for k, v in sorted(ks.items(), key=lambda kv: -kv[1]["ms"]):
            if v["launches"]:
                gbs = v["bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 else 0
                print(f"[profile] {k:14s} {v['ms']:9.3f} ms {100 * v['ms'] / tot:5.1f}%  launches {v['launches']:6d} "
                      f" alg {v['bytes'] / 1e9:8.3f} GB  {gbs:8.1f} GB/s", file=sys.stderr)
    return {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(dom, workload),
            "avg_launch_ms": round(avg_ms, 4), "alg_bytes_per_launch": int(bytes_per_launch),
            "runner_up": kernel_roofline(ranked[1]) if len(ranked) > 1 else None,
            "whole_sort": {"alg_bytes": int(sort_bytes), "device_ms": round(sort_ms, 4),
                           "achieved": round(sort_bytes / (sort_ms * 1e-3) / 1e9, 2),
                           "frac": round(sort_bytes / (sort_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}


if __name__ == "__main__":
    sys.exit(main())
