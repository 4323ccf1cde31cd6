/*
 * ps5_mi355x.h -- C ABI of the MI355X pS5 string sample sort.
 *
 * Reference interface being replaced (paths relative to the reference
 * checkout, bingmann/parallel-string-sorting):
 *   - Contest plug-in signatures, src/tools/contest.hpp:88-91
 *       typedef void (*func_type)(unsigned char** strings, size_t n);
 *       typedef void (*func_lcp_type)(unsigned char** strings, uintptr_t* lcp, size_t n);
 *   - pS5-Unroll contestants, src/parallel/bingmann-parallel_sample_sort.cpp:31-42
 *     (parallel_sample_sortBTCUI, formerly parallel_sample_sortBTCU2) and
 *     :97-108 (parallel_sample_sortBTCUI_lcp, which discards its LCPs); the
 *     LCP-returning entry is parallel_sample_sort_lcp_base,
 *     src/parallel/bingmann-parallel_sample_sort.hpp:1733-1748.
 *
 * Output contract (all entry points): strings are ordered by unsigned
 * byte-wise comparison up to the NUL terminator; byte-identical strings keep
 * their input order; lcp[i] (i >= 1) is the length of the longest common
 * prefix of output strings i-1 and i; lcp[0] is never written.
 *
 * No HIP or torch types appear here: device pointers are plain pointers and
 * streams are passed as void* (a hipStream_t, or NULL for the context's own).
 */
#ifndef PS5_MI355X_H
#define PS5_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    PS5_OK = 0,
    PS5_ERR_ARG = 1,        /* invalid argument (size limits, NULL, alignment) */
    PS5_ERR_HIP = 2,        /* a HIP runtime call failed */
    PS5_ERR_CAPACITY = 3,   /* an internal work list overflowed */
    PS5_ERR_INTERNAL = 4,   /* a device-side invariant check failed */
    PS5_ERR_NOMEM = 5       /* device allocation failed / over budget */
};

/* Message for the last error on the calling thread ("" if none). */
const char* ps5_last_error(void);

/* ---- Contest drop-ins (host pointers) ---------------------------------- */

/* Replaces parallel_sample_sortBTCUI (bingmann-parallel_sample_sort.cpp:31-37),
 * Contest func_type (contest.hpp:88).  Permutes strings[] in place.  Like the
 * reference it returns void: on failure it prints the error and abort()s. */
void mi355x_ps5_sort(unsigned char** strings, size_t n);

/* Replaces parallel_sample_sort_lcp_base (bingmann-parallel_sample_sort.hpp:
 * 1733-1748) behind the Contest func_lcp_type signature (contest.hpp:89):
 * sorts and fills lcp[1..n-1]; lcp[0] untouched.  abort()s on failure. */
void mi355x_ps5_sort_lcp(unsigned char** strings, uintptr_t* lcp, size_t n);

/* Replaces parallel_sample_sortBTCUI_out (bingmann-parallel_sample_sort.cpp:
 * 44-62), Contest func_type (contest.hpp:88): sorts into a separate output
 * array, then copies it back into strings[] (as that contestant does, for
 * psstest's checks).  abort()s on failure. */
void mi355x_ps5_sort_out(unsigned char** strings, size_t n);

/* Replaces parallel_sample_sort_out_base (bingmann-parallel_sample_sort.hpp:
 * 1704-1714): the sorted pointers go to output[0..n), strings[] is left as it
 * was.  abort()s on failure. */
void mi355x_ps5_sort_out_base(unsigned char** strings, unsigned char** output, size_t n);

/* Behind the Contest func_lcp_cache_type signature (contest.hpp:90-91): sorts,
 * fills lcp[1..n-1] and the distinguishing characters cache[i] = s_i[lcp[i]]
 * for i >= 1 (0 for equal neighbours), the output of
 * StringShadowLcpCacheOutPtr (stringptr.hpp:860-969) checked by
 * verify_lcp_cache (stringptr.hpp:977-1031).  lcp[0], cache[0] untouched.
 * abort()s on failure. */
void mi355x_ps5_sort_lcp_cache(unsigned char** strings, uintptr_t* lcp, unsigned char* cache, size_t n);

/* Host form of the merge, in place of the reference's
 * parallelLcpMerge(const LcpCacheStringPtr* input, unsigned numInputs,
 * string* output, size_t length) (eberle-parallel-lcp-merge-lcp-splitting.hpp:
 * 443-469): runs[r][0 .. sizes[r]) are sorted strings with LCP arrays lcps[r]
 * (lcps[r][0] ignored); the merged strings go to output and, if out_lcp is
 * not NULL, their LCPs to out_lcp[1..] (out_lcp[0] untouched).  Returns a
 * status code. */
int ps5_lcp_merge_strings(unsigned char* const* const* runs, const uintptr_t* const* lcps, const size_t* sizes,
                          unsigned nruns, unsigned char** output, uintptr_t* out_lcp);

/* Status-returning form of the three above: lcp and output may be NULL. */
int ps5_sort_strings(unsigned char** strings, size_t n, uintptr_t* lcp,
                     unsigned char** output);
/* ... and with the distinguishing characters (cache requires lcp). */
int ps5_sort_strings_cache(unsigned char** strings, size_t n, uintptr_t* lcp,
                           unsigned char* cache, unsigned char** output);

/* ---- device-resident interface ----------------------------------------- */

typedef struct ps5_ctx ps5_ctx;

/* A context owns a HIP stream and the device work buffers of one sort at a
 * time; buffers grow on demand and are kept between calls. */
int ps5_ctx_create(ps5_ctx** out, int device);
void ps5_ctx_destroy(ps5_ctx* ctx);
/* Pre-size the work buffers for up to n strings (optional). */
int ps5_ctx_reserve(ps5_ctx* ctx, uint64_t n);
/* The context's stream as a hipStream_t cast to void*. */
void* ps5_ctx_stream(ps5_ctx* ctx);

/* Sort n strings that live in device memory.
 *   d_arena : characters; 8-byte aligned; every string NUL-terminated; the
 *             byte at d_arena[arena_bytes-1] must be a NUL.  Never written.
 *   d_offs  : n uint32 start offsets into the arena (arena_bytes < 2^32).
 *   d_perm  : n uint32 out: input index of the i-th smallest string.
 *   d_lcp   : n uint32 out (may be NULL): lcp[i] for i >= 1, lcp[0] untouched.
 * Runs on the context's stream; returns after the result is complete. */
int ps5_sort_device(ps5_ctx* ctx, const uint8_t* d_arena, uint64_t arena_bytes,
                    const uint32_t* d_offs, uint64_t n,
                    uint32_t* d_perm, uint32_t* d_lcp);

/* ps5_sort_device without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Arguments
 * and results are those of ps5_sort_device; the work buffers come from a
 * static device array of 96 MiB that every GPU the library runs on holds
 * (one copy per GPU and process) until the process ends, so n is limited to
 * ps5_noalloc_max_strings() (2^20); a larger n is rejected with
 * PS5_ERR_ARG.  The call makes no hipMalloc/hipFree or other allocating call
 * (the context, created beforehand, holds the small pinned host buffers it
 * uses).  All calls on one GPU share that array and its contents persist
 * from call to call; two calls that overlapped on one GPU would race on it,
 * so the library serialises them per device (a call returns only when its
 * sort has finished). */
int ps5_sort_device_noalloc(ps5_ctx* ctx, const uint8_t* d_arena, uint64_t arena_bytes,
                            const uint32_t* d_offs, uint64_t n,
                            uint32_t* d_perm, uint32_t* d_lcp);
/* Largest n ps5_sort_device_noalloc accepts. */
uint64_t ps5_noalloc_max_strings(void);

/* Where a sort's levels are planned.  0 (host): after every level the large
 * segments are read back and the next level is planned on the host, with
 * grids sized exactly.  1 (device): a kernel plans each level from the
 * device's own lists and every kernel loops over the work it finds, so a
 * round of levels runs without any host synchronisation (one per round, to
 * learn whether another round is needed).  2 (auto): device planning for a
 * single sort of fewer than 4M strings, host planning for larger sorts,
 * batches and profiled sorts.  The default is 0.  All modes give the same
 * results.  No counterpart in the reference (its levels are recursion on
 * the CPU).  The environment variable PS5_DEV_LEVELS=0/1/2 sets a new
 * context's mode.  Returns PS5_ERR_ARG for any other mode. */
int ps5_ctx_set_planning(ps5_ctx* ctx, int mode);
/* The current planning mode (0, 1 or 2), -1 for a NULL context. */
int ps5_ctx_get_planning(ps5_ctx* ctx);

/* Capture the first `rounds` rounds of a sort (8 characters of depth each)
 * as one HIP graph, planned on the device, and replay it for later sorts with
 * the same buffers, string count, set boundaries and settings (0, the
 * default: off; at most 64).  A sort then needs one host synchronisation for
 * those rounds; strings that still tie after them continue round by round,
 * and a captured round that needs more sample-sort levels than the graph
 * holds is sorted again with host-planned levels, so results are always those
 * of ps5_sort_device.  The first sort of a shape pays the capture.  No
 * counterpart in the reference; for callers that sort the same shape
 * repeatedly (batches of fixed-size sets).  Not used by
 * ps5_sort_device_noalloc, since instantiating a graph allocates. */
int ps5_ctx_set_graph_rounds(ps5_ctx* ctx, uint32_t rounds);
/* The captured round count, -1 for a NULL context. */
int ps5_ctx_get_graph_rounds(ps5_ctx* ctx);

/* Distinguishing characters of a finished sort (inputs as ps5_sort_device,
 * d_perm and d_lcp its results): d_cache[i] = s_i[lcp[i]] for i >= 1, the
 * first character of the i-th smallest string that its predecessor does not
 * share, 0 when the two are equal; d_cache[0] untouched.  Replaces the
 * character cache of StringShadowLcpCacheOutPtr (src/tools/stringptr.hpp:860-969,
 * set by sample_sort_lcp, src/parallel/bingmann-parallel_sample_sort.hpp:366,385). */
int ps5_dist_chars_device(ps5_ctx* ctx, const uint8_t* d_arena, uint64_t arena_bytes,
                          const uint32_t* d_offs, uint64_t n,
                          const uint32_t* d_perm, const uint32_t* d_lcp, uint8_t* d_cache);

/* Merge nruns sorted runs into one sorted sequence with its LCP array, on
 * the device: the binary LCP merge of bingmann-lcp_mergesort_binary.hpp
 * (lcp_merge_binary / lcp_compare), applied pairwise in rounds with merge
 * path partitioning (the reference merges up to 8 streams at a time with an
 * LCP loser tree, eberle-parallel-lcp-merge*.hpp, for its NUMA sort,
 * eberle-ps5-parallel-toplevel-merge.hpp:55-176).
 *   d_arena, arena_bytes, d_offs: the strings, as for ps5_sort_device;
 *   run_first: host array of nruns+1 positions, run_first[0] = 0; run r is
 *     d_idx[run_first[r] .. run_first[r+1]), string indices in sorted order,
 *     with d_lcp[q] = LCP of the strings at q-1 and q (the first of a run
 *     unused) -- exactly what ps5_sort_device returns for a subset;
 *   d_idx_out, d_lcp_out: the merged order and its LCPs (d_lcp_out[0] unused).
 * Equal strings keep run order, so runs that are sorted index ranges in
 * ascending order merge to the same result as sorting everything at once. */
int ps5_merge_device(ps5_ctx* ctx, const uint8_t* d_arena, uint64_t arena_bytes,
                     const uint32_t* d_offs, const uint64_t* run_first, uint32_t nruns,
                     const uint32_t* d_idx, const uint32_t* d_lcp,
                     uint32_t* d_idx_out, uint32_t* d_lcp_out);

/* Batch: sort nsets independent string sets in one pass over the device.
 * Set s is the strings with input indices [set_first[s], set_first[s+1])
 * (set_first: host array of nsets+1 non-decreasing values, set_first[0] = 0);
 * d_offs holds uint64 offsets into one arena (which may exceed 4 GiB).
 * Results per set, at the set's own positions [set_first[s], set_first[s+1]):
 * d_perm = index inside the set of the i-th smallest string of the set,
 * d_lcp (may be NULL) = LCP with its predecessor inside the set; the first
 * LCP slot of every set is never written. */
int ps5_sort_batch_device(ps5_ctx* ctx, const uint8_t* d_arena, uint64_t arena_bytes,
                          const uint64_t* d_offs, const uint64_t* set_first, uint32_t nsets,
                          uint32_t* d_perm, uint32_t* d_lcp);

/* Per-kernel accounting of the last ps5_sort_device call when profiling is
 * on (ps5_ctx_set_profile(ctx, 1)): each launch is bracketed by HIP events on
 * the context's stream. */
int ps5_ctx_set_profile(ps5_ctx* ctx, int on);
int ps5_ctx_kernel_count(ps5_ctx* ctx);
/* name (static string), summed milliseconds, launches, algorithmic bytes */
int ps5_ctx_kernel_stat(ps5_ctx* ctx, int i, const char** name, double* ms,
                        uint64_t* launches, uint64_t* alg_bytes);
/* Work statistics of the last sort: rounds (8-byte depth steps), levels,
 * sample-sort segments, leaves, and the number of host synchronisations. */
int ps5_ctx_work_stats(ps5_ctx* ctx, uint64_t* rounds, uint64_t* levels,
                       uint64_t* segments, uint64_t* leaves, uint64_t* syncs);

/* ---- device memory helpers (so callers need no HIP headers) ------------ */
int ps5_dev_malloc(void** p, uint64_t bytes);
int ps5_dev_free(void* p);
int ps5_memcpy_h2d(void* dst, const void* src, uint64_t bytes);
int ps5_memcpy_d2h(void* dst, const void* src, uint64_t bytes);
int ps5_device_synchronize(void);
/* Device-to-device copy by a plain streaming kernel (16-byte aligned). */
int ps5_dev_copy(void* dst, const void* src, uint64_t bytes);
/* Milliseconds of one ps5_sort_device call measured with HIP events on the
 * context's stream (device-resident time). */
int ps5_sort_device_timed(ps5_ctx* ctx, const uint8_t* d_arena, uint64_t arena_bytes,
                          const uint32_t* d_offs, uint64_t n,
                          uint32_t* d_perm, uint32_t* d_lcp, double* ms);

/* ---- tuning ------------------------------------------------------------ */

/* Shape parameters of a context's sorts (they never change the output):
 *   leaf_target    strings per bucket the multi-workgroup levels aim for,
 *                  in [64, 1024] (default 768); sets how many levels run and
 *                  how large the on-chip blocks are;
 *   max_treebits   splitter-tree depth limit of one level, in [1, 10]
 *                  (default 10: up to 1023 splitters, 2047 buckets, the
 *                  reference's DefaultTreebits, bingmann-sample_sort.hpp:53);
 *   max_oversample sample keys per splitter, even, in [2, 32] (default 32;
 *                  the reference samples 2 per splitter, .hpp:1384-1410). */
int ps5_ctx_set_tuning(ps5_ctx* ctx, uint32_t leaf_target, uint32_t max_treebits, uint32_t max_oversample);
int ps5_ctx_get_tuning(ps5_ctx* ctx, uint32_t* leaf_target, uint32_t* max_treebits, uint32_t* max_oversample);

/* Tile shapes and key width of a context's sorts (they never change the
 * output).  No counterpart in the reference (its tiles are per-thread
 * buckets and its key cache is one uint64, bingmann-sample_sort.hpp:49):
 *   chunk_max      keys per wave chunk of a multi-workgroup level, in
 *                  [2048, 9216] (default 9216): the unit of stable ranks
 *                  and of the scatter's LDS reorder;
 *   chunk_max_root the same in a level of roots, in [2048, 12032] (default
 *                  12032; the root scatter stages 2-byte positions);
 *   min_tiles      tiles a level is spread over at least, its chunks
 *                  shrinking (not below 2048 keys) when it is small, in
 *                  [1, 65536] (default 512);
 *   leaf_keys      keys (8 characters each) a leaf pass fetches per string
 *                  at a time, in [1, 8] (default 8). */
int ps5_ctx_set_shape(ps5_ctx* ctx, uint32_t chunk_max, uint32_t chunk_max_root, uint32_t min_tiles,
                      uint32_t leaf_keys);
int ps5_ctx_get_shape(ps5_ctx* ctx, uint32_t* chunk_max, uint32_t* chunk_max_root, uint32_t* min_tiles,
                      uint32_t* leaf_keys);

/* Autotuner: sorts the given input (arguments as ps5_sort_device /
 * ps5_sort_batch_device, outputs written as by them) under candidate
 * settings -- coordinate descent over the three tuning parameters, the three
 * tile shapes and the key width, each candidate timed as the best of two
 * sorts -- keeps the fastest in the context and returns its tuning
 * parameters and its time (ps5_ctx_get_shape reads the chosen shape).  No
 * counterpart in the reference. */
int ps5_autotune_device(ps5_ctx* ctx, const uint8_t* d_arena, uint64_t arena_bytes,
                        const uint32_t* d_offs, uint64_t n, uint32_t* d_perm, uint32_t* d_lcp,
                        uint32_t* leaf_target, uint32_t* max_treebits, uint32_t* max_oversample,
                        double* best_ms);
int ps5_autotune_batch_device(ps5_ctx* ctx, const uint8_t* d_arena, uint64_t arena_bytes,
                              const uint64_t* d_offs, const uint64_t* set_first, uint32_t nsets,
                              uint32_t* d_perm, uint32_t* d_lcp, uint32_t* leaf_target,
                              uint32_t* max_treebits, uint32_t* max_oversample, double* best_ms);

/* ---- serving: continuous batching under an HBM budget ------------------ */

typedef struct ps5_server ps5_server;

/* `slots` workers each own a sort context (HIP stream + device buffers).  A
 * free slot takes every waiting request that fits under hbm_budget bytes of
 * staged data (in arrival order; a request larger than the budget runs alone)
 * and at most max_batch_strings strings (0: no limit), and sorts them as one
 * batch; a request joins a batch after its first only while the batch then
 * holds at most 2^20 strings, so larger requests run as batches of their own
 * (batching them would save only the per-sort overhead and hold their
 * results back for their batch mates).  Slots run concurrently on the device.  Admission is decided on the
 * device: submitted requests wait in a ring in mapped host memory, and a
 * kernel launched by the free slot claims them against a ledger of admitted
 * and released bytes kept in device memory.  With max_batch_strings > 0
 * every slot allocates its device buffers for such a batch at creation (up to
 * min(hbm_budget / slots, 64 * max_batch_strings) characters), so that the
 * first requests do not wait for allocations.  No counterpart in the
 * reference (it sorts one set per call). */
int ps5_server_create(ps5_server** out, int device, uint64_t hbm_budget, uint32_t slots,
                      uint64_t max_batch_strings);
void ps5_server_destroy(ps5_server* srv);
/* Submit host strings (arena + uint32 offsets, caller keeps them alive until
 * ps5_server_wait returns); results go to perm[n] and lcp[n] (lcp may be
 * NULL, lcp[0] untouched).  Thread-safe.  Returns a request id, or 0 when the
 * request is rejected (ps5_last_error says why): a malformed request -- an
 * offset at or past arena_bytes, or an arena whose last byte is not a NUL --
 * never enters a batch.  Failure isolation: a batch that fails is run again
 * one request at a time, so only a request that fails on its own reports an
 * error; when no slot could start, every request completes with an error. */
uint64_t ps5_server_submit(ps5_server* srv, const uint8_t* arena, uint64_t arena_bytes,
                           const uint32_t* offs, uint64_t n, uint32_t* perm, uint32_t* lcp);
/* Block until request `id` is done; its latency from submit to completion.
 * Returns the request's own status (its message in ps5_last_error). */
int ps5_server_wait(ps5_server* srv, uint64_t id, double* latency_ms);
/* Latency percentiles over completed requests, batches run, and throughput
 * (arena bytes of completed requests / time from first submit to last finish). */
int ps5_server_stats(ps5_server* srv, double* p50_ms, double* p99_ms, uint64_t* completed,
                     uint64_t* batches, double* throughput_gbs);
/* Slots running now, and requests completed with an error so far. */
int ps5_server_health(ps5_server* srv, uint32_t* live_slots, uint64_t* failed);
/* The device-side admission ledger: requests admitted so far, and bytes
 * admitted and not yet released (0 once every request is done).  No
 * counterpart in the reference. */
int ps5_server_ledger(ps5_server* srv, uint64_t* admitted, uint64_t* in_use_bytes);
/* Test hook: the request that will receive id `id` (ids count up from 1 in
 * submission order) fails inside its batch as a device error would, which
 * exercises the failure isolation from the host.  No counterpart in the
 * reference. */
int ps5_server_inject_fault(ps5_server* srv, uint64_t id);

/* ---- synthetic inputs (seeded, host memory) ---------------------------- */

/* Uniform strings: lengths uniform in [len_lo, len_hi], characters drawn
 * uniformly from alphabet[0..alpha_n); strings are appended (each followed by
 * a NUL) until the next would not fit in arena_cap bytes.  PRNG: xoshiro256**
 * seeded through splitmix64; a uniform integer in [lo,hi] is
 * lo + (next() >> 11) % (hi - lo + 1).  Writes at most offs_cap offsets;
 * returns the number of strings and the used arena bytes in *arena_used.
 * arena must hold arena_cap bytes; the unused tail is zeroed. */
uint64_t ps5_gen_uniform(uint64_t seed, uint64_t arena_cap, uint32_t len_lo,
                         uint32_t len_hi, const uint8_t* alphabet, uint32_t alpha_n,
                         uint8_t* arena, uint32_t* offs, uint64_t offs_cap,
                         uint64_t* arena_used);

/* High-LCP windows: a text of text_len i.i.d. uniform ACGT characters; copies
 * of nseg random segments of seg_len characters overwrite destinations drawn
 * uniformly until at least repeat_frac of the text has been written; then n
 * windows of win_len characters at uniform offsets become strings (each
 * followed by a NUL) in arena (n*(win_len+1) bytes). */
int ps5_gen_windows(uint64_t seed, uint64_t text_len, uint32_t nseg, uint32_t seg_len,
                    double repeat_frac, uint64_t n, uint32_t win_len,
                    uint8_t* arena, uint32_t* offs);

/* The reference's own artificial inputs random2/4/10/62/255
 * (src/tools/input.hpp:425-468): LCGRandom(1234567), string lengths 16..18
 * including the NUL, letter = letters[(rng()/100) % k].  size bytes are
 * generated and 9 NULs appended (arena must hold size+9 bytes).  Returns the
 * string count and fills offs (capacity offs_cap). */
uint64_t ps5_gen_reference_random(uint64_t size, const uint8_t* letters, uint32_t k,
                                  uint8_t* arena, uint32_t* offs, uint64_t offs_cap);

/* Prefix-heavy serving content: each string is one of nprefix random host
 * prefixes (length uniform in [8,40], chars 33..126) chosen Zipf(s) over ranks,
 * followed by a random path of length uniform in [10,60] over 33..126.
 * Fills up to arena_cap bytes as ps5_gen_uniform does. */
uint64_t ps5_gen_prefixed(uint64_t seed, uint64_t arena_cap, uint32_t nprefix,
                          double zipf_s, uint8_t* arena, uint32_t* offs,
                          uint64_t offs_cap, uint64_t* arena_used);

/* Batch workload: nsets sets; with s = splitmix64(seed + k), set k draws its
 * string count floor(exp(U(ln count_lo, ln count_hi))) and its alphabet (ACGT,
 * '0'..'9' or bytes 33..126, uniformly) from xoshiro256** seeded with s, its
 * string lengths U[len_lo, len_hi] from a stream seeded with
 * s ^ 0x5851F42D4C957F2D and its characters from one seeded with
 * s ^ 0xC4A2C7E5.  plan: fills counts[nsets] and bytes[nsets] (arena bytes of
 * each set: its strings with their NULs, then zero padding to a multiple of 8,
 * so that every set starts 8-byte aligned).  fill: writes set k at
 * arena + sum(bytes[0..k)) with uint64 offsets at offs + sum(counts[0..k)),
 * in parallel over sets. */
int ps5_gen_batch_plan(uint64_t seed, uint32_t nsets, uint64_t count_lo, uint64_t count_hi,
                       uint32_t len_lo, uint32_t len_hi, uint64_t* counts, uint64_t* bytes);
int ps5_gen_batch_fill(uint64_t seed, uint32_t nsets, uint64_t count_lo, uint64_t count_hi,
                       uint32_t len_lo, uint32_t len_hi, uint8_t* arena, uint64_t* offs);

#ifdef __cplusplus
}
#endif

#endif /* PS5_MI355X_H */
