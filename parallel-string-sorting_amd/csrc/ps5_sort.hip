// ps5_sort.hip -- pS5 string sample sort on MI355X (gfx950): HIP kernels and
// the host orchestration behind the C ABI in include/ps5_mi355x.h.
//
// Reference algorithm (paths relative to the reference checkout):
//   SampleSortStep sample/count/distribute  src/parallel/bingmann-parallel_sample_sort.hpp:1264-1611
//   splitter tree classify                   src/sequential/bingmann-sample_sortBTC.hpp:45-72,297-399
//   bucket recursion rules                   src/parallel/bingmann-parallel_sample_sort.hpp:1492-1584
//   bucket-boundary LCPs                     src/parallel/bingmann-parallel_sample_sort.hpp:311-396
//   LCP insertion sort base case             src/sequential/bingmann-lcp_inssort.hpp:36-187
//   key extraction                           src/tools/stringset.hpp:201-231
// The GPU formulation keeps the reference's outputs but reorganises the work
// breadth-first: every round handles all segments at one 8-byte depth, every
// pass is a kernel over a device-side work list, and buckets small enough for
// LDS are finished by one workgroup without another trip through HBM.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ps5_device.hpp"
#include "ps5_internal.hpp"
#include "ps5_kernels.hpp"
#include "../../include/ps5_mi355x.h"

using namespace ps5;

// ===========================================================================
// kernels
// ===========================================================================

// Round 0 keys: key0[i] = 8 characters of string i at depth 0.  The index
// at every root position is the position itself (root items carry F_IDENT),
// so idx0 is not written.
__global__ __launch_bounds__(kWG) void k_init(Bufs B, uint32_t n)
{
    const uint32_t i = blockIdx.x * kWG + tid_x();
    if (i >= n) return;
    B.key0[i] = load_key(B.arena, B.alen, str_off(B, i));
}

// Round 0 keys of the positions of listed items (the sets of a batch that
// start as tiny leaves or blocks, when the first count pass extracts the keys
// of the large ones), one workgroup per item at a time.
__global__ __launch_bounds__(kWG) void k_init_items(const Item* items, uint32_t nitems, Bufs B)
{
    for (uint32_t ii = blockIdx.x; ii < nitems; ii += gridDim.x) {
        const Item it = items[ii];
        if (!item_ok(it, B)) {
            if (tid_x() == 0) flag_error(B.cnt, ERR_BOUNDS);
            continue;
        }
        for (uint32_t p = it.begin + tid_x(); p < it.begin + it.size; p += kWG)
            B.key0[p] = load_key(B.arena, B.alen, str_off(B, p));
    }
}

// Reset counters and place the root item.
__global__ void k_setup(uint32_t* cnt, Item* list, int c_index, uint32_t n)
{
    if (tid_x() < C_NCNT) cnt[tid_x()] = 0;
    __syncthreads();
    if (tid_x() == 0) {
        Item it = { 0u, n, 0u, F_IDENT };
        list[0] = it;
        cnt[c_index] = 1;
    }
}

// Counters and the first kPeek items of a work list in one buffer, so that a
// single copy brings both to the host: out[0, C_NCNT) = counters, then
// min(cnt[ci], kPeek) items from byte kStageHead on.
constexpr uint32_t kPeek = 2048;
constexpr uint32_t kStageHead = 128;
constexpr uint32_t kStageSeq = kStageHead - 4;    // the copy's sequence number, written last
static_assert(C_NCNT * 4 <= kStageSeq, "counters must fit the stage head");
// Copy the counters and the head of a list into mapped host memory, then
// publish `seq`: every thread makes its stores visible to the host first, so
// a host that reads seq (acquire) reads the whole copy.  The stream runs
// k_stage after everything enqueued before it, so seq also says that all of
// that work is complete.
__global__ __launch_bounds__(kWG) void k_stage(const uint32_t* cnt, const Item* list, int ci, unsigned char* out,
                                               uint32_t seq)
{
    if (tid_x() < C_NCNT) reinterpret_cast<uint32_t*>(out)[tid_x()] = cnt[tid_x()];
    if (list) {
        const uint32_t nl = min(cnt[ci], kPeek);
        Item* dst = reinterpret_cast<Item*>(out + kStageHead);
        for (uint32_t i = tid_x(); i < nl; i += kWG) dst[i] = list[i];
    }
    __threadfence_system();
    __syncthreads();
    if (tid_x() == 0)
        __hip_atomic_store(reinterpret_cast<uint32_t*>(out + kStageSeq), seq, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
}

// copy 8-byte words of a level plan from mapped host memory to the device:
// the header with the descriptors, then the tile table
__device__ __forceinline__ void fetch_plan(const uint64_t* src, uint64_t* dst, uint64_t words, const uint64_t* src2,
                                           uint64_t* dst2, uint64_t words2)
{
    const uint64_t stride = (uint64_t)gridDim.x * kWG;
    for (uint64_t i = (uint64_t)blockIdx.x * kWG + tid_x(); i < words; i += stride) dst[i] = src[i];
    for (uint64_t i = (uint64_t)blockIdx.x * kWG + tid_x(); i < words2; i += stride) dst2[i] = src2[i];
}

__global__ __launch_bounds__(kWG) void k_fetch(const uint64_t* src, uint64_t* dst, uint64_t words,
                                               const uint64_t* src2, uint64_t* dst2, uint64_t words2)
{
    fetch_plan(src, dst, words, src2, dst2, words2);
}

// The same for the root level of a single root planned on the host, the
// sort's first kernel: its first workgroup also does k_setup's work
// (counters, the root item), one launch fewer before the root level.
__global__ __launch_bounds__(kWG) void k_root_fetch(const uint64_t* src, uint64_t* dst, uint64_t words,
                                                    const uint64_t* src2, uint64_t* dst2, uint64_t words2,
                                                    uint32_t* cnt, Item* list, uint32_t n)
{
    if (blockIdx.x == 0) {
        if (tid_x() < C_NCNT) cnt[tid_x()] = 0;
        __syncthreads();
        if (tid_x() == 0) {
            list[0] = Item{ 0u, n, 0u, F_IDENT };
            cnt[C_LROUND] = 1;
        }
    }
    fetch_plan(src, dst, words, src2, dst2, words2);
}

// zero the counters whose bit is set in `mask`
__global__ void k_reset(uint32_t* cnt, uint32_t mask)
{
    const int t = tid_x();
    if (t < C_NCNT && ((mask >> t) & 1u)) cnt[t] = 0;
}

// ---- multi-workgroup sample-sort step (segments > kBlkT) -------------------

// ---- device-planned levels ---------------------------------------------------

// Treebits of a multi-workgroup step on m strings (device copy of the host's
// large_treebits): the fewest levels of at most maxbits that bring buckets
// down to about `target` strings, the fan-out spread evenly over them.
__device__ __forceinline__ uint32_t dev_treebits(uint32_t m, uint32_t target, uint32_t maxbits)
{
    const double need = log2(fmax(2.0, (double)m / target));
    const int levels = max(1, (int)ceil(need / maxbits));
    const int k = (int)ceil(need / levels);
    return (uint32_t)max(1, min((int)maxbits, k));
}

struct PlanArgs {
    const Item* list;        // the level's large segments
    int ci;                  // their counter
    uint32_t cap_large;      // capacity of that list
    unsigned char* plan;     // header, descriptors, tile table (tiles at tiles_off)
    uint64_t tiles_off, tiles_cap, hist_cap, trees_cap;
    uint32_t leaf_target, max_treebits, max_oversample;
    uint32_t chunk_max, chunk_max_root, min_tiles;    // tile shapes (ps5_ctx_set_shape)
    uint32_t kmax;           // bound of the count launch (treebits are clamped to it)
    int done_rp;             // >= 0: mark the blocks of round parity done_rp listed so far as sorted
};

// Plan one level on the device (one workgroup of kBigWG threads): the same
// descriptors, tile table and totals that the host's large_step computes
// from a read-back list.  A level that does not fit the buffers is flagged
// (ERR_CAP) and planned empty.
__global__ __launch_bounds__(kBigWG) void k_plan(PlanArgs A, uint32_t* cnt)
{
    __shared__ uint32_t s_tmp[kBigWaves];
    __shared__ uint32_t s_tb[kBigWG];        // tile offsets of a block of segments, for its tile table
    __shared__ unsigned long long s_tot;
    __shared__ uint32_t s_ident, s_kmax;
    LevelHdr* hdr = reinterpret_cast<LevelHdr*>(A.plan);
    LargeDesc* desc = reinterpret_cast<LargeDesc*>(A.plan + sizeof(LevelHdr));
    TileRef* tiles = reinterpret_cast<TileRef*>(A.plan + A.tiles_off);
    const uint32_t tid = tid_x();
    if (A.done_rp >= 0 && tid < kBlkClasses) cnt[C_BDONE + tid] = cnt[C_BLOCK + 2 * tid + A.done_rp];
    uint32_t nl = cnt[A.ci];
    if (nl > A.cap_large) {      // (an overflow of the list has been flagged)
        nl = 0;
        if (tid == 0) flag_error(cnt, ERR_CAP);
    }
    if (tid == 0) {
        s_tot = 0;
        s_ident = 1;
        s_kmax = 1;
    }
    __syncthreads();
    unsigned long long tot = 0;
    bool ident = true;
    for (uint32_t i = tid; i < nl; i += kBigWG) {
        tot += A.list[i].size;
        ident = ident && (A.list[i].flags & F_IDENT);
    }
    atomicAdd(&s_tot, tot);
    if (!ident) atomicAnd(&s_ident, 0u);
    __syncthreads();
    const uint64_t cmax = s_ident ? A.chunk_max_root : A.chunk_max;
    const uint64_t climit = min(cmax, max((uint64_t)kChunkMin, (s_tot + kWaves * A.min_tiles - 1) / (kWaves * A.min_tiles)));
    // descriptors: tile and histogram bases are exclusive scans over the
    // segments (blocks of kBigWG with carries); pad holds the tile base
    uint32_t tcarry = 0;
    uint64_t hcarry = 0;
    for (uint32_t b0 = 0; b0 < nl; b0 += kBigWG) {
        const uint32_t i = b0 + tid;
        uint32_t T = 0, hc = 0;
        LargeDesc D{};
        if (i < nl) {
            D.seg = A.list[i];
            T = (uint32_t)((D.seg.size + climit * kWaves - 1) / (climit * kWaves));
            D.T = T;
            D.chunk = (uint32_t)((D.seg.size + (uint64_t)T * kWaves - 1) / ((uint64_t)T * kWaves));
            D.k = min(dev_treebits(D.seg.size, A.leaf_target, A.max_treebits), A.kmax);
            const uint32_t nspl = (1u << D.k) - 1;
            D.oversample = max(2u, min(A.max_oversample, (kMaxSamples / nspl) & ~1u));
            hc = ((2u << D.k) - 1) * T * kWaves;
            atomicMax(&s_kmax, D.k);
        }
        uint32_t ttot, htot;
        const uint32_t tex = block_excl_scan<kBigWaves>(T, &ttot, s_tmp);
        const uint32_t hex = block_excl_scan<kBigWaves>(hc, &htot, s_tmp);
        if (i < nl) {
            D.pad = tcarry + tex;
            D.hist_base = (uint32_t)(hcarry + hex);
            desc[i] = D;
        }
        // this block's entries of the tile table, written by the whole
        // workgroup (within its capacity; an overflow is flagged below):
        // tile t of the block belongs to the last segment whose offset is
        // <= t (every segment has at least one tile)
        s_tb[tid] = tex;
        __syncthreads();
        const uint32_t nseg = min(nl - b0, (uint32_t)kBigWG);
        for (uint32_t t = tid; t < ttot && tcarry + t < A.tiles_cap; t += kBigWG) {
            uint32_t lo = 0, hi = nseg;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_tb[mid] <= t) lo = mid;
                else hi = mid;
            }
            tiles[tcarry + t] = TileRef{ b0 + lo, t - s_tb[lo] };
        }
        __syncthreads();     // s_tb is written again by the next block
        tcarry += ttot;
        hcarry += htot;
    }
    const bool fits = tcarry <= A.tiles_cap && hcarry <= A.hist_cap && hcarry <= 0xFFFFFFFFull && nl <= A.trees_cap;
    if (!fits && tid == 0) flag_error(cnt, ERR_CAP);
    const uint32_t ntiles = fits ? tcarry : 0u;
    if (tid == 0)
        *hdr = LevelHdr{ fits ? nl : 0u, ntiles, fits ? (uint32_t)hcarry : 0u, s_kmax, { 0u, 0u, 0u, 0u } };
}

// Splitter trees of the large segments of a level, one workgroup each.
// D.oversample * nspl random keys are sorted (4-key runs per thread by a
// sorting network, then merge-path rounds in LDS); splitter i is sample
// oversample/2 + oversample*i (the reference's choice for oversample factor 2,
// bingmann-sample_sort_tree_builder.hpp:44-58) and the tree is stored in
// level order: trees[seg*2048 + t], t = 1..nspl, followed by the splitters in
// sorted order at trees[seg*2048 + 1024 + i], i < nspl.  With from_arena (the root
// of a sort whose keys are not extracted yet) the samples are read from the
// arena directly.
__device__ __forceinline__ void tree_of_segment(const LargeDesc D, const Bufs& B, uint64_t seed, key64* tree,
                                                int from_arena, key64* s_samp)
{
    const Item seg = D.seg;
    if (!item_ok(seg, B)) {
        if (tid_x() == 0) flag_error(B.cnt, ERR_BOUNDS);
        return;
    }
    const key64* K = (seg.flags & F_BUFB) ? B.key1 : B.key0;
    const int k = (int)D.k, nspl = (1 << k) - 1, a = (int)D.oversample;
    const uint32_t ss = (uint32_t)(a * nspl);
    uint32_t P = 64;
    while (P < ss) P <<= 1;
    if (P > (uint32_t)kMaxSamples) {
        if (tid_x() == 0) flag_error(B.cnt, ERR_BOUNDS);
        return;
    }
    // every sample load is issued first (kPer consecutive samples per
    // thread), each thread sorts its kPer keys with a network, then runs are
    // merged pairwise by merge path: a co-rank search places a thread's kPer
    // outputs, then it merges them sequentially (padding ~0 sorts last; equal
    // keys need no tie order, only the splitters are kept)
    constexpr int kPer = kMaxSamples / kBigWG;
    static_assert(kMaxSamples % kBigWG == 0 && kPer == 4, "four samples per thread");
    const uint32_t o = tid_x() * kPer;
    key64 v[kPer];
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const uint32_t j = o + e;
        v[e] = ~0ull;
        if (j < ss) {
            const uint64_t h = mix64(seed ^ ((uint64_t)seg.begin << 21) ^ (uint64_t)j);
            const uint32_t pos = seg.begin + (uint32_t)(((h >> 32) * (uint64_t)seg.size) >> 32);
            v[e] = from_arena ? load_key(B.arena, B.alen, str_off(B, pos)) : K[pos];
        }
    }
    const bool act = o < P;
    if (act) {
        constexpr int net[][2] = { { 0, 1 }, { 2, 3 }, { 0, 2 }, { 1, 3 }, { 1, 2 } };
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const key64 x = v[net[c][0]], y = v[net[c][1]];
            v[net[c][0]] = min(x, y);
            v[net[c][1]] = max(x, y);
        }
#pragma unroll
        for (int e = 0; e < kPer; ++e) s_samp[o + e] = v[e];
    }
    __syncthreads();
    key64* sa = s_samp;
    for (uint32_t L = kPer; L < P; L <<= 1) {
        if (act) {
            const uint32_t a0 = (o / (2 * L)) * (2 * L), b0 = a0 + L, diag = o - a0;
            uint32_t lo = diag > L ? diag - L : 0u, hi = min(diag, L);
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sa[a0 + mid] <= sa[b0 + diag - 1 - mid]) lo = mid + 1;
                else hi = mid;
            }
            // branch-free steps on byte offsets, as in the block sort: one
            // read of the consumed run's next head (unused past the run's
            // end; s_samp has one spare slot)
            const unsigned char* sbase = reinterpret_cast<const unsigned char*>(sa);
            uint32_t pa = (a0 + lo) * 8u, pb = (b0 + diag - lo) * 8u;
            const uint32_t ea = (a0 + L) * 8u, eb = (b0 + L) * 8u;
            key64 ka = pa < ea ? *reinterpret_cast<const key64*>(sbase + pa) : 0ull;
            key64 kb = pb < eb ? *reinterpret_cast<const key64*>(sbase + pb) : 0ull;
#pragma unroll
            for (int e = 0; e < kPer; ++e) {
                const bool take_a = pb >= eb || (pa < ea && ka <= kb);
                const uint32_t src = take_a ? pa : pb;
                v[e] = take_a ? ka : kb;
                const key64 nx = *reinterpret_cast<const key64*>(sbase + src + 8u);
                pa += take_a ? 8u : 0u;
                pb += take_a ? 0u : 8u;
                ka = take_a ? nx : ka;
                kb = take_a ? kb : nx;
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int e = 0; e < kPer; ++e) sa[o + e] = v[e];
        }
        __syncthreads();
    }
    const key64* sorted = sa;
    for (int t = tid_x() + 1; t <= nspl; t += kBigWG) tree[t] = sorted[a / 2 + a * inorder_of(t, k)];
    for (int i = tid_x(); i < nspl; i += kBigWG) tree[1024 + i] = sorted[a / 2 + a * i];
}

__global__ __launch_bounds__(kBigWG) void k_tree(Plan P, Bufs B, uint64_t seed, key64* trees, int from_arena)
{
    __shared__ key64 s_samp[kMaxSamples + 1];     // + 1: a merge step reads the slot after the one it takes
    const uint32_t nlarge = P.hdr->nlarge;
    for (uint32_t si = blockIdx.x; si < nlarge; si += gridDim.x) {
        __syncthreads();     // the previous segment's samples are consumed
        tree_of_segment(P.desc[si], B, seed, trees + (uint64_t)si * kTreeSlots, from_arena, s_samp);
    }
}

// Classify one tile of kWaves wave chunks.  Each wave ranks its own chunk
// of D.chunk keys stably (ballot matching, ranks follow input order) and the
// chunk counts go to hist[hist_base + (b * T + t) * kWaves + w]: bucket-major,
// then tile, then wave -- the order of the reference's prefix sum
// (.hpp:1437-1460) with wave chunks as the parts -- so an exclusive scan of
// hist gives every chunk's base offset per bucket.  bktrank gets bucket and
// rank inside the chunk.  With from_arena (the root of a sort, positions =
// input indices) the keys are extracted from the arena here and stored: this
// replaces the separate extraction pass and saves reading the keys back.
// The tree and the sorted splitters sit in static LDS (their addresses are
// constants, so a descent step needs no address arithmetic); the per-wave
// 16-bit counts[kWaves][2^(kmax+1)] in dynamic LDS, sized by the largest
// tree of the level (kmax treebits).
__host__ __device__ constexpr uint32_t count_lds_bytes(int kmax)
{
    return kWaves * (2u << kmax) * 2u;
}

// Load a read-only value that every lane of the wave reads at the same
// address, through the constant address space (scalar loads).
template <typename T>
__device__ __forceinline__ T load_uniform(const T* p)
{
    static_assert(sizeof(T) % 4 == 0, "whole dwords");
    using cword = const __attribute__((address_space(4))) uint32_t;
    cword* q = (cword*)(uintptr_t)p;
    uint32_t w[sizeof(T) / 4];
#pragma unroll
    for (uint32_t i = 0; i < sizeof(T) / 4; ++i) w[i] = q[i];
    T v;
    __builtin_memcpy(&v, w, sizeof(T));
    return v;
}

// NB: bits of a bucket id in this launch (kmax + 1, at least 8).  Lanes of
// one bucket are matched through a per-wave LDS table of lane masks on the
// id's low kPeerBits bits and by ballots on the rest (a few vector
// instructions per bit); ids of 11 bits use ballots only, since their
// launch's LDS leaves no room for the tables at full occupancy.
template <int NB>
constexpr int kPeerBits = NB <= 10 ? (NB < 8 ? NB : 8) : 0;

template <int NB>
__device__ __forceinline__ void count_tile(uint32_t ti, const Plan& P, const Bufs& B, const key64* trees,
                                           uint32_t* hist, int from_arena, int kmax, key64* s_tree, key64* s_sorted,
                                           uint16_t* s_hist_base, uint64_t* s_peer)
{
    constexpr uint32_t kNoBkt = (1u << NB) - 1;   // id of inactive lanes, above every bucket of the launch
    const uint32_t nb2 = 2u << kmax;               // bucket slots per wave
    ti = __builtin_amdgcn_readfirstlane(ti);
    const TileRef tr = load_uniform(P.tiles + ti);     // (read-only here, the same for the workgroup)
    const LargeDesc D = load_uniform(P.desc + tr.seg);
    const Item seg = D.seg;
    const int k = (int)D.k, nbkt = (2 << k) - 1;
    if (k > kmax || k + 1 > NB) {   // the launch must size LDS and ids for every tree of the level
        if (tid_x() == 0) flag_error(B.cnt, ERR_BOUNDS);
        return;
    }
    const key64* tree = trees + (uint64_t)tr.seg * kTreeSlots;
    const int nspl = (1 << k) - 1;
    for (int i = tid_x(); i <= nspl; i += kWG) s_tree[i] = tree[i];
    for (int i = tid_x(); i < nspl; i += kWG) s_sorted[i] = tree[1024 + i];
    for (uint32_t i = tid_x(); i < kWaves * nb2; i += kWG) s_hist_base[i] = 0;
    __syncthreads();
    const key64* K = (seg.flags & F_BUFB) ? B.key1 : B.key0;
    const uint32_t t = tr.t;
    const uint32_t w = wave_id(), lane = lane_id();
    uint16_t* s_hist = s_hist_base + w * nb2;      // this wave's counts
    uint64_t* peer_tab = s_peer + (w << kPeerBits<NB>);
    const uint64_t lanebit = 1ull << lane;
    const uint32_t cb = (t * kWaves + w) * D.chunk;
    const uint32_t cmax = (seg.flags & F_IDENT) ? kChunkMaxIdent : kChunkMax;
    const uint32_t ce = (item_ok(seg, B) && t < D.T && D.chunk <= cmax) ? min(seg.size, cb + D.chunk) : 0u;
    // kGroups 64-key groups per iteration: all loads are in flight before the
    // tree descents, then the groups are ranked in order (stable)
    constexpr int kGroups = 8;
    // extracting at the root, the string offsets of the next iteration are
    // loaded one iteration ahead (the first link of offset -> characters)
    uint64_t off_next[kGroups];
    if (from_arena) {
#pragma unroll
        for (int j = 0; j < kGroups; ++j) {
            const uint32_t p = cb + 64 * j + lane;
            off_next[j] = p < ce ? str_off(B, seg.begin + p) : 0ull;
        }
    }
    for (uint32_t g = cb; g < ce; g += 64 * kGroups) {
        key64 kv[kGroups];
        if (from_arena) {
            uint64_t off[kGroups];
#pragma unroll
            for (int j = 0; j < kGroups; ++j) off[j] = off_next[j];
#pragma unroll
            for (int j = 0; j < kGroups; ++j) {
                const uint32_t p = g + 64 * kGroups + 64 * j + lane;
                off_next[j] = p < ce ? str_off(B, seg.begin + p) : 0ull;
            }
#pragma unroll
            for (int j = 0; j < kGroups; ++j) {
                const uint32_t p = g + 64 * j + lane;
                kv[j] = p < ce ? load_key(B.arena, B.alen, off[j]) : 0ull;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kGroups; ++j) {
                const uint32_t p = g + 64 * j + lane;
                kv[j] = p < ce ? K[seg.begin + p] : 0ull;
            }
        }
        uint32_t bv[kGroups];
        classify_keys<kGroups>(kv, s_tree, s_sorted, k, bv);
#pragma unroll
        for (int j = 0; j < kGroups; ++j) {
            const uint32_t p = g + 64 * j + lane;
            if (p >= ce) bv[j] = kNoBkt;
        }
#pragma unroll
        for (int j = 0; j < kGroups; ++j) {
            const uint32_t p = g + 64 * j + lane;
            uint64_t peers;
            if constexpr (kPeerBits<NB> > 0) peers = match_bucket_lds<NB, kPeerBits<NB>>(bv[j], peer_tab, lanebit);
            else peers = match_bucket<NB>(bv[j]);
            const uint32_t r = (uint32_t)__popcll(peers & lanemask_lt());
            if (p < ce) {
                const uint32_t old = s_hist[bv[j]];
                if (r == 0) s_hist[bv[j]] = (uint16_t)(old + (uint32_t)__popcll(peers));
                B.bktrank[seg.begin + p] = (bv[j] << kRankBits) | (old + r);
            }
        }
        if (from_arena) {     // the extracted keys, for the scatter
#pragma unroll
            for (int j = 0; j < kGroups; ++j) {
                const uint32_t p = g + 64 * j + lane;
                if (p < ce) B.key0[seg.begin + p] = kv[j];
            }
        }
    }
    __syncthreads();
    uint32_t* h = hist + D.hist_base;
    for (int i = tid_x(); i < nbkt * kWaves; i += kWG) {
        const uint32_t b = (uint32_t)i / kWaves, ww = (uint32_t)i % kWaves;
        h[((uint32_t)b * D.T + t) * kWaves + ww] = s_hist_base[ww * nb2 + b];
    }
}

// XCD-contiguous tiles: workgroup g runs on XCD g % kXcds (the grid is a
// multiple of kXcds), and XCD x takes the tiles [N x / kXcds, N (x+1) /
// kXcds): one per workgroup, or with LOOP (levels planned on the device)
// each workgroup loops over its XCD's share.
template <int NB, bool LOOP>
__global__ __launch_bounds__(kWG) void k_count(Plan P, Bufs B, const key64* trees, uint32_t* hist, int from_arena,
                                                int kmax)
{
    static_assert(NB >= 2 && NB <= kMaxTreebits + 1, "bucket ids of up to 2047 buckets");
    static_assert(kChunkMax <= 65535 && kChunkMaxIdent <= 65535, "per-wave bucket counts must fit 16 bits");
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ key64 s_tree[1 << (NB - 1)];        // trees of up to NB - 1 levels
    __shared__ key64 s_sorted[1 << (NB - 1)];
    __shared__ uint64_t s_peer[kWaves << kPeerBits<NB>];     // lane-mask tables, kept zero between uses
    uint16_t* s_hist_base = reinterpret_cast<uint16_t*>(smem);
    if constexpr (kPeerBits<NB> > 0)
        for (int i = tid_x(); i < (kWaves << kPeerBits<NB>); i += kWG) s_peer[i] = 0;   // (a barrier follows in count_tile)
    const uint32_t N = P.hdr->ntiles, per = gridDim.x / kXcds, xcd = blockIdx.x % kXcds;
    const uint32_t thi = (uint32_t)((uint64_t)N * (xcd + 1) / kXcds);
    uint32_t ti = (uint32_t)((uint64_t)N * xcd / kXcds) + blockIdx.x / kXcds;
    if (!LOOP) {
        if (ti < thi) count_tile<NB>(ti, P, B, trees, hist, from_arena, kmax, s_tree, s_sorted, s_hist_base, s_peer);
        return;
    }
    for (; ti < thi; ti += per) {
        __syncthreads();     // the previous tile's counts are written out
        count_tile<NB>(ti, P, B, trees, hist, from_arena, kmax, s_tree, s_sorted, s_hist_base, s_peer);
    }
}

// exclusive scan of hist[0..len): block sums
constexpr int kScanPer = 16;
constexpr int kScanBlock = kWG * kScanPer;   // 4096 entries per block

// (the length is the level's hl; workgroups loop over the blocks of kScanBlock entries)
__global__ __launch_bounds__(kWG) void k_scan_reduce(const uint32_t* in, const LevelHdr* hdr, uint32_t* part)
{
    __shared__ uint32_t s_tmp[16];
    const uint32_t len = hdr->hl, nb = (len + kScanBlock - 1) / kScanBlock;
    for (uint32_t blk = blockIdx.x; blk < nb; blk += gridDim.x) {
        const uint32_t base = blk * kScanBlock;
        uint32_t sum = 0;
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) {
            const uint32_t i = base + j * kWG + tid_x();
            if (i < len) sum += in[i];
        }
        uint32_t tot;
        block_excl_scan<kWaves>(sum, &tot, s_tmp);
        if (tid_x() == 0) part[blk] = tot;
    }
}

// exclusive scan of the block sums (one workgroup, any count)
__global__ __launch_bounds__(kWG) void k_scan_top(uint32_t* part, const LevelHdr* hdr)
{
    __shared__ uint32_t s_tmp[16];
    const uint32_t nparts = (hdr->hl + kScanBlock - 1) / kScanBlock;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nparts; base += kWG) {
        const uint32_t i = base + tid_x();
        const uint32_t v = i < nparts ? part[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<kWaves>(v, &tot, s_tmp);
        if (i < nparts) part[i] = carry + ex;
        carry += tot;
    }
}

// exclusive scan inside each block, offset by its block prefix (in place)
// (data is 16-byte aligned: a thread's kScanPer consecutive entries move as
// 16-byte vectors, except in a block cut short by len)
__global__ __launch_bounds__(kWG) void k_scan_down(uint32_t* data, const LevelHdr* hdr, const uint32_t* part)
{
    static_assert(kScanPer % 4 == 0, "whole 16-byte vectors per thread");
    __shared__ uint32_t s_tmp[16];
    const uint32_t len = hdr->hl, nb = (len + kScanBlock - 1) / kScanBlock;
    for (uint32_t blk = blockIdx.x; blk < nb; blk += gridDim.x) {
        const uint32_t base = blk * kScanBlock + tid_x() * kScanPer;
        const bool full = base + kScanPer <= len;
        uint32_t v[kScanPer];
        if (full) {
            const uint4* p = reinterpret_cast<const uint4*>(data + base);
#pragma unroll
            for (int q = 0; q < kScanPer / 4; ++q) {
                const uint4 x = p[q];
                v[4 * q] = x.x;
                v[4 * q + 1] = x.y;
                v[4 * q + 2] = x.z;
                v[4 * q + 3] = x.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kScanPer; ++j) v[j] = base + j < len ? data[base + j] : 0u;
        }
        uint32_t sum = 0;
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) sum += v[j];
        uint32_t tot;
        uint32_t run = block_excl_scan<kWaves>(sum, &tot, s_tmp) + part[blk];
        uint32_t o[kScanPer];
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) {
            o[j] = run;
            run += v[j];
        }
        if (full) {
            uint4* p = reinterpret_cast<uint4*>(data + base);
#pragma unroll
            for (int q = 0; q < kScanPer / 4; ++q) p[q] = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
        } else {
#pragma unroll
            for (int j = 0; j < kScanPer; ++j)
                if (base + j < len) data[base + j] = o[j];
        }
    }
}

// LDS layout of k_scatter (dynamic shared memory).  In a level of root
// segments (IDENT: indices = positions) a chunk stages 2-byte positions
// instead of 4-byte indices, which leaves room for longer chunks: longer
// runs per bucket, fewer partial cache lines written.
template <bool IDENT>
struct ScatLds {
    static constexpr uint32_t kMax = IDENT ? kChunkMaxIdent : kChunkMax;
    static constexpr int kPer = (kMax + kBigWG - 1) / kBigWG;   // keys per thread
    static constexpr uint32_t kKeyOff = 0;
    static constexpr uint32_t kIdxOff = kKeyOff + kMax * 8;
    static constexpr uint32_t kBktOff = kIdxOff + kMax * (IDENT ? 2 : 4);
    static constexpr uint32_t kStartOff = (kBktOff + kMax * 2 + 15) / 16 * 16;
    static constexpr uint32_t kBaseOff = kStartOff + 2048 * 4;
    static constexpr uint32_t kTmpOff = kBaseOff + 2048 * 4;
    static constexpr uint32_t kBytes = kTmpOff + 64 * 4;
};
static_assert(ScatLds<false>::kBytes <= 160 * 1024 && ScatLds<true>::kBytes <= 160 * 1024,
              "a chunk image fits a CU's LDS");

// Move wave chunks (at most kChunkMax keys each, kChunkMaxIdent in a level
// of roots) to their buckets (distribute, .hpp:1464-1490, but stable).  A
// chunk is first reordered by bucket in
// LDS, then written out position by position, so consecutive lanes store
// consecutive addresses inside each bucket: full cache lines instead of the
// scattered 8-byte stores of a direct scatter.
//
// The chunk's LDS image takes most of a CU's LDS, so one workgroup runs per
// CU; to keep memory busy, each workgroup loops over chunks and loads the
// next chunk (keys, indices, ranks and histogram entries) into registers
// before it reorders and writes the current one.  Chunks are split into
// kXcds contiguous ranges, one per XCD (workgroup g runs on XCD g % kXcds),
// so the chunks in flight on an XCD are neighbours that share histogram
// lines and bucket output lines in its L2.

// tid_x() through an opaque move: addresses derived from it are
// recomputed per chunk instead of being hoisted out of the chunk loop,
// where they would hold registers the two chunk buffers need
__device__ __forceinline__ uint32_t scat_tid()
{
    uint32_t t;
    asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"((uint32_t)threadIdx.x));
    return t;
}

template <int PER>
struct ScatChunk {       // one chunk's description and its prefetched inputs
    LargeDesc D;
    uint32_t t, w, cb, m;
    uint32_t br[PER], iv[PER];
    key64 kv[PER];
    uint32_t h0, h0n, h1, h1n, hseg;   // histogram entries of buckets tid, tid + kBigWG; segment base
};

// Issue the loads of chunk c.  Every load is unconditional (indices are
// clamped into the arrays and invalid lanes discard their values when the
// chunk is processed): the same number of loads every time lets the
// compiler wait for exactly the older loads when the chunk is used, while
// the next chunk's loads stay in flight.
template <bool IDENT>
__device__ __forceinline__ void scat_fetch(const LargeDesc* __restrict__ desc, const TileRef* __restrict__ tiles,
                                           const Bufs& B, const uint32_t* __restrict__ hscan, uint32_t c,
                                           ScatChunk<ScatLds<IDENT>::kPer>& S)
{
    const uint32_t tid = scat_tid();
    // the chunk's tile and descriptor are read-only in this kernel and the
    // same for the whole workgroup: scalar loads, counted apart from the
    // vector loads and stores of the chunks
    c = __builtin_amdgcn_readfirstlane(c);
    const TileRef tr = load_uniform(tiles + c / kWaves);
    S.w = c % kWaves;
    S.t = tr.t;
    S.D = load_uniform(desc + tr.seg);
    const Item seg = S.D.seg;
    S.cb = (S.t * kWaves + S.w) * S.D.chunk;
    const bool ok = item_ok(seg, B) && S.t < S.D.T && S.D.chunk <= ScatLds<IDENT>::kMax && S.cb < seg.size;
    S.m = ok ? min(seg.size - S.cb, S.D.chunk) : 0u;
    const key64* K = (seg.flags & F_BUFB) ? B.key1 : B.key0;
    // a root's indices are implicit: its third load repeats the rank load
    const uint32_t* I = (seg.flags & F_IDENT) ? B.bktrank : (seg.flags & F_BUFB) ? B.idx1 : B.idx0;
    const uint32_t qmax = B.n - 1;
#pragma unroll
    for (int e = 0; e < ScatLds<IDENT>::kPer; ++e) {
        const uint32_t q = min(seg.begin + S.cb + e * kBigWG + tid, qmax);
        S.br[e] = B.bktrank[q];
        S.kv[e] = K[q];
        if (!IDENT) S.iv[e] = I[q];     // (a level of roots stages positions)
    }
    // histogram entries; entry e's count is scan[e + 1] - scan[e], and the
    // entry after the segment's last one is segbase + size (fixed up on use)
    const uint32_t* h = hscan + S.D.hist_base;
    const uint32_t T = max(S.D.T, 1u), nbkt = (2u << S.D.k) - 1;
    const uint32_t last = ((nbkt - 1) * T + (T - 1)) * kWaves + (kWaves - 1);
    const uint32_t t = min(S.t, T - 1);
    const uint32_t e0 = (min(tid, nbkt - 1) * T + t) * kWaves + S.w;
    const uint32_t e1 = (min(tid + kBigWG, nbkt - 1) * T + t) * kWaves + S.w;
    S.hseg = h[0];
    S.h0 = h[e0];
    S.h0n = h[min(e0 + 1, last)];
    S.h1 = h[e1];
    S.h1n = h[min(e1 + 1, last)];
}

// Stage one prefetched chunk in LDS, reordered by bucket (after the
// previous chunk's image has been written out).  Returns false (uniformly)
// for an empty or inconsistent chunk.
template <bool IDENT>
__device__ __forceinline__ bool scat_stage(const ScatChunk<ScatLds<IDENT>::kPer>& S, const Bufs& B,
                                           unsigned char* smem)
{
    using L = ScatLds<IDENT>;
    const uint32_t tid = scat_tid();
    key64* s_key = reinterpret_cast<key64*>(smem + L::kKeyOff);
    uint32_t* s_idx = reinterpret_cast<uint32_t*>(smem + L::kIdxOff);
    uint16_t* s_pos = reinterpret_cast<uint16_t*>(smem + L::kIdxOff);
    uint16_t* s_bkt = reinterpret_cast<uint16_t*>(smem + L::kBktOff);
    uint32_t* s_start = reinterpret_cast<uint32_t*>(smem + L::kStartOff);
    uint32_t* s_base = reinterpret_cast<uint32_t*>(smem + L::kBaseOff);
    uint32_t* s_tmp = reinterpret_cast<uint32_t*>(smem + L::kTmpOff);
    __syncthreads();     // the previous chunk's LDS image has been written out
    const uint32_t m = S.m;
    if (m == 0) return false;
    const Item seg = S.D.seg;
    const uint32_t nbkt = (2u << S.D.k) - 1;
    const uint32_t segend = S.hseg + seg.size;
    const uint32_t last = ((nbkt - 1) * S.D.T + (S.D.T - 1)) * kWaves + (kWaves - 1);
    const uint32_t b0 = tid, b1 = tid + kBigWG;
    uint32_t c0 = 0, c1 = 0;
    if (b0 < nbkt) {
        const uint32_t e = (b0 * S.D.T + S.t) * kWaves + S.w;
        c0 = (e == last ? segend : S.h0n) - S.h0;
        s_base[b0] = S.h0 - S.hseg;
    }
    if (b1 < nbkt) {
        const uint32_t e = (b1 * S.D.T + S.t) * kWaves + S.w;
        c1 = (e == last ? segend : S.h1n) - S.h1;
        s_base[b1] = S.h1 - S.hseg;
    }
    // local bucket starts: buckets 0..1023 first, then 1024..2046
    uint32_t tot0, tot1;
    const uint32_t ex0 = block_excl_scan<kBigWaves>(c0, &tot0, s_tmp);
    const uint32_t ex1 = block_excl_scan<kBigWaves>(c1, &tot1, s_tmp);
    if (b0 < nbkt) s_start[b0] = ex0;
    if (b1 < nbkt) s_start[b1] = tot0 + ex1;
    __syncthreads();
    if (tot0 + tot1 != m) {   // the histogram must account for every key of the chunk
        if (tid == 0) flag_error(B.cnt, ERR_BOUNDS);
        return false;
    }
#pragma unroll
    for (int e = 0; e < L::kPer; ++e) {
        const uint32_t p = e * kBigWG + tid;
        if (p >= m) continue;
        const uint32_t b = S.br[e] >> kRankBits;
        const uint32_t l = b < nbkt ? s_start[b] + (S.br[e] & kRankMask) : 0xFFFFFFFFu;
        if (l >= m) { flag_error(B.cnt, ERR_BOUNDS); continue; }
        s_key[l] = S.kv[e];
        if (IDENT) s_pos[l] = (uint16_t)p;
        else s_idx[l] = (seg.flags & F_IDENT) ? seg.begin + S.cb + p : S.iv[e];
        s_bkt[l] = (uint16_t)b;
    }
    __syncthreads();
    return true;
}

// Write a staged chunk of m keys of segment seg (chunk base cb) out, position
// by position.
template <bool IDENT>
__device__ __forceinline__ void scat_write(const Item& seg, uint32_t m, uint32_t cb, const Bufs& B,
                                           unsigned char* smem)
{
    using L = ScatLds<IDENT>;
    const uint32_t tid = scat_tid();
    const key64* s_key = reinterpret_cast<const key64*>(smem + L::kKeyOff);
    const uint32_t* s_idx = reinterpret_cast<const uint32_t*>(smem + L::kIdxOff);
    const uint16_t* s_pos = reinterpret_cast<const uint16_t*>(smem + L::kIdxOff);
    const uint16_t* s_bkt = reinterpret_cast<const uint16_t*>(smem + L::kBktOff);
    const uint32_t* s_start = reinterpret_cast<const uint32_t*>(smem + L::kStartOff);
    const uint32_t* s_base = reinterpret_cast<const uint32_t*>(smem + L::kBaseOff);
    key64* KD = (seg.flags & F_BUFB) ? B.key0 : B.key1;
    uint32_t* ID = (seg.flags & F_BUFB) ? B.idx0 : B.idx1;
#pragma unroll
    for (int e = 0; e < L::kPer; ++e) {
        const uint32_t j = e * kBigWG + tid;
        if (j >= m) continue;
        const uint32_t b = s_bkt[j];
        const uint32_t rel = s_base[b] + (j - s_start[b]);
        if (rel >= seg.size) { flag_error(B.cnt, ERR_BOUNDS); continue; }
        KD[seg.begin + rel] = s_key[j];
        ID[seg.begin + rel] = IDENT ? seg.begin + cb + s_pos[j] : s_idx[j];
    }
}

template <bool IDENT>
__global__ __launch_bounds__(kBigWG) void k_scatter(Plan P, Bufs B, const uint32_t* __restrict__ hscan)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const LargeDesc* __restrict__ desc = P.desc;
    const TileRef* __restrict__ tiles = P.tiles;
    const uint32_t nchunks = P.hdr->ntiles * kWaves;
    const uint32_t per = gridDim.x / kXcds;      // workgroups per XCD (the grid is a multiple of kXcds)
    const uint32_t xcd = blockIdx.x % kXcds, slot = blockIdx.x / kXcds;
    const uint32_t chi = (uint32_t)((uint64_t)nchunks * (xcd + 1) / kXcds);
    uint32_t c = (uint32_t)((uint64_t)nchunks * xcd / kXcds) + slot;
    if (slot >= per || c >= chi) return;
    // one register set: once a chunk is staged in LDS its registers take the
    // next chunk's loads, which overlap the write-out (past the end of its
    // range a workgroup fetches its last chunk again, unused, so that every
    // iteration issues the same loads)
    ScatChunk<ScatLds<IDENT>::kPer> R;
    scat_fetch<IDENT>(desc, tiles, B, hscan, c, R);
    for (;;) {
        const Item seg = R.D.seg;
        const uint32_t m = R.m, cb = R.cb;
        const bool staged = scat_stage<IDENT>(R, B, smem);
        scat_fetch<IDENT>(desc, tiles, B, hscan, min(c + per, chi - 1), R);
        if (staged) scat_write<IDENT>(seg, m, cb, B, smem);
        c += per;
        if (c >= chi) break;
    }
}

// Emit the children buckets of a multi-workgroup step.  start(b) =
// bstart[b*stride] - base relative to seg.begin; the last bucket ends at
// seg.size.  Tiny buckets become tiny leaves, other equal buckets (odd b)
// equal-key leaves in chunks, other less-than buckets blocks or large
// segments of the next level (distribute_finished, .hpp:1492-1584).
//
// spl: the segment's sorted splitters; equal bucket b holds the key spl[b >> 1],
// and one whose key holds no NUL continues into the next round (counted).
constexpr int kChildScan = 32;     // scan scratch of emit_children (at least kWaves * 7 entries)

__device__ void emit_children(const Item& seg, int nbkt, const uint32_t* bstart, uint32_t stride, uint32_t base,
                              const key64* spl, const Sink& s, uint32_t* s_tmp)
{
    constexpr int kPer = (kMaxBkt + kWG - 1) / kWG;   // 8
    const uint32_t dstbuf = (seg.flags & F_BUFB) ^ F_BUFB;
    uint32_t c_t = 0, c_l = 0, c_r = 0, c_b[kBlkClasses] = {}, c_cont = 0;
    uint32_t st[kPer], sz[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int b = tid_x() * kPer + j;
        st[j] = 0;
        sz[j] = 0;
        if (b < nbkt) {
            st[j] = bstart[(uint32_t)b * stride] - base;
            const uint32_t en = (b + 1 < nbkt) ? bstart[(uint32_t)(b + 1) * stride] - base : seg.size;
            sz[j] = en - st[j];
            const uint32_t m = sz[j];
            if (m == 0) continue;
            if (m <= kTinyT) c_t++;
            else if (b & 1) {
                c_r += (m + kRunChunk - 1) / kRunChunk;
                if ((spl[b >> 1] & 0xFFull) != 0) c_cont++;
            }
            else if (m <= kBlkT) c_b[block_class(m)]++;
            else c_l++;
        }
    }
    // the list offsets of this thread's children: one scan of all seven
    // counts (tiny, large, equal-key chunks, the block classes), then one
    // atomic per list for the segment's place in it
    constexpr int kLists = 3 + kBlkClasses;
    static_assert(kWaves * kLists <= kChildScan, "scan scratch");
    uint32_t cv[kLists], ov[kLists], tv[kLists];
    cv[0] = c_t;
    cv[1] = c_l;
    cv[2] = c_r;
#pragma unroll
    for (int c = 0; c < kBlkClasses; ++c) cv[3 + c] = c_b[c];
    block_excl_scan_k<kWaves, kLists>(cv, ov, tv, s_tmp);
    if (c_cont) atomicAdd(&s.cnt[s.c_runcont], c_cont);
    uint32_t* s_base = s_tmp + kChildScan;
    if (tid_x() == 0) {
        s_base[0] = tv[0] ? atomicAdd(&s.cnt[s.c_tiny], tv[0]) : 0;
        s_base[1] = tv[1] ? atomicAdd(&s.cnt[s.c_large], tv[1]) : 0;
        s_base[2] = tv[2] ? atomicAdd(&s.cnt[s.c_run], tv[2]) : 0;
#pragma unroll
        for (int c = 0; c < kBlkClasses; ++c)
            s_base[3 + c] = tv[3 + c] ? atomicAdd(&s.cnt[s.c_block + 2 * c], tv[3 + c]) : 0;
    }
    __syncthreads();
    uint32_t o_t = ov[0] + s_base[0], o_l = ov[1] + s_base[1], o_r = ov[2] + s_base[2], o_b[kBlkClasses];
#pragma unroll
    for (int c = 0; c < kBlkClasses; ++c) o_b[c] = ov[3 + c] + s_base[3 + c];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int b = tid_x() * kPer + j;
        const uint32_t m = sz[j];
        if (b >= nbkt || m == 0) continue;
        Item it;
        it.begin = seg.begin + st[j];
        it.size = m;
        it.rbeg = seg.rbeg;
        it.flags = dstbuf;
        if (m <= kTinyT) {
            put_item(s.tiny, s.cap_tiny, o_t++, it, s.cnt);
        } else if (b & 1) {
            const uint32_t nch = (m + kRunChunk - 1) / kRunChunk;
            for (uint32_t c = 0; c < nch; ++c) {
                Item r = it;
                r.flags = dstbuf | F_SORTED | (c << kFlagShift);
                put_item(s.run, s.cap_run, o_r++, r, s.cnt);
            }
        } else if (m <= kBlkT) {
            const int c = block_class(m);
            put_item(s.block + (uint64_t)c * s.cap_block, s.cap_block, o_b[c]++, it, s.cnt);
        } else {
            put_item(s.large, s.cap_large, o_l++, it, s.cnt);
        }
    }
}

__global__ __launch_bounds__(kWG) void k_children(Plan P, const uint32_t* hscan, const key64* trees, Sink s)
{
    __shared__ uint32_t s_tmp[kChildScan + 8];     // the counts' scan, then the lists' bases
    const uint32_t nlarge = P.hdr->nlarge;
    for (uint32_t si = blockIdx.x; si < nlarge; si += gridDim.x) {
        const LargeDesc D = P.desc[si];
        const uint32_t* h = hscan + D.hist_base;
        emit_children(D.seg, (2 << D.k) - 1, h, D.T * kWaves, h[0], trees + (uint64_t)si * kTreeSlots + 1024, s,
                      s_tmp);
        __syncthreads();     // s_tmp is reused by the next segment
    }
}

// ---- leaves --------------------------------------------------------------------

// Finish one sorted position at depth d: LCP with the predecessor when the
// keys differ; inside a run of equal keys the LCP is known when the key holds
// a NUL (fill_lcp with lcpKeyDepth, .hpp:1546-1547), otherwise the run moves
// on to depth d+8 and the string's next key is gathered for round r+1.
__device__ __forceinline__ void finish_position(const Bufs& B, uint32_t pos, uint32_t id, key64 k, bool has_prev,
                                                key64 pk, bool run_single, uint32_t depth)
{
    const bool eqp = has_prev && pk == k;
    if (has_prev && !eqp && B.lcp) B.lcp[pos] = depth + lcp_key(pk, k);
    const bool term = (k & 0xFFull) == 0;
    if (run_single || term) {
        B.perm[pos] = id;
        if (eqp && B.lcp) B.lcp[pos] = depth + key_depth(k);
    } else {
        B.idx0[pos] = id;
        if (id < B.n) B.key0[pos] = load_key(B.arena, B.alen, str_off(B, id) + depth + 8);
        else flag_error(B.cnt, ERR_BOUNDS);
    }
}

// LDS exchange between the lanes of one wave: its LDS operations execute in
// order, so keeping the compiler from reordering them is enough.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// One wave finishes a leaf of m <= 64 strings at positions [begin, begin+m)
// completely: a rank sort by broadcast compares (the slot each string takes
// in an LCP insertion sort, lcp_inssort.hpp:36-187) at depth `cur`, then,
// while runs of equal keys without a NUL remain, their strings fetch their
// next 8 characters and every such run is re-sorted in place.  Runs never
// leave the wave, so deep shared prefixes cost one gather per string per 8
// characters and no work-list traffic (the reference likewise finishes small
// buckets by insertion sort).  k, id: lane's key at `cur` and input index
// (lanes >= m ignored).  s_k, s_i, s_r: this wave's 64-entry LDS arrays.
// write_bkey: record the round-depth keys of the two end positions for the
// leaf-boundary pass.  All 64 lanes must call this.
//
// The wave may also finish several independent runs at once (packed runs of
// a block): slots [run0, run0 + length) of each run are contiguous, `pos` is
// the position this lane's slot stands for (consecutive within a run), and
// every lane of a run passes the run's first slot as run0.  A single leaf is
// run0 = 0, pos = begin + lane.
// NK: keys fetched per string at a time (1..8), or 0 for B.leaf_keys.
template <int NK>
__device__ void wave_finish_leaf(const Bufs& B, uint32_t pos, uint32_t m, key64 k, uint32_t id, uint32_t cur,
                                 bool write_bkey, uint32_t run0, key64* s_k, uint32_t* s_i, uint32_t* s_r)
{
    const uint32_t lane = lane_id();
    const bool act = lane < m;
    // every string starts in its run; (key, slot) orders a run stably
    uint32_t run = run0;
    bool live = act;         // position still has an undecided successor relation
    wave_lds_sync();
    if (act) {
        s_k[lane] = k;
        s_r[lane] = run0;
    }
    bool first_pass = write_bkey;
    constexpr int kLeafKeys = 8;
    key64 kbuf[kLeafKeys] = {};     // string bid's next keys (depths cur+8, cur+16, ...)
    // keys fetched per string at a time (the tuned key width of leaf passes)
    const int nk = NK ? NK : (int)min(max(B.leaf_keys, 1u), (uint32_t)kLeafKeys);
    uint32_t bid = 0xFFFFFFFFu;
    int nbuf = 0;
    for (;;) {
        // re-sort every live run by (key, slot); finished positions keep their slot
        wave_lds_sync();
        uint32_t rank = lane;
        if (live) {
            rank = run;
            for (uint32_t j = run; j < m && s_r[j] == run; ++j) {
                const key64 kj = s_k[j];
                rank += (kj < k || (kj == k && j < lane)) ? 1u : 0u;
            }
        }
        wave_lds_sync();
        if (live) {
            s_k[rank] = k;
            s_i[rank] = id;
        } else if (act) {
            s_i[lane] = id;
        }
        wave_lds_sync();
        if (act) {
            k = s_k[lane];
            id = s_i[lane];
        }
        wave_lds_sync();
        // neighbours inside the same run at this depth
        key64 pk = 0;
        bool eqp = false, eqn = false, same_prev = false;
        if (act && live) {
            same_prev = lane > 0 && s_r[lane - 1] == run;
            if (same_prev) pk = s_k[lane - 1];
            eqp = same_prev && pk == k;
            eqn = lane + 1 < m && s_r[lane + 1] == run && s_k[lane + 1] == k;
        }
        const bool term = (k & 0xFFull) == 0;
        if (act && live) {
            if (first_pass && (lane == 0 || lane == m - 1)) B.bkey[pos] = k;
            if (same_prev && !eqp && B.lcp) B.lcp[pos] = cur + lcp_key(pk, k);
            if (eqp && term && B.lcp) B.lcp[pos] = cur + key_depth(k);
            if ((!eqp && !eqn) || term) {
                B.perm[pos] = id;     // decided: a single string or identical strings
                live = false;
            }
        }
        // new runs: a live position starts one where it differs from its predecessor
        const uint64_t starts = __ballot(act && live && !eqp);
        const uint64_t alive = __ballot(act && live);
        if (alive == 0) break;
        if (live) {
            const uint64_t below = starts & ((2ull << lane) - 1);   // start bits at or below this lane
            run = 63u - (uint32_t)__clzll((long long)below);
        }
        wave_lds_sync();
        if (act) s_r[lane] = live ? run : 0xFFFFFFFFu;
        // the next 8 characters of every string still in a run: fetched
        // nk keys at a time, so deep shared prefixes cost one memory round
        // trip per 8 * nk characters.  A lane's buffer belongs
        // to the string it was loaded for (bid); when a re-sort moved any
        // string to another lane, the wave loads afresh.
        first_pass = false;
        cur += 8;
        if (nbuf == 0 || __ballot(live && id != bid) != 0) {     // (uniform)
            if (live) {
                if (id < B.n) load_keys<kLeafKeys>(B.arena, B.alen, str_off(B, id) + cur, kbuf, nk);
                else { flag_error(B.cnt, ERR_BOUNDS); for (int i = 0; i < kLeafKeys; ++i) kbuf[i] = 0; }
                bid = id;
            }
            nbuf = nk;
            // depths where, in every run, all strings agree with the run's
            // first one and none ends decide nothing: skip them together
            // (identical strings then cost one pass per 8 * (nk - 1) characters)
            int skip = nk - 1;
#pragma unroll
            for (int i = 0; i < kLeafKeys - 1; ++i) {
                if (i >= nk - 1) break;     // (uniform)
                const uint32_t lo = __shfl((uint32_t)kbuf[i], (int)run, 64);
                const uint32_t hi = __shfl((uint32_t)(kbuf[i] >> 32), (int)run, 64);
                const bool same = kbuf[i] == (((uint64_t)hi << 32) | lo) && (kbuf[i] & 0xFFull) != 0;
                if (skip == nk - 1 && __ballot(live && !same) != 0) skip = i;     // (uniform)
            }
            for (int s = 0; s < skip; ++s) {     // (uniform)
#pragma unroll
                for (int i = 0; i + 1 < kLeafKeys; ++i) kbuf[i] = kbuf[i + 1];
            }
            nbuf -= skip;
            cur += 8 * (uint32_t)skip;
        }
        k = kbuf[0];
#pragma unroll
        for (int i = 0; i + 1 < kLeafKeys; ++i) kbuf[i] = kbuf[i + 1];
        --nbuf;
        if (live) s_k[lane] = k;
    }
}

// One wave per tiny leaf (at most 64 strings), finished completely.
template <int NK>
__global__ __launch_bounds__(kWG) void k_leaf_tiny(const Item* items, const uint32_t* cnt_items, uint32_t cap,
                                                   Bufs B, uint32_t depth)
{
    __shared__ key64 s_k[kWaves][64];
    __shared__ uint32_t s_i[kWaves][64];
    __shared__ uint32_t s_r[kWaves][64];     // first sorted position of the run holding a position
    const uint32_t w = wave_id(), lane = lane_id();
    if (B.pending && *B.pending) {      // the round's levels are not finished
        if (B.pending_err && tid_x() == 0) flag_error(B.cnt, ERR_LEVELS);
        return;
    }
    const uint32_t nitems = min(*cnt_items, cap);     // (an overflow has been flagged)
    // each wave takes leaves it, it + stride, ... (no workgroup barriers below)
    for (uint32_t it = blockIdx.x * kWaves + w; it < nitems; it += gridDim.x * kWaves) {
        Item L = items[it];
        L.begin = (uint32_t)__builtin_amdgcn_readfirstlane((int)L.begin);
        L.size = (uint32_t)__builtin_amdgcn_readfirstlane((int)L.size);
        L.flags = (uint32_t)__builtin_amdgcn_readfirstlane((int)L.flags);
        if (!item_ok(L, B) || L.size > kTinyT) {
            if (lane == 0) flag_error(B.cnt, ERR_BOUNDS);
            continue;      // (wave-uniform)
        }
        const uint32_t m = L.size;
        const int src = L.flags & F_BUFB;
        key64 k = ~0ull;
        uint32_t id = 0;
        if (lane < m) {
            k = src ? B.key1[L.begin + lane] : B.key0[L.begin + lane];
            id = (L.flags & F_IDENT) ? L.begin + lane : src ? B.idx1[L.begin + lane] : B.idx0[L.begin + lane];
        }
        wave_finish_leaf<NK>(B, L.begin + lane, m, k, id, depth, true, 0u, s_k[w], s_i[w], s_r[w]);
    }
}

// Sorting network over E (key, position) pairs in one thread's registers
// (Batcher's odd-even merge sort: 5 comparators for 4, 19 for 8).  Positions
// are distinct, so the (key, position) order is total and the result stable.
template <int E>
__device__ __forceinline__ void thread_sort(key64 (&k)[E], uint32_t (&p)[E])
{
    static_assert(E == 4 || E == 8, "networks for 4 and 8 elements");
    constexpr int n4[][2] = { { 0, 1 }, { 2, 3 }, { 0, 2 }, { 1, 3 }, { 1, 2 } };
    constexpr int n8[][2] = { { 0, 1 }, { 2, 3 }, { 4, 5 }, { 6, 7 }, { 0, 2 }, { 1, 3 }, { 4, 6 },
                              { 5, 7 }, { 1, 2 }, { 5, 6 }, { 0, 4 }, { 1, 5 }, { 2, 6 }, { 3, 7 },
                              { 2, 4 }, { 3, 5 }, { 1, 2 }, { 3, 4 }, { 5, 6 } };
    constexpr int NC = E == 4 ? 5 : 19;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int a = E == 4 ? n4[c][0] : n8[c][0], b = E == 4 ? n4[c][1] : n8[c][1];
        const bool sw = kp_less(k[b], p[b], k[a], p[a]);
        const key64 ka = k[a], kb = k[b];
        const uint32_t pa = p[a], pb = p[b];
        k[a] = sw ? kb : ka;
        k[b] = sw ? ka : kb;
        p[a] = sw ? pb : pa;
        p[b] = sw ? pa : pb;
    }
}

constexpr uint32_t kBlockRuns = 64;     // short runs a block finishes itself (per batch of runs)

// Long runs (> kTinyT strings) a block in tail mode keeps to sort itself at
// their next depth.  Pending runs are disjoint intervals of the block, so
// at most kBlkT / (kTinyT + 1) = 31 wait at once.
constexpr uint32_t kTailStack = 32;
struct TailStack {
    uint32_t pos[kTailStack], len[kTailStack], depth[kTailStack];
    uint32_t n;
};
// keys per thread of the four block classes (256, 512, 1024, 2048 keys)
constexpr int kBlkE0 = 4, kBlkE1 = 4, kBlkE2 = 4, kBlkE3 = 4;

// Finish a block's short runs of equal keys (each at most 64 strings, at
// sorted positions beg + (s_runs[r] >> 16)) at depth + 8: the runs are
// packed into batches of at most 64 strings, and a wave finishes all runs of
// a batch in one pass (duplicates come in many short runs).  The batch
// table lives in s_k after the waves' leaf arrays.
template <int NK>
__device__ __forceinline__ void block_finish_runs(const Bufs& B, key64* s_k, const uint32_t* s_runs,
                                                  uint32_t nruns, uint32_t beg, uint32_t depth, uint32_t nwaves)
{
    const uint32_t tid = tid_x(), w = wave_id(), lane = lane_id();
    // the sorted block is no longer needed: each wave's leaf arrays live in
    // s_k, and after them the run batches
    uint32_t* s_rslot = reinterpret_cast<uint32_t*>(s_k + nwaves * 128);      // first slot of a run in its batch
    uint32_t* s_bfirst = s_rslot + kBlockRuns;                                 // first run of each batch
    uint32_t* s_nbatch_p = s_bfirst + kBlockRuns + 1;
    // pack the runs into batches of at most 64 strings: a wave finishes all
    // runs of a batch in one pass (duplicates come in many short runs)
    if (tid == 0) {
        uint32_t nb = 0, fill = 64;
        for (uint32_t r = 0; r < nruns; ++r) {
            const uint32_t len = s_runs[r] & 0xFFFFu;
            if (fill + len > 64) {
                s_bfirst[nb++] = r;
                fill = 0;
            }
            s_rslot[r] = fill;
            fill += len;
        }
        s_bfirst[nb] = nruns;
        *s_nbatch_p = nb;
    }
    __syncthreads();
    key64* wk = s_k + w * 128;
    uint32_t* wi = reinterpret_cast<uint32_t*>(wk + 64);
    uint32_t* wr = wi + 64;
    const uint32_t nbatch = *s_nbatch_p;
    for (uint32_t bt = w; bt < nbatch; bt += nwaves) {
        const uint32_t r0 = s_bfirst[bt], r1 = s_bfirst[bt + 1];
        uint32_t slot0 = 0, total = 0, rs = 0;
        bool mine = false;
        for (uint32_t r = r0; r < r1; ++r) {     // the run whose slots hold this lane
            const uint32_t so = s_rslot[r], len = s_runs[r] & 0xFFFFu;
            if (lane >= so && lane < so + len) {
                mine = true;
                slot0 = so;
                rs = s_runs[r] >> 16;
            }
            total = max(total, so + len);
        }
        key64 kk = 0;
        uint32_t id = 0, pos = 0;
        if (mine) {
            pos = beg + rs + (lane - slot0);
            kk = B.key0[pos];
            id = B.idx0[pos];
        }
        wave_finish_leaf<NK>(B, pos, total, kk, id, depth + 8, false, slot0, wk, wi, wr);
    }
}

// Merge sort of a block's m keys, in s_k in position order (padding ~0 up
// to the next multiple of E sorts last): each of MAXM / E threads sorts E
// consecutive (key, position) pairs with a sorting network, then runs are
// merged pairwise by merge path: a thread finds where its E outputs start in
// the two runs by binary search (co-rank) and merges them sequentially; on
// equal keys the left run goes first, so the order stays stable.  Leaves the
// sorted keys in s_k and their positions in s_p.
template <uint32_t MAXM, int E>
__device__ __forceinline__ void block_merge_sort(uint32_t m, key64* s_k, uint16_t* s_p)
{
    constexpr uint32_t T = MAXM / E;
    (void)T;
    const uint32_t tid = tid_x();
    const uint32_t Ta = (m + E - 1) / E, Pm = Ta * E;   // threads in use, padded size
    key64 k[E];
    uint32_t p[E];
    const uint32_t o = tid * E;
    if (tid < Ta) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            k[e] = s_k[o + e];
            p[e] = o + e;
        }
        thread_sort<E>(k, p);
    }
    __syncthreads();
    if (tid < Ta) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            s_k[o + e] = k[e];
            s_p[o + e] = (uint16_t)p[e];
        }
    }
    __syncthreads();
    for (uint32_t Lr = E; Lr < Pm; Lr <<= 1) {
        if (tid < Ta) {
            const uint32_t a0 = (o / (2 * Lr)) * (2 * Lr), la = min(Lr, Pm - a0);
            const uint32_t b0 = a0 + Lr, lb = b0 < Pm ? min(Lr, Pm - b0) : 0u;
            const uint32_t diag = o - a0;
            uint32_t lo = diag > lb ? diag - lb : 0u, hi = min(diag, la);
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_k[a0 + mid] <= s_k[b0 + diag - 1 - mid]) lo = mid + 1;
                else hi = mid;
            }
            // the heads as byte offsets into s_k, and the runs' ends
            uint32_t pa = (a0 + lo) * 8u, pb = (b0 + diag - lo) * 8u;
            const uint32_t ea = (a0 + la) * 8u, eb = (b0 + lb) * 8u;
            const unsigned char* kbase = reinterpret_cast<const unsigned char*>(s_k);
            const unsigned char* pbase = reinterpret_cast<const unsigned char*>(s_p);
            key64 ka = pa < ea ? *reinterpret_cast<const key64*>(kbase + pa) : 0ull;
            key64 kb = pb < eb ? *reinterpret_cast<const key64*>(kbase + pb) : 0ull;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                // branch-free step: the consumed run's next head is the slot
                // after the one taken (past a run's end it is never used:
                // take_a tests pa < ea and pb >= eb; s_k has one spare slot)
                const bool take_a = pb >= eb || (pa < ea && ka <= kb);
                const uint32_t src = take_a ? pa : pb;
                k[e] = take_a ? ka : kb;
                p[e] = *reinterpret_cast<const uint16_t*>(pbase + (src >> 2));
                const key64 nx = *reinterpret_cast<const key64*>(kbase + src + 8u);
                pa += take_a ? 8u : 0u;
                pb += take_a ? 0u : 8u;
                ka = take_a ? nx : ka;
                kb = take_a ? kb : nx;
            }
        }
        __syncthreads();
        if (tid < Ta) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                s_k[o + e] = k[e];
                s_p[o + e] = (uint16_t)p[e];
            }
        }
        __syncthreads();
    }
}

// Radix partition of a block.  The keys of a block lie between two
// neighbouring splitters, so their offsets from the block's smallest key
// usually fit 48 bits; then v = offset << 16 | position orders the block
// exactly as (key, position) does, in one 64-bit word.  Values are grouped
// by their top log2(MAXM) offset bits and ranked inside their groups.  A
// thread's ranking work grows with the sizes of the groups its slots fall
// in, so a block is merge-sorted instead when its offsets need more than 48
// bits, a group holds more than kRadixMaxGroup keys, or the key-weighted
// mean group size (sum of squared group sizes over the keys) exceeds
// kRadixMeanGroup.
constexpr uint32_t kRadixMaxGroup = 128;
constexpr uint32_t kRadixMeanGroup = 16;

// Exclusive block scan of a, the block totals of a and b and the block
// maximum of b (one uint32 each per thread, NW waves), with one barrier.  The
// caller writes s_tmp again only after another barrier.
template <int NW>
__device__ __forceinline__ uint32_t block_scan_sum_max(uint32_t a, uint32_t b, uint32_t* tot_a, uint32_t* tot_b,
                                                       uint32_t* max_b, uint32_t* s_tmp)
{
    const uint32_t lane = lane_id(), w = wave_id();
    uint32_t x = a, sb = b, mb = b;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        sb += (uint32_t)__shfl_xor((int)sb, d, 64);
        mb = max(mb, (uint32_t)__shfl_xor((int)mb, d, 64));
    }
    if (lane == 63) {
        s_tmp[w] = x;
        s_tmp[NW + w] = sb;
        s_tmp[2 * NW + w] = mb;
    }
    __syncthreads();
    uint32_t wpre = 0, ta = 0, tb = 0, mx = 0;
#pragma unroll 1
    for (int i = 0; i < NW; ++i) {     // (not unrolled: few registers for the 8-wave class)
        const uint32_t t = s_tmp[i];
        wpre += (i < (int)w) ? t : 0u;
        ta += t;
        tb += s_tmp[NW + i];
        mx = max(mx, s_tmp[2 * NW + i]);
    }
    *tot_a = ta;
    *tot_b = tb;
    *max_b = mx;
    return wpre + x - a;
}

__device__ __forceinline__ key64 shfl_xor64(key64 v, int d)
{
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, d, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), d, 64);
    return ((key64)hi << 32) | lo;
}

// Radix-partitioned rank sort of a block's m keys (position i = key K[beg+i]):
// values v are grouped in LDS by d = v >> (16 + sh), the top log2(MAXM)
// bits of the offset (slots from LDS atomics, in no particular order), and
// each value's sorted slot is its group's start plus the group members below
// it -- exact, so the result is the stable sort.  Returns false when the
// block is to be merge-sorted; s_k then holds the keys in position order,
// padded with ~0 up to the next multiple of E, for block_merge_sort.  On
// success s_k holds the sorted keys and s_p their positions.  s_cnt: MAXM +
// 2 counters (group counts, the total, a spare for lanes past m).
template <uint32_t MAXM, int E>
__device__ __forceinline__ bool block_radix_sort(const key64* __restrict__ K, uint32_t beg, uint32_t m,
                                                 key64* s_k, uint16_t* s_p, uint32_t* s_cnt, uint32_t* s_tmp)
{
    constexpr uint32_t T = MAXM / E;
    constexpr int NW = (int)(T / 64);
    constexpr int D = __builtin_ctz(MAXM);      // digit bits: MAXM groups
    static_assert((MAXM & (MAXM - 1)) == 0 && MAXM / T == (uint32_t)E, "a power-of-two block, E counters per thread");
    // (through an opaque move: indices derived from it are recomputed per
    // block instead of being hoisted out of a tail workgroup's loop)
    const uint32_t tid = scat_tid();
    const uint32_t Pm = (m + E - 1) / E * E;
    // w: this thread's keys (positions e*T + tid), later their packed values
    // every load is unconditional (clamped index), so all E are in flight at once
    key64 w[E];
#pragma unroll
    for (int e = 0; e < E; ++e) w[e] = K[beg + min((uint32_t)e * T + tid, m - 1)];
    key64 kmin = ~0ull, kmax = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        kmin = min(kmin, w[e]);     // (clamped copies of the last key change neither)
        kmax = max(kmax, w[e]);
    }
    for (uint32_t j = tid; j < MAXM + 2; j += T) s_cnt[j] = 0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        kmin = min(kmin, shfl_xor64(kmin, d));
        kmax = max(kmax, shfl_xor64(kmax, d));
    }
    key64* s_mm = reinterpret_cast<key64*>(s_tmp);      // [NW] minima, then [NW] maxima
    if (lane_id() == 0) {
        s_mm[wave_id()] = kmin;
        s_mm[NW + wave_id()] = kmax;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        kmin = min(kmin, s_mm[i]);
        kmax = max(kmax, s_mm[NW + i]);
    }
    // (s_tmp is written again only after the counting barrier below)
    const key64 range = kmax - kmin;
    if ((range >> 48) != 0) {     // (uniform) offsets too wide to pack: merge sort
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t i = (uint32_t)e * T + tid;
            if (i < Pm) s_k[i] = i < m ? w[e] : ~0ull;
        }
        __syncthreads();
        return false;
    }
    // the packed values, and the shift that leaves a value's group
    const int sh = 16 + max(0, 64 - (int)__clzll((long long)(range | 1)) - D);
    // (positions past m count into the spare counter s_cnt[MAXM + 1])
    uint32_t slot[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t i = (uint32_t)e * T + tid;
        w[e] = ((w[e] - kmin) << 16) | i;
        slot[e] = atomicAdd(&s_cnt[i < m ? (uint32_t)(w[e] >> sh) : MAXM + 1], 1u);
    }
    __syncthreads();
    // group starts (exclusive scan in place, E counters per thread); the
    // largest per-thread sum of squared group sizes and their total
    uint32_t sum = 0, sq = 0;
#pragma unroll
    for (int j = 0; j < E; ++j) {
        const uint32_t c = s_cnt[tid * E + j];
        sum += c;
        sq += c * c;
    }
    uint32_t tot, tsq, over;
    // (all counts are read before the scan's barrier: starts may overwrite them)
    uint32_t run = block_scan_sum_max<NW>(sum, sq, &tot, &tsq, &over, s_tmp);
#pragma unroll
    for (int j = 0; j < E; ++j) {
        const uint32_t c = s_cnt[tid * E + j];
        s_cnt[tid * E + j] = run;
        run += c;
    }
    if (tid == 0) s_cnt[MAXM] = tot;
    __syncthreads();
    if (over > kRadixMaxGroup * kRadixMaxGroup || tsq > kRadixMeanGroup * m) {   // (uniform) merge sort instead
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t i = (uint32_t)e * T + tid;
            if (i < Pm) s_k[i] = i < m ? (w[e] >> 16) + kmin : ~0ull;
        }
        __syncthreads();
        return false;
    }
    key64* s_v = s_k;     // packed values, grouped
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t i = (uint32_t)e * T + tid;
        if (i < m) s_v[s_cnt[(uint32_t)(w[e] >> sh)] + slot[e]] = w[e];
    }
    __syncthreads();
    // sorted places: thread t ranks the slots t + e T (consecutive lanes read
    // consecutive values, free of LDS bank conflicts).  Groups are ordered
    // by their top bits, so a slot's place is its group's start plus the
    // group's values below it; value and place stay in registers until every
    // value has been read.
    uint32_t place[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        if ((uint32_t)e * T >= m) break;     // (uniform) no slot of this pass is in the block
        const uint32_t j = (uint32_t)e * T + tid;
        const uint32_t jc = min(j, m - 1);
        const key64 v = s_v[jc];
        const uint32_t g = (uint32_t)(v >> sh);
        const uint32_t G1 = s_cnt[g + 1];
        uint32_t q = s_cnt[g], r = q;
        for (; q + 1 < G1; q += 2) {     // two values per step: their reads overlap
            const key64 va = s_v[q], vb = s_v[q + 1];
            r += (va < v ? 1u : 0u) + (vb < v ? 1u : 0u);
        }
        if (q < G1) r += s_v[q] < v ? 1u : 0u;
        w[e] = v;
        place[e] = r;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
        if ((uint32_t)e * T >= m) break;     // (uniform)
        if ((uint32_t)e * T + tid < m) {
            s_k[place[e]] = (w[e] >> 16) + kmin;
            s_p[place[e]] = (uint16_t)w[e];
        }
    }
    __syncthreads();
    return true;
}

// The block sort's LDS (one workgroup).
template <uint32_t MAXM, int E, bool TAIL>
struct BlockLds {
    static constexpr uint32_t T = MAXM / E;
    key64 k[MAXM + 1];     // + 1: a merge step reads the slot after the one it takes
    uint16_t p[MAXM];
    uint32_t runs[TAIL ? MAXM / 2 : kBlockRuns];      // (start << 16) | length; in tail mode every short run
    uint32_t nruns;
    uint32_t cnt[MAXM + 2];           // radix partition: group counts, then starts
    alignas(8) uint32_t tmp[4 * (T / 64) + 4];
};

// One workgroup of MAXM / E threads sorts a block of 65..MAXM strings
// completely on chip and finishes every position at this depth (the
// reference's insertion-sort base case, lcp_inssort.hpp:36-187, and the
// equal-bucket rules, .hpp:1531-1557, on sorted keys).  The sort is the
// radix partition with in-group ranks (block_radix_sort) when the block's
// keys spread well over their range, else the merge sort (block_merge_sort);
// both give the stable (key, position) order.  Finishing a sorted position:
// LCP to its neighbour, identical strings done, runs of equal keys without a
// NUL handed to round r+1.
//
// TAIL (a round whose blocks are its last work): a run longer than kTinyT
// is pushed on the workgroup's stack ts at depth + 8 instead of going to the
// next round (k_block sorts it there), and every short run is finished here.
// top: a block of the round's lists (false: a run of one, sorted deeper),
// whose first and last keys at this depth go to bkey for k_leafbound.
template <uint32_t MAXM, int E, int NK, bool TAIL>
__device__ __forceinline__ void block_one(const Item L, const Bufs& B, const Sink& nxt, uint32_t depth,
                                          BlockLds<MAXM, E, TAIL>& S, TailStack* ts, bool top)
{
    constexpr uint32_t T = MAXM / E;
    static_assert(T % 64 == 0 && MAXM <= 65536, "whole waves; positions fit 16 bits");
    // short equal runs finished here (more go to the next round; all of them in tail mode)
    constexpr uint32_t kMaxRuns = TAIL ? MAXM / 2 : kBlockRuns;
    static_assert(MAXM * 8 >= 1024 * (T / 64) + 4 * (2 * kBlockRuns + 2),
                  "the per-wave leaf arrays (1 KB each) and the run batches reuse s_k");
    key64* s_k = S.k;
    uint16_t* s_p = S.p;
    uint32_t* s_runs = S.runs;
    uint32_t& s_nruns = S.nruns;
    uint32_t* s_cnt = S.cnt;
    uint32_t* s_tmp = S.tmp;
    if (!item_ok(L, B) || L.size > MAXM) {
        if (tid_x() == 0) flag_error(B.cnt, ERR_BOUNDS);
        return;
    }
    const uint32_t m = L.size, beg = L.begin;
    uint32_t tid = scat_tid();     // (opaque: see block_radix_sort)
    const int src = L.flags & F_BUFB;
    const key64* K = src ? B.key1 : B.key0;
    const uint32_t* I = src ? B.idx1 : B.idx0;
    const bool ident = L.flags & F_IDENT;
    if (tid == 0) s_nruns = 0;
    // the block's input indices are loaded beside its keys and kept in
    // registers through the sort, then staged in s_cnt (free after either
    // sort), so that sorted positions read them from LDS instead of waiting
    // for a second global gather
    constexpr int kPerThread = MAXM / T;
    uint32_t idv[kPerThread];
    if (!ident) {
#pragma unroll
        for (int e = 0; e < kPerThread; ++e) idv[e] = I[beg + min(tid + (uint32_t)e * T, m - 1)];
    }
    if (!block_radix_sort<MAXM, E>(K, beg, m, s_k, s_p, s_cnt, s_tmp))
        block_merge_sort<MAXM, E>(m, s_k, s_p);
    tid = scat_tid();     // (a fresh copy: positions from it are not held through the sort)
    if (!ident) {     // (uniform)
#pragma unroll
        for (int e = 0; e < kPerThread; ++e) {
            const uint32_t i = tid + (uint32_t)e * T;
            if (i < m) s_cnt[i] = idv[e];
        }
        __syncthreads();
    }
    // the input index of every sorted position, read before finish_position
    // may overwrite its slot in idx0 (continuing runs store theirs)
    uint32_t ids[kPerThread];
#pragma unroll
    for (int e = 0; e < kPerThread; ++e) {
        const uint32_t i = tid + (uint32_t)e * T;
        ids[e] = 0;
        if (i < m) ids[e] = ident ? beg + s_p[i] : s_cnt[s_p[i]];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < kPerThread; ++e) {
        const uint32_t i = tid + (uint32_t)e * T;
        if ((uint32_t)e * T >= m) break;     // (uniform) no position of this pass is in the block
        const bool act = i < m;     // sorted position
        bool eqp = false, eqn = false;
        key64 kk = 0, pk = 0;
        if (act) {
            kk = s_k[i];
            if (i > 0) pk = s_k[i - 1];
            eqp = i > 0 && pk == kk;
            eqn = i + 1 < m && s_k[i + 1] == kk;
            finish_position(B, beg + i, ids[e], kk, i > 0, pk, !eqp && !eqn, depth);
            if (top && (i == 0 || i == m - 1)) B.bkey[beg + i] = kk;
        }
        // a run of equal keys without a NUL continues at depth + 8 (its keys
        // there are in place, finish_position fetched them): a run of at most
        // kTinyT strings is finished below by this workgroup, a longer one is
        // a segment of the next round
        const bool nr = act && !eqp && eqn && (kk & 0xFFull) != 0;
        if (__ballot(nr) == 0) continue;     // (uniform) nothing to append in this wave
        uint32_t rl = 0;
        if (nr) {     // the run's last position: binary search in the sorted keys
            uint32_t lo = i + 1, hi = m;     // s_k[lo] == kk, s_k[hi] != kk (or hi == m)
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_k[mid] == kk) lo = mid;
                else hi = mid;
            }
            rl = lo - i + 1;
        }
        const bool small = nr && rl <= kTinyT;
        const uint32_t slot = small ? atomicAdd(&s_nruns, 1u) : kMaxRuns;
        const bool keep = slot < kMaxRuns;
        if (keep) s_runs[slot] = (i << 16) | rl;
        if (TAIL) {
            if (nr && !small) {
                const uint32_t t = atomicAdd(&ts->n, 1u);
                if (t < kTailStack) {
                    ts->pos[t] = beg + i;
                    ts->len[t] = rl;
                    ts->depth[t] = depth + 8;
                } else {
                    flag_error(B.cnt, ERR_BOUNDS);     // (pending runs are disjoint: cannot happen)
                }
            }
        } else {
            Item ni = { beg + i, rl, beg + i, 0u };
            wave_append_segment(nr && !keep, ni, nxt);
        }
    }
    __syncthreads();     // the runs' keys and indices written above are visible to the whole group
    const uint32_t nruns = min(s_nruns, kMaxRuns);
    for (uint32_t r0 = 0; r0 < nruns; r0 += kBlockRuns) {     // (uniform) batches of at most kBlockRuns runs
        block_finish_runs<NK>(B, s_k, s_runs + r0, min(kBlockRuns, nruns - r0), beg, depth, T / 64);
        if (TAIL) __syncthreads();     // the next batch reuses the batch table
    }
}

// Blocks [first, last) of a list (one per workgroup), or with LOOP (levels
// planned on the device) from *p_first to the list's count, each workgroup
// looping over them; a count over the list's capacity has been flagged and
// is clamped.
//
// TAIL (host-planned, the round's last work: no large segments and no
// equal-key chunks are left): the workgroup finishes its block completely,
// sorting each long run again at its next depth until every string is
// decided, so the sort needs no further round.
template <uint32_t MAXM, int E, bool LOOP, int NK, bool TAIL>
__global__ __launch_bounds__(MAXM / E) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_block(const Item* items, uint32_t first, uint32_t last,
                                                     const uint32_t* p_first, const uint32_t* cnt_items, uint32_t cap,
                                                     Bufs B, Sink nxt, uint32_t depth)
{
    __shared__ BlockLds<MAXM, E, TAIL> S;
    if constexpr (TAIL) {
        static_assert(!LOOP, "tail mode is host-planned");
        __shared__ TailStack ts;
        const uint32_t idx = first + blockIdx.x;
        if (idx >= min(last, cap)) return;
        if (tid_x() == 0) ts.n = 0;
        // (every thread reads ts.n after block_one, which returns at once on
        // a malformed item: the count is stored before a barrier each time)
        __syncthreads();
        Item cur = items[idx];
        uint32_t d = depth;
        for (bool top = true;; top = false) {     // one call site, so that block_one stays inline
            block_one<MAXM, E, NK, true>(cur, B, nxt, d, S, &ts, top);
            // (the pushes precede a barrier in block_one's finish pass)
            const uint32_t n = min(ts.n, kTailStack);
            if (n == 0) break;     // (uniform)
            __syncthreads();     // the last pass's LDS is no longer read, every thread has read ts.n
            cur = Item{ ts.pos[n - 1], ts.len[n - 1], ts.pos[n - 1], 0u };
            d = ts.depth[n - 1];
            if (tid_x() == 0) ts.n = n - 1;     // (pushes into the freed entry come after block_one's barriers)
            __syncthreads();     // every thread has read the top entry; the count is stored
        }
    } else if constexpr (!LOOP) {
        const uint32_t idx = first + blockIdx.x;
        if (idx >= min(last, cap)) return;
        block_one<MAXM, E, NK, false>(items[idx], B, nxt, depth, S, nullptr, true);
    } else {
        const uint32_t lo = *p_first, hi = min(*cnt_items, cap);
        for (uint32_t idx = lo + blockIdx.x; idx < hi; idx += gridDim.x) {
            __syncthreads();     // the previous block's LDS is no longer read
            block_one<MAXM, E, NK, false>(items[idx], B, nxt, depth, S, nullptr, true);
        }
    }
}

// One workgroup per chunk of an equal-key leaf (all keys equal).
__global__ __launch_bounds__(kWG) void k_leaf_run(const Item* items, const uint32_t* cnt_items, uint32_t cap,
                                                  Bufs B, Sink nxt, uint32_t depth)
{
    if (B.pending && *B.pending) {      // the round's levels are not finished
        if (B.pending_err && tid_x() == 0) flag_error(B.cnt, ERR_LEVELS);
        return;
    }
    const uint32_t nitems = min(*cnt_items, cap);     // (an overflow has been flagged)
    for (uint32_t ii = blockIdx.x; ii < nitems; ii += gridDim.x) {
        __syncthreads();     // the previous chunk's key read is complete
        const Item L = items[ii];
        const uint32_t c = L.flags >> kFlagShift;
        if (!item_ok(L, B) || (uint64_t)c * kRunChunk >= L.size) {
            if (tid_x() == 0) flag_error(B.cnt, ERR_BOUNDS);
            continue;
        }
        const int src = L.flags & F_BUFB;
        const key64* K = src ? B.key1 : B.key0;
        const uint32_t* I = src ? B.idx1 : B.idx0;
        const uint32_t rend = L.begin + L.size;
        const uint32_t cb = L.begin + c * kRunChunk;
        const uint32_t ce = min(rend, cb + kRunChunk);
        const key64 key = K[cb];    // this chunk's own copy of the run key
        __syncthreads();
        const bool term = (key & 0xFFull) == 0;
        const bool single = L.size == 1;
        for (uint32_t p = cb + tid_x(); p < ce; p += kWG) {
            const uint32_t id = (L.flags & F_IDENT) ? p : I[p];
            if (single || term) {
                B.perm[p] = id;
                if (p > L.begin && B.lcp) B.lcp[p] = depth + key_depth(key);
            } else {
                if (src || (L.flags & F_IDENT)) B.idx0[p] = id;
                if (id < B.n) B.key0[p] = load_key(B.arena, B.alen, str_off(B, id) + depth + 8);
                else flag_error(B.cnt, ERR_BOUNDS);
            }
        }
        if (tid_x() < 64) {
            const bool nr = tid_x() == 0 && c == 0 && !term && !single;
            Item ni = { L.begin, L.size, L.begin, 0u };
            wave_append_segment(nr, ni, nxt);
            if (tid_x() == 0) {
                if (c == 0) B.bkey[L.begin] = key;
                if (ce == rend) B.bkey[rend - 1] = key;
            }
        }
    }
}

// LCP between neighbouring leaves of one round segment:
// depth + lcp_key(last key of the left leaf, first key of the right leaf)
// (sample_sort_lcp, .hpp:311-396).
struct LeafLists {        // the leaf lists of a round, concatenated in this order
    const Item* list[kBlkClasses + 2];   // tiny, blocks by class, equal-key chunks
    int cnt[kBlkClasses + 2];            // their counters (indices into the counter array)
    uint32_t cap[kBlkClasses + 2];       // their capacities (an overflow has been flagged)
};

__global__ __launch_bounds__(kWG) void k_leafbound(LeafLists LL, Bufs B, uint32_t depth)
{
    constexpr int NL = kBlkClasses + 2;
    if (B.pending && *B.pending) {      // the round's levels are not finished
        if (B.pending_err && tid_x() == 0) flag_error(B.cnt, ERR_LEVELS);
        return;
    }
    uint32_t end[NL];        // running totals of the list lengths
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < NL; ++j) end[j] = run += min(B.cnt[LL.cnt[j]], LL.cap[j]);
    for (uint32_t i = blockIdx.x * kWG + tid_x(); i < end[NL - 1]; i += gridDim.x * kWG) {
        int l = 0;
#pragma unroll
        for (int j = 0; j < NL - 1; ++j) l += i >= end[j] ? 1 : 0;
        const Item L = LL.list[l][i - (l ? end[l - 1] : 0u)];
        if (l == NL - 1 && (L.flags >> kFlagShift) != 0) continue;   // only the first chunk of an equal-key leaf
        if (L.begin <= L.rbeg || !B.lcp) continue;
        if (L.begin >= B.n) { flag_error(B.cnt, ERR_BOUNDS); continue; }
        const key64 a = B.bkey[L.begin - 1], c = B.bkey[L.begin];
        if (a == c) { flag_error(B.cnt, ERR_LCP); continue; }
        B.lcp[L.begin] = depth + lcp_key(a, c);
    }
}

// Batch outputs: input index -> index inside its set (perm[pos] - first of
// the set holding pos; a set occupies the same positions as its inputs).
// A workgroup takes kLocTile positions: two lanes find the sets holding its
// first and last position, and each position searches only between them
// (no search at all in the common case of a tile inside one set).
constexpr uint32_t kLocTile = 16 * kWG;
__global__ __launch_bounds__(kWG) void k_localize(uint32_t* perm, uint32_t n, const uint64_t* set_first,
                                                   uint32_t nsets)
{
    __shared__ uint32_t s_set[2];
    const uint32_t t0 = blockIdx.x * kLocTile;
    if (t0 >= n) return;
    const uint32_t t1 = (uint32_t)min<uint64_t>(n, (uint64_t)t0 + kLocTile);
    if (tid_x() < 2) {
        const uint32_t pos = tid_x() == 0 ? t0 : t1 - 1;
        uint32_t lo = 0, hi = nsets;   // largest s with set_first[s] <= pos
        while (lo + 1 < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (set_first[mid] <= pos) lo = mid;
            else hi = mid;
        }
        s_set[tid_x()] = lo;
    }
    __syncthreads();
    const uint32_t s0 = s_set[0], s1 = s_set[1];
    if (s0 == s1 && ((uintptr_t)perm & 15) == 0) {     // (uniform) one set: one base, 16-byte accesses
        const uint32_t base = (uint32_t)set_first[s0];
        for (uint32_t q = t0 + 4 * tid_x(); q < t1; q += 4 * kWG) {
            if (q + 4 <= t1) {
                uint4 v = *reinterpret_cast<const uint4*>(perm + q);
                v.x -= base;
                v.y -= base;
                v.z -= base;
                v.w -= base;
                *reinterpret_cast<uint4*>(perm + q) = v;
            } else {
                for (uint32_t pos = q; pos < t1; ++pos) perm[pos] -= base;
            }
        }
        return;
    }
    for (uint32_t pos = t0 + tid_x(); pos < t1; pos += kWG) {
        uint32_t lo = s0, hi = s1 + 1;
        while (lo + 1 < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (set_first[mid] <= pos) lo = mid;
            else hi = mid;
        }
        perm[pos] -= (uint32_t)set_first[lo];
    }
}

// Distinguishing characters of a sorted order: cache[i] = s_i[lcp[i]] for
// i >= 1, the first character in which the i-th string differs from its
// predecessor, NUL when they are equal (set_cache in sample_sort_lcp,
// bingmann-parallel_sample_sort.hpp:366,385; fill_lcp, stringptr.hpp:950-958;
// checked by verify_lcp_cache, stringptr.hpp:1015-1021).  lcp[i] never
// exceeds the length of s_i, so the byte read is inside the string.
__global__ __launch_bounds__(kWG) void k_dist_chars(const uint8_t* arena, const uint32_t* offs,
                                                    const uint32_t* perm, const uint32_t* lcp, uint32_t n,
                                                    uint8_t* cache, uint32_t* cnt, uint64_t alen)
{
    const uint32_t i = blockIdx.x * kWG + tid_x();
    if (i == 0 || i >= n) return;
    const uint64_t p = (uint64_t)offs[perm[i]] + lcp[i];
    if (p >= alen) { flag_error(cnt, ERR_BOUNDS); return; }
    cache[i] = arena[p];
}

// ptrs[i] = base + offs[perm[i]]: the sorted strings' host pointers
__global__ __launch_bounds__(kWG) void k_gather_ptrs(const uint32_t* perm, const uint32_t* offs, uint64_t base,
                                                     uint64_t n, uint64_t* ptrs)
{
    const uint64_t i = (uint64_t)blockIdx.x * kWG + tid_x();
    if (i < n) ptrs[i] = base + offs[perm[i] < n ? perm[i] : 0u];
}

// wide[i] = lcp[i] as uint64 (entry 0 is copied too; callers skip it)
__global__ __launch_bounds__(kWG) void k_widen_lcp(const uint32_t* lcp, uint64_t n, uint64_t* wide)
{
    const uint64_t i = (uint64_t)blockIdx.x * kWG + tid_x();
    if (i < n) wide[i] = lcp[i];
}

// offs64[i] = base[s] + offs32[i] for the set s holding string i
__global__ __launch_bounds__(kWG) void k_widen(const uint32_t* offs32, const uint64_t* first, const uint64_t* base,
                                               uint32_t nsets, uint64_t n, uint64_t* offs64)
{
    const uint64_t i = (uint64_t)blockIdx.x * kWG + tid_x();
    if (i >= n) return;
    uint32_t lo = 0, hi = nsets;   // largest s with first[s] <= i
    while (lo + 1 < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (first[mid] <= i) lo = mid;
        else hi = mid;
    }
    offs64[i] = base[lo] + offs32[i];
}

// ---- LCP-aware merge of sorted runs ---------------------------------------

// Compare strings a and b (arena offsets) that share their first h
// characters: returns true if a <= b, and their LCP in *lcp.
__device__ __forceinline__ bool str_leq_from(const uint8_t* arena, uint64_t alen, uint64_t a, uint64_t b,
                                             uint32_t h, uint32_t* lcp)
{
    for (;;) {
        const key64 ka = load_key(arena, alen, a + h), kb = load_key(arena, alen, b + h);
        if (ka != kb) {
            *lcp = h + lcp_key(ka, kb);
            return ka < kb;
        }
        if ((ka & 0xFFull) == 0) {   // both end inside this key: equal strings
            *lcp = h + key_depth(ka);
            return true;
        }
        h += 8;
    }
}

// One round of pairwise merges.  Pair p merges runs A = [first[p], mid[p])
// and B = [mid[p], first[p+1]) of (idx, lcp) into the same positions of the
// output; lcp[q] is the LCP of the strings at q-1 and q within a run (its
// first entry unused).  The output of a pair is cut into tiles of kMergeT
// positions, one workgroup each.  k_merge_part finds, for the first output
// of every tile, how many strings of A precede it (a merge-path co-rank,
// by string comparisons; equal strings take A first, so ties keep run
// order).  k_merge_tile stages the tile's strings of A and B in LDS (index,
// LCP and their first 8 characters), places each thread's kMergeE outputs by
// a co-rank inside the tile, and emits them with the binary LCP merge
// (lcp_merge_binary and lcp_compare, bingmann-lcp_mergesort_binary.hpp),
// which compares characters only where the LCPs to the previous output do
// not decide; the staged characters decide most of those compares without
// touching the arena.  The first round takes the 8 characters from keys
// extracted in input order by one streaming pass (a gather of 8 bytes per
// string instead of an offset and an arena window); every round writes them
// beside its outputs for the next.
constexpr uint32_t kMergeE = 8;                    // outputs per thread
constexpr uint32_t kMergeT = kWG * kMergeE;        // outputs per tile

// a <= b for strings a, b (input indices) whose first 8 characters are ka, kb
__device__ __forceinline__ bool merge_leq(const Bufs& B, uint32_t a, uint32_t b, key64 ka, key64 kb,
                                          uint32_t* lcp)
{
    if (ka != kb) {
        *lcp = lcp_key(ka, kb);
        return ka < kb;
    }
    if ((ka & 0xFFull) == 0) {
        *lcp = key_depth(ka);
        return true;
    }
    return str_leq_from(B.arena, B.alen, str_off(B, a), str_off(B, b), 8, lcp);
}

// first 8 characters of the string at position pos of the source: from the
// previous round's output, else from in_key (indexed by input index)
__device__ __forceinline__ key64 merge_key(const Bufs& B, const uint32_t* src_idx, const key64* src_key,
                                           const key64* in_key, uint64_t pos)
{
    if (src_key) return src_key[pos];
    const uint32_t id = src_idx[pos];
    if (id < B.n) return in_key[id];
    flag_error(B.cnt, ERR_BOUNDS);
    return 0;
}

// the pair holding tile g: the largest p with tstart[p] <= g
__device__ __forceinline__ uint32_t merge_pair(const uint64_t* tstart, uint32_t npairs, uint64_t g)
{
    uint32_t lo = 0, hi = npairs;
    while (lo + 1 < hi) {
        const uint32_t m = (lo + hi) / 2;
        if (tstart[m] <= g) lo = m;
        else hi = m;
    }
    return lo;
}

// co[g] = strings of A (of g's pair) that precede the first output of tile g
__global__ __launch_bounds__(kWG) void k_merge_part(Bufs B, const uint32_t* src_idx, const key64* src_key,
                                                    const key64* in_key,
                                                    const uint64_t* first, const uint64_t* mid,
                                                    const uint64_t* tstart, uint32_t npairs, uint32_t* co)
{
    const uint64_t g = (uint64_t)blockIdx.x * kWG + tid_x();
    if (g >= tstart[npairs]) return;
    const uint32_t p = merge_pair(tstart, npairs, g);
    const uint64_t a0 = first[p], b0 = mid[p];
    const uint64_t la = b0 - a0, lb = first[p + 1] - b0;
    const uint64_t diag = (g - tstart[p]) * kMergeT;
    uint64_t i0 = diag > lb ? diag - lb : 0, i1 = min(diag, la);
    while (i0 < i1) {
        const uint64_t mi = (i0 + i1) / 2;
        const uint64_t pa = a0 + mi, pb = b0 + diag - 1 - mi;
        uint32_t l;
        if (merge_leq(B, src_idx[pa], src_idx[pb], merge_key(B, src_idx, src_key, in_key, pa),
                      merge_key(B, src_idx, src_key, in_key, pb), &l))
            i0 = mi + 1;
        else
            i1 = mi;
    }
    co[g] = (uint32_t)i0;
}

__global__ __launch_bounds__(kWG) void k_merge_tile(Bufs B, const uint32_t* src_idx, const uint32_t* src_lcp,
                                                    const key64* src_key, const key64* in_key,
                                                    uint32_t* dst_idx, uint32_t* dst_lcp,
                                                    key64* dst_key, const uint64_t* first, const uint64_t* mid,
                                                    const uint64_t* tstart, uint32_t npairs, const uint32_t* co)
{
    // slots [0, na+1): A[ia0-1 .. ia1), slots [na+1, na+nb+2): B[jb0-1 .. jb1);
    // slot 0 of a part is unused when the part starts its run
    __shared__ key64 s_key[kMergeT + 2];
    __shared__ uint32_t s_idx[kMergeT + 2];
    __shared__ uint32_t s_lcp[kMergeT + 2];
    __shared__ uint16_t s_oslot[kMergeT];
    __shared__ uint32_t s_olcp[kMergeT];
    const uint64_t g = blockIdx.x;
    const uint32_t p = merge_pair(tstart, npairs, g);
    const uint64_t a0 = first[p], b0 = mid[p];
    const uint64_t la = b0 - a0, lb = first[p + 1] - b0;
    const uint64_t diag0 = (g - tstart[p]) * kMergeT;
    const uint32_t len = (uint32_t)min((uint64_t)kMergeT, la + lb - diag0);
    const uint64_t ia0 = co[g], ia1 = g + 1 < tstart[p + 1] ? co[g + 1] : la;
    const uint64_t jb0 = diag0 - ia0;
    const uint32_t na = (uint32_t)(ia1 - ia0), nb = len - na;
    const uint32_t sb = na + 1;        // first slot of B
    for (uint32_t k = tid_x(); k < len + 2; k += kWG) {
        uint64_t pos;
        bool valid;
        if (k < sb) {
            valid = k > 0 || ia0 > 0;
            pos = a0 + ia0 + k - 1;
        } else {
            valid = k > sb || jb0 > 0;
            pos = b0 + jb0 + (k - sb) - 1;
        }
        if (valid) {
            s_idx[k] = src_idx[pos];
            s_lcp[k] = src_lcp[pos];
            s_key[k] = merge_key(B, src_idx, src_key, in_key, pos);
        }
    }
    __syncthreads();
    const uint32_t d = tid_x() * kMergeE;
    if (d < len) {
        auto leq = [&](uint32_t x, uint32_t y, uint32_t* l) {
            return merge_leq(B, s_idx[x], s_idx[y], s_key[x], s_key[y], l);
        };
        uint32_t i0 = d > nb ? d - nb : 0, i1 = min(d, na);
        while (i0 < i1) {
            const uint32_t mi = (i0 + i1) / 2;
            uint32_t l;
            if (leq(1 + mi, sb + d - mi, &l)) i0 = mi + 1;     // A[mi] <= B[d-1-mi]
            else i1 = mi;
        }
        uint32_t i = i0, j = d - i0;
        // LCPs of the next A and B string with the previous output
        uint32_t lcpa = 0, lcpb = 0;
        const bool hasa = ia0 + i > 0, hasb = jb0 + j > 0;
        if (hasa || hasb) {
            bool prev_is_a;
            uint32_t l;
            if (!hasb) prev_is_a = true;
            else if (!hasa) prev_is_a = false;
            else prev_is_a = !leq(i, sb + j, &l);
            const uint32_t prev = prev_is_a ? i : sb + j;
            if (i < na) {
                if (prev_is_a) lcpa = s_lcp[1 + i];
                else leq(prev, 1 + i, &lcpa);
            }
            if (j < nb) {
                if (!prev_is_a) lcpb = s_lcp[sb + 1 + j];
                else leq(prev, sb + 1 + j, &lcpb);
            }
        }
        const uint32_t nout = min(kMergeE, len - d);
        for (uint32_t q = 0; q < nout; ++q) {
            bool take_a;
            uint32_t lab = 0;     // LCP of the two heads
            if (j >= nb) take_a = true;
            else if (i >= na) take_a = false;
            else if (lcpa == lcpb) {
                const uint32_t x = 1 + i, y = sb + 1 + j;
                if (lcpa < 8) take_a = leq(x, y, &lab);
                else take_a = str_leq_from(B.arena, B.alen, str_off(B, s_idx[x]), str_off(B, s_idx[y]), lcpa, &lab);
            }
            else if (lcpa > lcpb) { take_a = true; lab = lcpb; }
            else { take_a = false; lab = lcpa; }
            if (take_a) {
                s_oslot[d + q] = (uint16_t)(1 + i);
                s_olcp[d + q] = lcpa;
                ++i;
                lcpa = i < na ? s_lcp[1 + i] : 0u;
                lcpb = lab;
            } else {
                s_oslot[d + q] = (uint16_t)(sb + 1 + j);
                s_olcp[d + q] = lcpb;
                ++j;
                lcpb = j < nb ? s_lcp[sb + 1 + j] : 0u;
                lcpa = lab;
            }
        }
    }
    __syncthreads();
    const uint64_t o0 = a0 + diag0;
    for (uint32_t k = tid_x(); k < len; k += kWG) {
        const uint32_t s = s_oslot[k];
        dst_idx[o0 + k] = s_idx[s];
        dst_lcp[o0 + k] = s_olcp[k];
        if (dst_key) dst_key[o0 + k] = s_key[s];
    }
}

// plain copy of 16-byte words (calibrates the HBM byte counters)
__global__ __launch_bounds__(kWG) void k_copy16(const uint4* src, uint4* dst, uint64_t n16)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kWG + tid_x(); i < n16; i += (uint64_t)gridDim.x * kWG)
        dst[i] = src[i];
}

// sum of the sizes of a work list range (profiling only)
__global__ __launch_bounds__(kWG) void k_sum_sizes(const Item* items, uint32_t first, const uint32_t* cnt_items,
                                                    unsigned long long* out)
{
    __shared__ uint32_t s_tmp[16];
    const uint32_t n = *cnt_items;
    uint32_t v = 0;
    for (uint32_t i = first + blockIdx.x * kWG + tid_x(); i < n; i += gridDim.x * kWG) {
        const Item it = items[i];
        if (it.flags & F_SORTED) {   // a chunk of an equal-key leaf
            const uint32_t c = it.flags >> kFlagShift, cb = c * kRunChunk;
            v += cb < it.size ? min(kRunChunk, it.size - cb) : 0u;
        } else {
            v += it.size;
        }
    }
    uint32_t tot;
    block_excl_scan<kWaves>(v, &tot, s_tmp);
    if (tid_x() == 0 && tot) atomicAdd(out, (unsigned long long)tot);
}

// ===========================================================================
// host side
// ===========================================================================

namespace {

thread_local std::string g_err;

int set_err(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t _e = (x);                                                       \
        if (_e != hipSuccess)                                                      \
            return set_err(PS5_ERR_HIP, "%s failed: %s (%s:%d)", #x,               \
                           hipGetErrorString(_e), __FILE__, __LINE__);             \
    } while (0)

enum KernelId {
    K_INIT, K_TREE, K_COUNT, K_SCAN, K_SCATTER, K_CHILDREN, K_BLOCK,
    K_LEAF_TINY, K_LEAF_RUN, K_LEAFBOUND, K_MISC, K_NUM
};
const char* kKernelNames[K_NUM] = {
    "k_init", "k_tree", "k_count", "k_scan", "k_scatter", "k_children", "k_block",
    "k_leaf_tiny", "k_leaf_run", "k_leafbound", "k_misc"
};

enum { S_BLOCK, S_TINY, S_RUN, S_NUM };

// The work buffers of a sort of up to `cap` strings.  A context owns one set
// (grown with hipMalloc on demand); the allocation-free entry point swaps in
// a set carved from the static device scratch for the duration of a call.
struct WorkBuffers {
    bool fixed = false;              // carved from the static scratch: never grown or freed
    uint64_t cap = 0;
    key64* key[2] = { nullptr, nullptr };
    uint32_t* idx[2] = { nullptr, nullptr };
    key64* bkey = nullptr;
    uint32_t* bktrank = nullptr;
    // large segments: first level of a round [round parity], later levels [level parity]
    Item* l_round[2] = { nullptr, nullptr };
    Item* l_lvl[2] = { nullptr, nullptr };
    // leaves of a round [round parity]
    Item* l_tiny[2] = { nullptr, nullptr };
    Item* l_block[2] = { nullptr, nullptr };
    Item* l_run[2] = { nullptr, nullptr };
    uint32_t cap_large = 0, cap_block = 0, cap_tiny = 0, cap_run = 0;
    Item* h_large = nullptr;         // pinned copy of a large list
    // plan of a level: its nlarge descriptors followed by its tile table, in
    // one buffer so that one copy uploads both
    unsigned char* d_plan = nullptr;
    unsigned char* h_plan = nullptr;     // pinned, mapped: the host's plans are pulled by k_fetch
    unsigned char* h_plan_dev = nullptr; // its device address
    uint64_t tiles_cap = 0;
    key64* trees = nullptr;          // kTreeSlots splitter slots per large segment
    uint64_t trees_cap = 0;
    uint32_t* hist = nullptr;
    uint64_t hist_cap = 0;
    uint32_t* part = nullptr;
    uint64_t part_cap = 0;
};

// List capacities of a sort of up to cap strings (each list bounds what the
// device can append to it; an overflow is flagged, never written).
constexpr uint32_t cap_large_of(uint64_t cap) { return (uint32_t)(cap / kBlkT + 16); }
constexpr uint32_t cap_block_of(uint64_t cap) { return (uint32_t)(cap / kTinyT + 16); }
constexpr uint32_t cap_tiny_of(uint64_t cap) { return (uint32_t)(cap + 16); }
// an equal-key run of m > kTinyT strings yields ceil(m / kRunChunk) <= m / 64 chunks
constexpr uint32_t cap_run_of(uint64_t cap) { return (uint32_t)(cap / 64 + 16); }
// every large segment has at most size/(kWaves * kChunkMin) + 1 tiles
constexpr uint64_t tiles_cap_of(uint64_t cap) { return cap / ((uint64_t)kWaves * kChunkMin) + cap_large_of(cap) + 1; }
// a plan: the level header, cap_large descriptors, the tile table
constexpr uint64_t plan_tiles_off(uint64_t cap)
{
    return sizeof(LevelHdr) + (uint64_t)cap_large_of(cap) * sizeof(LargeDesc);
}
constexpr uint64_t plan_bytes_of(uint64_t cap) { return plan_tiles_off(cap) + tiles_cap_of(cap) * sizeof(TileRef); }
static_assert(sizeof(LevelHdr) % 8 == 0 && sizeof(LargeDesc) % 8 == 0 && sizeof(TileRef) % 8 == 0,
              "a plan is copied in 8-byte words");
Plan plan_view(const unsigned char* base, uint64_t cap)
{
    Plan P;
    P.hdr = reinterpret_cast<const LevelHdr*>(base);
    P.desc = reinterpret_cast<const LargeDesc*>(base + sizeof(LevelHdr));
    P.tiles = reinterpret_cast<const TileRef*>(base + plan_tiles_off(cap));
    return P;
}

}  // namespace

struct ps5_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;         // a round's leaves run here beside its last blocks
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    WorkBuffers w;                   // the buffers run_sort uses
    // the allocation-free entry point: pinned plan buffers sized for the
    // scratch's largest sort (allocated with the context), and the scratch's
    // device address on this context's device
    Item* na_h_large = nullptr;
    unsigned char* na_h_plan = nullptr;
    unsigned char* na_h_plan_dev = nullptr;
    unsigned char* na_scratch = nullptr;
    uint32_t* d_cnt = nullptr;
    uint32_t* h_cnt = nullptr;       // pinned
    // counters + head of a list, written by k_stage straight into mapped
    // host memory (no copy engine on the readback path)
    unsigned char* h_stage = nullptr;    // pinned, mapped, coherent
    unsigned char* h_stage_dev = nullptr;    // its device address
    uint32_t stage_seq = 0;                  // sequence number of the last staged copy
    uint32_t pending_root = 0;               // strings of a single root whose set-up the next plan fetch does
    uint32_t* h_upcnt = nullptr;     // pinned: counters uploaded with batch roots
    Item* h_roots = nullptr;         // pinned: root items of a batch
    uint64_t roots_cap = 0;
    uint64_t* d_merge = nullptr;     // plan of a merge round
    uint64_t merge_cap = 0;
    uint32_t scatter_wgs = 256;         // k_scatter workgroups: one per CU, a multiple of kXcds
    uint64_t* d_sets = nullptr;      // set boundaries of a batch
    uint64_t sets_cap = 0;
    unsigned long long* d_sums = nullptr;   // profiling: elements per kernel class
    unsigned long long* h_sums = nullptr;
    // tuning
    uint32_t leaf_target = 768;      // strings per block the multi-workgroup levels aim for
    uint32_t max_treebits = kMaxTreebits;   // splitter-tree depth limit of one level
    uint32_t max_oversample = 32;    // sample keys per splitter (limited by kMaxSamples)
    // tile shapes of the multi-workgroup levels and the leaf passes' key width
    uint32_t chunk_max = kChunkMax;          // keys per wave chunk, ordinary levels
    uint32_t chunk_max_root = kChunkMaxIdent;    // the same in a level of roots
    uint32_t min_tiles = kMinTiles;          // a small level's chunks shrink to give this many tiles
    uint32_t leaf_keys = 8;                  // keys a leaf pass fetches per string at a time
    uint64_t seed = 0x5EED5EEDull;
    int planning = 0;                // 0: levels planned on the host, 1: on the device (no host synchronisation
                                     // inside a round), 2: on the device for a single sort of fewer than
                                     // kAutoDevMax strings, else on the host
    // a sort's first rounds as one captured HIP graph (0: off), and the
    // instantiated graph of the last sort, replayed while its inputs match
    uint32_t graph_rounds = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_key;
    uint64_t w_gen = 0;              // bumped whenever a work buffer is allocated again
    bool tail_blocks = true;         // a round's last blocks finish every string (PS5_TAIL_BLOCKS=0: off)
    bool chain_levels = true;        // a host-planned level's successor is planned on the device, without a
                                     // read-back between them (PS5_CHAIN_LEVELS=0: off)
    uint32_t graph_levels = 0;       // non-zero: fewer levels per captured round (PS5_GRAPH_LEVELS; tests
                                     // use it to reach the host-planned fallback)
    // profiling
    bool profile = false;
    bool trace = false;              // PS5_TRACE set: per-round work to stderr (diagnostics)
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;
    double k_ms[K_NUM] = {};
    uint64_t k_launch[K_NUM] = {};
    uint64_t k_bytes[K_NUM] = {};
    uint64_t st_rounds = 0, st_levels = 0, st_segments = 0, st_leaves = 0, st_syncs = 0;
};

namespace {

void free_buffers(ps5_ctx* c)
{
    if (c->w.fixed) {     // views into the static scratch: nothing to free
        c->w = WorkBuffers();
        return;
    }
    auto f = [](void* p) { if (p) (void)hipFree(p); };
    for (int i = 0; i < 2; ++i) {
        f(c->w.key[i]); f(c->w.idx[i]); f(c->w.l_round[i]); f(c->w.l_lvl[i]);
        f(c->w.l_tiny[i]); f(c->w.l_block[i]); f(c->w.l_run[i]);
        c->w.key[i] = nullptr; c->w.idx[i] = nullptr; c->w.l_round[i] = nullptr; c->w.l_lvl[i] = nullptr;
        c->w.l_tiny[i] = nullptr; c->w.l_block[i] = nullptr; c->w.l_run[i] = nullptr;
    }
    f(c->w.bkey); f(c->w.bktrank); f(c->w.hist); f(c->w.part); f(c->w.trees); f(c->w.d_plan);
    c->w.bkey = nullptr; c->w.bktrank = nullptr; c->w.hist = nullptr; c->w.part = nullptr; c->w.trees = nullptr;
    c->w.d_plan = nullptr;
    if (c->w.h_large) (void)hipHostFree(c->w.h_large);
    if (c->w.h_plan) (void)hipHostFree(c->w.h_plan);
    c->w.h_large = nullptr; c->w.h_plan = nullptr; c->w.h_plan_dev = nullptr;
    c->w.cap = 0;
    c->w.hist_cap = 0;
    c->w.part_cap = 0;
    c->w.tiles_cap = 0;
    c->w.trees_cap = 0;
}

// Splitter trees for `nlarge` segments and `hl` histogram counters of one
// level.  Called between levels, when the stream has drained (the level's
// counts were just read back), so old buffers are idle.
int grow_level(ps5_ctx* c, uint32_t nlarge, uint64_t hl)
{
    if (c->w.fixed && (nlarge > c->w.trees_cap || hl > c->w.hist_cap))
        return set_err(PS5_ERR_CAPACITY, "a level needs %u trees and %llu histogram entries; the scratch holds "
                       "%llu and %llu", nlarge, (unsigned long long)hl, (unsigned long long)c->w.trees_cap,
                       (unsigned long long)c->w.hist_cap);
    if (nlarge > c->w.trees_cap) {
        if (c->w.trees) (void)hipFree(c->w.trees);
        c->w.trees = nullptr;
        c->w.trees_cap = 0;
        const uint64_t want = std::max<uint64_t>(nlarge + nlarge / 2, 64);
        HIPCHK(hipMalloc(&c->w.trees, want * kTreeSlots * sizeof(key64)));
        c->w.trees_cap = want;
        c->w_gen++;
    }
    if (hl > c->w.hist_cap) {
        if (c->w.hist) (void)hipFree(c->w.hist);
        if (c->w.part) (void)hipFree(c->w.part);
        c->w.hist = nullptr;
        c->w.part = nullptr;
        c->w.hist_cap = c->w.part_cap = 0;
        const uint64_t want = std::min<uint64_t>(std::max<uint64_t>(hl + hl / 2, 1u << 20), 0xFFFFFFFFull);
        HIPCHK(hipMalloc(&c->w.hist, want * sizeof(uint32_t)));
        c->w.part_cap = want / kScanBlock + 2;
        HIPCHK(hipMalloc(&c->w.part, c->w.part_cap * sizeof(uint32_t)));
        c->w.hist_cap = want;
        c->w_gen++;
    }
    return PS5_OK;
}

int ensure_capacity(ps5_ctx* c, uint64_t n)
{
    if (n <= c->w.cap) return PS5_OK;
    if (c->w.fixed)
        return set_err(PS5_ERR_CAPACITY, "%llu strings exceed the scratch's %llu", (unsigned long long)n,
                       (unsigned long long)c->w.cap);
    // a context that has to grow again grows by half at least
    uint64_t cap = std::max<uint64_t>(std::max<uint64_t>(n, 1024), c->w.cap + c->w.cap / 2);
    cap = std::min<uint64_t>(cap, std::max<uint64_t>(n, 0xFFFFFF00ull));
    free_buffers(c);
    for (int i = 0; i < 2; ++i) {
        HIPCHK(hipMalloc(&c->w.key[i], cap * sizeof(key64)));
        HIPCHK(hipMalloc(&c->w.idx[i], cap * sizeof(uint32_t)));
    }
    HIPCHK(hipMalloc(&c->w.bkey, cap * sizeof(key64)));
    HIPCHK(hipMalloc(&c->w.bktrank, cap * sizeof(uint32_t)));
    c->w.cap_large = cap_large_of(cap);
    c->w.cap_block = cap_block_of(cap);
    c->w.cap_tiny = cap_tiny_of(cap);
    c->w.cap_run = cap_run_of(cap);
    for (int i = 0; i < 2; ++i) {
        HIPCHK(hipMalloc(&c->w.l_round[i], (uint64_t)c->w.cap_large * sizeof(Item)));
        HIPCHK(hipMalloc(&c->w.l_lvl[i], (uint64_t)c->w.cap_large * sizeof(Item)));
        HIPCHK(hipMalloc(&c->w.l_tiny[i], (uint64_t)c->w.cap_tiny * sizeof(Item)));
        HIPCHK(hipMalloc(&c->w.l_block[i], (uint64_t)kBlkClasses * c->w.cap_block * sizeof(Item)));
        HIPCHK(hipMalloc(&c->w.l_run[i], (uint64_t)c->w.cap_run * sizeof(Item)));
    }
    // the splitter trees and histograms of a level grow on demand (grow_level)
    c->w.tiles_cap = tiles_cap_of(cap);
    const uint64_t plan_bytes = plan_bytes_of(cap);
    HIPCHK(hipMalloc(&c->w.d_plan, plan_bytes));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->w.h_large), (uint64_t)c->w.cap_large * sizeof(Item)));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->w.h_plan), plan_bytes, hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->w.h_plan_dev), c->w.h_plan, 0));
    c->w.cap = cap;
    c->w_gen++;
    return PS5_OK;
}

// ---- the allocation-free entry point's device scratch -----------------------
//
// ps5_sort_device_noalloc takes every work buffer of a sort of up to
// kNoAllocMaxStrings strings from `scratch`, a static device array: each GPU
// the library runs on holds its own copy from the time the code object is
// loaded until the process ends.  Host code never uses the array's name as a
// device address; the address is resolved once per device with
// hipGetSymbolAddress (scratch_address) and kept in the context.
constexpr uint64_t kScratchBytes = 96ull << 20;
constexpr uint64_t kNoAllocMaxStrings = 1ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Byte offsets of the work buffers of a sort of up to `cap` strings inside
// the scratch.  Histogram bound per large segment of m strings: at most
// min(2047, m/16) buckets (2^treebits < 2m/leaf_target, leaf_target >= 64)
// times at most m/8192 + 1 tiles of kWaves chunks, so cap + cap/4 entries
// cover a level; a level needs at most cap_large trees.
struct ScratchLayout {
    uint64_t key[2], idx[2], bkey, bktrank, l_round[2], l_lvl[2], l_tiny[2], l_block[2], l_run[2];
    uint64_t desc, trees, hist, part, total;
    uint64_t trees_cap, hist_cap, part_cap;
};
constexpr uint64_t align256(uint64_t x) { return (x + 255) / 256 * 256; }
constexpr ScratchLayout scratch_layout(uint64_t cap)
{
    ScratchLayout L{};
    uint64_t o = 0;
    auto take = [&o](uint64_t bytes) { const uint64_t at = o; o = align256(o + bytes); return at; };
    const uint64_t cl = cap_large_of(cap), cb = cap_block_of(cap), ct = cap_tiny_of(cap), cr = cap_run_of(cap);
    for (int i = 0; i < 2; ++i) {
        L.key[i] = take(cap * sizeof(key64));
        L.idx[i] = take(cap * sizeof(uint32_t));
    }
    L.bkey = take(cap * sizeof(key64));
    L.bktrank = take(cap * sizeof(uint32_t));
    for (int i = 0; i < 2; ++i) {
        L.l_round[i] = take(cl * sizeof(Item));
        L.l_lvl[i] = take(cl * sizeof(Item));
        L.l_tiny[i] = take(ct * sizeof(Item));
        L.l_block[i] = take(kBlkClasses * cb * sizeof(Item));
        L.l_run[i] = take(cr * sizeof(Item));
    }
    L.desc = take(plan_bytes_of(cap));
    L.trees_cap = cl;
    L.trees = take(L.trees_cap * kTreeSlots * sizeof(key64));
    L.hist_cap = cap + cap / 4 + 64;
    L.hist = take(L.hist_cap * sizeof(uint32_t));
    L.part_cap = L.hist_cap / kScanBlock + 2;
    L.part = take(L.part_cap * sizeof(uint32_t));
    L.total = o;
    return L;
}
static_assert(scratch_layout(kNoAllocMaxStrings).total <= kScratchBytes,
              "the scratch holds the work buffers of its largest sort");

// Device address of `scratch` on `device` (the caller has made it current):
// resolved on the first call per device and remembered.
int scratch_address(int device, unsigned char** out)
{
    static std::mutex mu;
    static std::map<int, unsigned char*> addr;
    std::lock_guard<std::mutex> g(mu);
    auto it = addr.find(device);
    if (it == addr.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(PS5_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                           hipGetErrorString(e));
        it = addr.emplace(device, static_cast<unsigned char*>(p)).first;
    }
    *out = it->second;
    return PS5_OK;
}

// Calls of the allocation-free entry point on one device share its scratch:
// they are serialised per device inside this process.
std::mutex& scratch_mutex(int device)
{
    static std::mutex mus[64];
    return mus[(unsigned)device % 64u];
}

// Work buffers of a sort of up to kNoAllocMaxStrings strings carved from the
// scratch at `base` (its device address), with the context's pinned plan
// buffers.  Nothing is allocated.
WorkBuffers scratch_buffers(unsigned char* base, Item* h_large, unsigned char* h_plan, unsigned char* h_plan_dev)
{
    constexpr ScratchLayout L = scratch_layout(kNoAllocMaxStrings);
    const uint64_t cap = kNoAllocMaxStrings;
    WorkBuffers w;
    w.fixed = true;
    w.cap = cap;
    for (int i = 0; i < 2; ++i) {
        w.key[i] = reinterpret_cast<key64*>(base + L.key[i]);
        w.idx[i] = reinterpret_cast<uint32_t*>(base + L.idx[i]);
        w.l_round[i] = reinterpret_cast<Item*>(base + L.l_round[i]);
        w.l_lvl[i] = reinterpret_cast<Item*>(base + L.l_lvl[i]);
        w.l_tiny[i] = reinterpret_cast<Item*>(base + L.l_tiny[i]);
        w.l_block[i] = reinterpret_cast<Item*>(base + L.l_block[i]);
        w.l_run[i] = reinterpret_cast<Item*>(base + L.l_run[i]);
    }
    w.bkey = reinterpret_cast<key64*>(base + L.bkey);
    w.bktrank = reinterpret_cast<uint32_t*>(base + L.bktrank);
    w.cap_large = cap_large_of(cap);
    w.cap_block = cap_block_of(cap);
    w.cap_tiny = cap_tiny_of(cap);
    w.cap_run = cap_run_of(cap);
    w.h_large = h_large;
    w.d_plan = base + L.desc;
    w.h_plan = h_plan;
    w.h_plan_dev = h_plan_dev;
    w.tiles_cap = tiles_cap_of(cap);
    w.trees = reinterpret_cast<key64*>(base + L.trees);
    w.trees_cap = L.trees_cap;
    w.hist = reinterpret_cast<uint32_t*>(base + L.hist);
    w.hist_cap = L.hist_cap;
    w.part = reinterpret_cast<uint32_t*>(base + L.part);
    w.part_cap = L.part_cap;
    return w;
}

struct Launch {   // counts a launch; brackets it with events when profiling
    ps5_ctx* c;
    int id;
    hipEvent_t a = nullptr, b = nullptr;
    Launch(ps5_ctx* c_, int id_, uint64_t bytes) : c(c_), id(id_)
    {
        c->k_launch[id]++;
        c->k_bytes[id] += bytes;
        if (!c->profile) return;
        a = take();
        b = take();
        (void)hipEventRecord(a, c->stream);
    }
    ~Launch()
    {
        if (!c->profile) return;
        (void)hipEventRecord(b, c->stream);
        c->ev_used.push_back({ id, { a, b } });
    }
    hipEvent_t take()
    {
        if (c->ev_pool.empty()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            return e;
        }
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
};

void collect_events(ps5_ctx* c)
{
    for (auto& u : c->ev_used) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, u.second.first, u.second.second) == hipSuccess) c->k_ms[u.first] += ms;
        c->ev_pool.push_back(u.second.first);
        c->ev_pool.push_back(u.second.second);
    }
    c->ev_used.clear();
}

// sink: large -> `large` list, leaves -> lists of round parity `rp`
Sink make_sink(ps5_ctx* c, Item* large, int c_large, int rp)
{
    Sink s;
    s.large = large;
    s.block = c->w.l_block[rp];
    s.tiny = c->w.l_tiny[rp];
    s.run = c->w.l_run[rp];
    s.cnt = c->d_cnt;
    s.c_large = c_large;
    s.c_block = C_BLOCK + rp;
    s.c_tiny = C_TINY + rp;
    s.c_run = C_RUN + rp;
    s.c_runcont = C_RUNCONT + rp;
    s.cap_large = c->w.cap_large;
    s.cap_block = c->w.cap_block;
    s.cap_tiny = c->w.cap_tiny;
    s.cap_run = c->w.cap_run;
    return s;
}

// copy the counters (and the large list `list` with counter `ci`) to the host
// Wait until the staged copy's sequence number reads `seq`, rather than for
// the stream to go idle: the host sees the number as soon as the GPU writes
// it.  The stream is queried every 1024 spins, so that a failed stream ends
// the wait with its error.
constexpr uint32_t kGraphSeqPending = 0xFFFFFFFFu;     // set by the host before a captured graph's replay
int wait_stage(ps5_ctx* c, uint32_t seq)
{
    const uint32_t* f = reinterpret_cast<const uint32_t*>(c->h_stage + kStageSeq);
    for (uint32_t spin = 1;; ++spin) {
        if (__atomic_load_n(f, __ATOMIC_ACQUIRE) == seq) return PS5_OK;
        if ((spin & 1023) == 0) {
            const hipError_t e = hipStreamQuery(c->stream);
            if (e == hipSuccess) {
                if (__atomic_load_n(f, __ATOMIC_ACQUIRE) == seq) return PS5_OK;
                return set_err(PS5_ERR_INTERNAL, "the staged copy's sequence number never arrived");
            }
            if (e != hipErrorNotReady) return set_err(PS5_ERR_HIP, "waiting for a staged copy: %s", hipGetErrorString(e));
        }
        __builtin_ia32_pause();
    }
}

int readback(ps5_ctx* c, const Item* list, int ci)
{
    // one staged copy holds the counters and the head of the list
    uint32_t seq = ++c->stage_seq;
    if (seq == 0 || seq == kGraphSeqPending) seq = c->stage_seq = 1;     // (values the graph path uses)
    hipLaunchKernelGGL(k_stage, dim3(1), dim3(kWG), 0, c->stream, c->d_cnt, list, ci, c->h_stage_dev, seq);
    HIPCHK(hipGetLastError());
    int rc = wait_stage(c, seq);
    if (rc) return rc;
    c->st_syncs++;
    memcpy(c->h_cnt, c->h_stage, C_NCNT * sizeof(uint32_t));
    if (c->h_cnt[C_ERR] & ERR_CAP) return set_err(PS5_ERR_CAPACITY, "work list capacity exceeded");
    if (c->h_cnt[C_ERR] & ERR_LCP) return set_err(PS5_ERR_INTERNAL, "equal keys across a leaf boundary");
    if (c->h_cnt[C_ERR] & ERR_BOUNDS) return set_err(PS5_ERR_INTERNAL, "device bounds check failed");
    if (c->h_cnt[C_ERR] & ERR_LEVELS) return set_err(PS5_ERR_INTERNAL, "a captured round needed more levels");
    if (list) {
        const uint32_t nl = c->h_cnt[ci];
        if (nl > c->w.cap_large) return set_err(PS5_ERR_CAPACITY, "large list overflow");
        memcpy(c->w.h_large, c->h_stage + kStageHead, std::min(nl, kPeek) * sizeof(Item));
        if (nl > kPeek) {
            HIPCHK(hipMemcpyAsync(c->w.h_large + kPeek, list + kPeek, (nl - kPeek) * sizeof(Item),
                                  hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            c->st_syncs++;
        }
    }
    return PS5_OK;
}

uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

void sum_sizes(ps5_ctx* c, const Item* list, uint32_t first, const uint32_t* cnt, uint32_t n, int slot)
{
    if (!c->profile || n == 0) return;
    hipLaunchKernelGGL(k_sum_sizes, dim3(std::min<uint32_t>(div_up(n, kWG), 1024)), dim3(kWG), 0, c->stream,
                       list, first, cnt, c->d_sums + slot);
}

// Treebits of a multi-workgroup step on m strings: the fewest levels of at
// most 1023 splitters that bring buckets down to about leaf_target strings,
// with the fan-out spread evenly over those levels.
int large_treebits(uint32_t m, uint32_t target, int maxbits)
{
    const double need = std::log2(std::max(2.0, (double)m / target));
    const int levels = std::max(1, (int)std::ceil(need / maxbits));
    const int k = (int)std::ceil(need / levels);
    return std::max(1, std::min(maxbits, k));
}

// k_count instance for bucket ids of NB = kmax + 1 bits (at least 8)
void launch_count(ps5_ctx* c, const Bufs& B, Plan P, dim3 g, int kmax, bool from_arena, bool loop)
{
    hipStream_t st = c->stream;
    const dim3 b(kWG);
    const uint32_t lds = count_lds_bytes(kmax);
#define PS5_COUNT_LAUNCH(NB)                                                                                          \
    do {                                                                                                              \
        if (loop)                                                                                                     \
            hipLaunchKernelGGL((k_count<NB, true>), g, b, lds, st, P, B, c->w.trees, c->w.hist, (int)from_arena,      \
                               kmax);                                                                                 \
        else                                                                                                          \
            hipLaunchKernelGGL((k_count<NB, false>), g, b, lds, st, P, B, c->w.trees, c->w.hist, (int)from_arena,     \
                               kmax);                                                                                 \
    } while (0)
    if (kmax <= 7) PS5_COUNT_LAUNCH(8);
    else if (kmax == 8) PS5_COUNT_LAUNCH(9);
    else if (kmax == 9) PS5_COUNT_LAUNCH(10);
    else PS5_COUNT_LAUNCH(11);
#undef PS5_COUNT_LAUNCH
}

// One multi-workgroup sample-sort step over all large segments of a level,
// planned on the host from the level's large list (read back into h_large).
int large_step(ps5_ctx* c, const Bufs& B, uint32_t nlarge, const Sink& sink, uint64_t seed, bool from_arena)
{
    hipStream_t st = c->stream;
    uint64_t ntiles = 0, hl = 0, elems = 0, elems_ident = 0;
    int kmax = 1;
    if (nlarge > c->w.cap_large) return set_err(PS5_ERR_CAPACITY, "large list overflow");
    LevelHdr* h_hdr = reinterpret_cast<LevelHdr*>(c->w.h_plan);
    LargeDesc* h_descs = reinterpret_cast<LargeDesc*>(c->w.h_plan + sizeof(LevelHdr));
    TileRef* h_tiles = reinterpret_cast<TileRef*>(c->w.h_plan + plan_tiles_off(c->w.cap));
    uint64_t total = 0;
    bool all_ident = true;     // a level of roots: the scatter stages positions, chunks may be longer
    for (uint32_t i = 0; i < nlarge; ++i) {
        total += c->w.h_large[i].size;
        all_ident = all_ident && (c->w.h_large[i].flags & F_IDENT);
    }
    const uint64_t cmax = all_ident ? c->chunk_max_root : c->chunk_max;
    const uint64_t climit = std::min<uint64_t>(cmax, std::max<uint64_t>(kChunkMin, div_up(total, kWaves * c->min_tiles)));
    for (uint32_t i = 0; i < nlarge; ++i) {
        const Item seg = c->w.h_large[i];
        const uint32_t T = div_up(seg.size, climit * kWaves);
        LargeDesc& D = h_descs[i];
        D.seg = seg;
        D.T = T;
        D.chunk = div_up(seg.size, (uint64_t)T * kWaves);   // <= climit: equal chunks, the last one short
        D.pad = 0;
        D.hist_base = (uint32_t)hl;
        D.k = (uint32_t)large_treebits(seg.size, c->leaf_target, (int)c->max_treebits);
        const uint32_t nspl = (1u << D.k) - 1;
        kmax = std::max(kmax, (int)D.k);
        D.oversample = std::max<uint32_t>(2, std::min<uint32_t>(c->max_oversample, (kMaxSamples / nspl) & ~1u));
        if (ntiles + T > c->w.tiles_cap) return set_err(PS5_ERR_CAPACITY, "tile table capacity");
        for (uint32_t t = 0; t < T; ++t) h_tiles[ntiles + t] = TileRef{ i, t };
        ntiles += T;
        hl += (uint64_t)((2u << D.k) - 1) * T * kWaves;
        elems += seg.size;
        if (seg.flags & F_IDENT) elems_ident += seg.size;
        if (hl > 0xFFFFFFFFull) return set_err(PS5_ERR_CAPACITY, "histogram capacity");
    }
    *h_hdr = LevelHdr{ nlarge, (uint32_t)ntiles, (uint32_t)hl, (uint32_t)kmax, { 0u, 0u, 0u, 0u } };
    if (int rc = grow_level(c, nlarge, hl)) return rc;
    {   // the plan: pulled from mapped host memory by a kernel (no copy engine)
        const uint64_t words = (sizeof(LevelHdr) + nlarge * sizeof(LargeDesc)) / 8, words2 = ntiles * sizeof(TileRef) / 8;
        const uint64_t toff = plan_tiles_off(c->w.cap);
        const dim3 g((uint32_t)std::min<uint64_t>(div_up(std::max(words, words2), kWG), 1024));
        const uint64_t* s1 = reinterpret_cast<const uint64_t*>(c->w.h_plan_dev);
        const uint64_t* s2 = reinterpret_cast<const uint64_t*>(c->w.h_plan_dev + toff);
        uint64_t* d1 = reinterpret_cast<uint64_t*>(c->w.d_plan);
        uint64_t* d2 = reinterpret_cast<uint64_t*>(c->w.d_plan + toff);
        if (c->pending_root)
            hipLaunchKernelGGL(k_root_fetch, g, dim3(kWG), 0, st, s1, d1, words, s2, d2, words2, c->d_cnt,
                               c->w.l_round[0], c->pending_root);
        else
            hipLaunchKernelGGL(k_fetch, g, dim3(kWG), 0, st, s1, d1, words, s2, d2, words2);
        c->pending_root = 0;
    }
    const Plan P = plan_view(c->w.d_plan, c->w.cap);
    const uint32_t hl32 = (uint32_t)hl;
    {
        Launch L(c, K_TREE, (uint64_t)nlarge * kMaxSamples * 8);
        hipLaunchKernelGGL(k_tree, dim3(nlarge), dim3(kBigWG), 0, st, P, B, seed, c->w.trees, (int)from_arena);
    }
    {
        // key 8 R + bktrank 4 W; extracting at the root: offs 4 R + 8 characters R + key 8 W + bktrank 4 W
        Launch L(c, K_COUNT, (from_arena ? elems * 24 : elems * 12) + hl * 4);
        launch_count(c, B, P, dim3(div_up(ntiles, kXcds) * kXcds), kmax, from_arena, false);
    }
    {
        const uint32_t nb = div_up(hl32, kScanBlock);
        Launch L(c, K_SCAN, hl * 12);
        hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kWG), 0, st, c->w.hist, P.hdr, c->w.part);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kWG), 0, st, c->w.part, P.hdr);
        hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kWG), 0, st, c->w.hist, P.hdr, c->w.part);
    }
    {
        // bktrank 4 + key 8 + idx 4 R (no idx at a root), key 8 + idx 4 W
        Launch L(c, K_SCATTER, elems * 28 - elems_ident * 4 + hl * 8);
        const uint32_t nchunks = (uint32_t)ntiles * kWaves;
        const uint32_t wgs = std::min(c->scatter_wgs, div_up(nchunks, kXcds) * kXcds);
        if (all_ident)
            hipLaunchKernelGGL(k_scatter<true>, dim3(wgs), dim3(kBigWG), ScatLds<true>::kBytes, st, P, B, c->w.hist);
        else
            hipLaunchKernelGGL(k_scatter<false>, dim3(wgs), dim3(kBigWG), ScatLds<false>::kBytes, st, P, B, c->w.hist);
    }
    {
        Launch L(c, K_CHILDREN, hl * 4 / kWaves);
        hipLaunchKernelGGL(k_children, dim3(nlarge), dim3(kWG), 0, st, P, c->w.hist, c->w.trees, sink);
    }
    c->st_segments += nlarge;
    return PS5_OK;
}

// Roots of a sort: one segment [0, n), or one per set of a batch (sets never
// mix: every set is its own round segment, so no LCP crosses a set border).
int setup_roots(ps5_ctx* c, uint32_t n, const uint64_t* set_first, uint32_t nsets)
{
    hipStream_t st = c->stream;
    if (!set_first) {
        Item* list; int ci;
        if (n <= kTinyT) { list = c->w.l_tiny[0]; ci = C_TINY; }
        else if (n <= kBlkT) {
            const int cls = block_class(n);
            list = c->w.l_block[0] + (uint64_t)cls * c->w.cap_block;
            ci = C_BLOCK + 2 * cls;
        }
        else { list = c->w.l_round[0]; ci = C_LROUND; }
        Launch L(c, K_MISC, 0);
        hipLaunchKernelGGL(k_setup, dim3(1), dim3(64), 0, st, c->d_cnt, list, ci, n);
        return PS5_OK;
    }
    // root items staged in pinned memory and copied on the context's stream
    // (a blocking copy on the null stream would wait for every other
    // context's work on the device)
    if ((uint64_t)nsets > c->roots_cap) {
        const uint64_t want = std::max<uint64_t>(nsets, 2 * c->roots_cap);
        if (c->h_roots) (void)hipHostFree(c->h_roots);
        c->h_roots = nullptr;
        c->roots_cap = 0;
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->h_roots), want * sizeof(Item)));
        c->roots_cap = want;
    }
    // kinds: 0 tiny, 1..kBlkClasses blocks by size class, last large
    constexpr int NK = kBlkClasses + 2;
    uint32_t cnt[NK] = {};
    uint32_t tot = 0;
    for (int kind = 0; kind < NK; ++kind) {
        for (uint32_t s = 0; s < nsets; ++s) {
            const uint64_t b = set_first[s], e = set_first[s + 1];
            if (e < b || e > n) return set_err(PS5_ERR_ARG, "set boundaries must be non-decreasing and <= n");
            if (e == b) continue;
            const Item it = { (uint32_t)b, (uint32_t)(e - b), (uint32_t)b, F_IDENT };
            const int k = it.size <= kTinyT ? 0 : it.size <= kBlkT ? 1 + block_class(it.size) : NK - 1;
            if (k != kind) continue;
            c->h_roots[tot++] = it;
            cnt[kind]++;
        }
    }
    if (cnt[NK - 1] > c->w.cap_large || cnt[0] > c->w.cap_tiny) return set_err(PS5_ERR_CAPACITY, "too many sets");
    for (int k = 1; k <= kBlkClasses; ++k)
        if (cnt[k] > c->w.cap_block) return set_err(PS5_ERR_CAPACITY, "too many sets");
    HIPCHK(hipStreamSynchronize(st));     // h_upcnt may still feed an earlier copy
    for (int i = 0; i < C_NCNT; ++i) c->h_upcnt[i] = 0;
    c->h_upcnt[C_TINY] = cnt[0];
    for (int k = 1; k <= kBlkClasses; ++k) c->h_upcnt[C_BLOCK + 2 * (k - 1)] = cnt[k];
    c->h_upcnt[C_LROUND] = cnt[NK - 1];
    const Item* h = c->h_roots;
    for (int kind = 0; kind < NK; ++kind) {
        Item* dst = kind == 0 ? c->w.l_tiny[0]
                  : kind == NK - 1 ? c->w.l_round[0]
                  : c->w.l_block[0] + (uint64_t)(kind - 1) * c->w.cap_block;
        if (cnt[kind]) HIPCHK(hipMemcpyAsync(dst, h, cnt[kind] * sizeof(Item), hipMemcpyHostToDevice, st));
        h += cnt[kind];
    }
    HIPCHK(hipMemcpyAsync(c->d_cnt, c->h_upcnt, C_NCNT * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    return PS5_OK;
}

constexpr uint32_t kMaxGraphRounds = 64;      // rounds one captured graph may hold
constexpr uint32_t kForkLeafGroups = 8192;    // leaf workgroups worth a side stream
constexpr uint64_t kChainMin = 1u << 22;       // strings of a level whose successor is chained on the device
// Automatic planning plans a single sort of fewer strings on the device.
// From about 4 M strings on the device plan's bounded grids cost more than
// the host's synchronisations (0.489 vs 0.50 ms on 4 M strings); below, the
// two are within a few percent either way, depending on the input (device
// planning 6-9% faster on 0.13 M digit strings and 0.64 M ASCII strings, 4-7%
// slower on 0.32 M ASCII strings and on high-LCP windows).
constexpr uint64_t kAutoDevMax = 1u << 22;

// Largest grids of the device-planned levels: their kernels loop over the
// work the device lists, so these bound the workgroups of a launch.
constexpr uint32_t kDevTreeWgs = 512, kDevCountWgs = 2048, kDevScanWgs = 1024, kDevChildWgs = 1024,
                   kDevTinyWgs = 8192, kDevRunWgs = 2048, kDevBoundWgs = 1024;
static_assert(kBlkE0 == kBlkE1 && kBlkE1 == kBlkE2 && kBlkE2 == kBlkE3, "one thread count per 256 keys");

// Leaf grids for nt tiny leaves and nr equal-key chunks: the leaf kernels
// loop over their lists, and a bounded grid of them is far faster than a
// workgroup per leaf when leaves number in the millions.
uint32_t leaf_tiny_grid(uint64_t nt) { return (uint32_t)std::min<uint64_t>(kDevTinyWgs, div_up(nt, kWaves)); }
uint32_t leaf_run_grid(uint64_t nr) { return (uint32_t)std::min<uint64_t>(kDevRunWgs, nr); }

// Looping block kernels of size class cls: four workgroups per wave slot
// of the GPU (32 per CU)
uint32_t block_grid_full(const ps5_ctx* c, int cls)
{
    const uint32_t waves = (256u << cls) / (uint32_t)kBlkE0 / 64;
    return (uint32_t)((uint64_t)c->scatter_wgs * 128 / waves);
}

// Blocks of size class cls and round parity rp: the list entries [first,
// first + count) (or, with dev_first, from the device counter C_BDONE + cls
// to the list's count), `grid` workgroups.
void launch_block_class(ps5_ctx* c, const Bufs& B, int rp, int cls, uint32_t first, uint32_t count, bool dev_first,
                        uint32_t grid, const Sink& nsink, uint32_t depth, bool tail = false, int flags = 0)
{
    const uint32_t last = first + count;
    hipStream_t st = c->stream;
    const Item* list = c->w.l_block[rp] + (uint64_t)cls * c->w.cap_block;
    const uint32_t* cnt = c->d_cnt + C_BLOCK + 2 * cls + rp;
    const uint32_t* pf = dev_first ? c->d_cnt + C_BDONE + cls : nullptr;
    const uint32_t cap = c->w.cap_block;
    Launch L(c, K_BLOCK, 0);
    // the default key width (8 keys) has its own instances; others share one
#define PS5_BLOCK_LAUNCH(M, E, LOOP, TAIL)                                                                         \
    do {                                                                                                           \
        if (c->leaf_keys == 8)                                                                                     \
            hipExtLaunchKernelGGL((k_block<M, E, LOOP, 8, TAIL>), dim3(grid), dim3(M / E), 0, st, nullptr, nullptr, \
                                  flags, list, first, last, pf, cnt, cap, B, nsink, depth);                          \
        else                                                                                                       \
            hipExtLaunchKernelGGL((k_block<M, E, LOOP, 0, TAIL>), dim3(grid), dim3(M / E), 0, st, nullptr, nullptr, \
                                  flags, list, first, last, pf, cnt, cap, B, nsink, depth);                          \
    } while (0)
    if (dev_first) {
        if (cls == 0) PS5_BLOCK_LAUNCH(256, kBlkE0, true, false);
        else if (cls == 1) PS5_BLOCK_LAUNCH(512, kBlkE1, true, false);
        else if (cls == 2) PS5_BLOCK_LAUNCH(1024, kBlkE2, true, false);
        else PS5_BLOCK_LAUNCH(kBlkT, kBlkE3, true, false);
    } else if (tail) {
        if (cls == 0) PS5_BLOCK_LAUNCH(256, kBlkE0, false, true);
        else if (cls == 1) PS5_BLOCK_LAUNCH(512, kBlkE1, false, true);
        else if (cls == 2) PS5_BLOCK_LAUNCH(1024, kBlkE2, false, true);
        else PS5_BLOCK_LAUNCH(kBlkT, kBlkE3, false, true);
    } else {
        if (cls == 0) PS5_BLOCK_LAUNCH(256, kBlkE0, false, false);
        else if (cls == 1) PS5_BLOCK_LAUNCH(512, kBlkE1, false, false);
        else if (cls == 2) PS5_BLOCK_LAUNCH(1024, kBlkE2, false, false);
        else PS5_BLOCK_LAUNCH(kBlkT, kBlkE3, false, false);
    }
#undef PS5_BLOCK_LAUNCH
}

// Tiny leaves and equal-key chunks of round parity rp on stream s (grids of
// g_tiny and g_run workgroups; 0 skips a list).
void launch_leaves(ps5_ctx* c, const Bufs& B, int rp, const Sink& nsink, uint32_t depth, uint32_t g_tiny,
                   uint32_t g_run, uint64_t n_tiny, uint64_t n_run, hipStream_t s, int flags = 0)
{
    if (g_tiny) {
        {
            Launch L(c, K_LEAF_TINY, 0);
            if (c->leaf_keys == 8)
                hipExtLaunchKernelGGL(k_leaf_tiny<8>, dim3(g_tiny), dim3(kWG), 0, s, nullptr, nullptr, flags,
                                      c->w.l_tiny[rp], c->d_cnt + C_TINY + rp, c->w.cap_tiny, B, depth);
            else
                hipExtLaunchKernelGGL(k_leaf_tiny<0>, dim3(g_tiny), dim3(kWG), 0, s, nullptr, nullptr, flags,
                                      c->w.l_tiny[rp], c->d_cnt + C_TINY + rp, c->w.cap_tiny, B, depth);
        }
        sum_sizes(c, c->w.l_tiny[rp], 0, c->d_cnt + C_TINY + rp, (uint32_t)n_tiny, S_TINY);
    }
    if (g_run) {
        {
            Launch L(c, K_LEAF_RUN, 0);
            hipExtLaunchKernelGGL(k_leaf_run, dim3(g_run), dim3(kWG), 0, s, nullptr, nullptr, flags, c->w.l_run[rp],
                                  c->d_cnt + C_RUN + rp, c->w.cap_run, B, nsink, depth);
        }
        sum_sizes(c, c->w.l_run[rp], 0, c->d_cnt + C_RUN + rp, (uint32_t)n_run, S_RUN);
    }
}

// LCPs between neighbouring leaves of round parity rp (grid workgroups).
void launch_leafbound(ps5_ctx* c, const Bufs& B, int rp, uint32_t depth, uint32_t grid, uint64_t nleaves)
{
    Launch L(c, K_LEAFBOUND, nleaves * 36);
    LeafLists LL;
    LL.list[0] = c->w.l_tiny[rp];
    LL.cnt[0] = C_TINY + rp;
    LL.cap[0] = c->w.cap_tiny;
    for (int cls = 0; cls < kBlkClasses; ++cls) {
        LL.list[1 + cls] = c->w.l_block[rp] + (uint64_t)cls * c->w.cap_block;
        LL.cnt[1 + cls] = C_BLOCK + 2 * cls + rp;
        LL.cap[1 + cls] = c->w.cap_block;
    }
    LL.list[kBlkClasses + 1] = c->w.l_run[rp];
    LL.cnt[kBlkClasses + 1] = C_RUN + rp;
    LL.cap[kBlkClasses + 1] = c->w.cap_run;
    hipLaunchKernelGGL(k_leafbound, dim3(grid), dim3(kWG), 0, c->stream, LL, B, depth);
}

// One level planned on the device: k_plan turns the large list `list`
// (counter ci) into the level's plan, and the step's kernels read their work
// from it, in grids bounded by what the host knows (at most bound_large
// segments holding S strings), so nothing comes back to the host.
void dev_level(ps5_ctx* c, const Bufs& B, const Item* list, int ci, uint32_t bound_large, uint64_t S,
               const Sink& sink, uint64_t seed, bool from_arena, bool ident, int kmax, int done_rp)
{
    hipStream_t st = c->stream;
    const uint64_t n = S;     // strings the level's segments hold at most
    PlanArgs A;
    A.list = list;
    A.ci = ci;
    A.cap_large = c->w.cap_large;
    A.plan = c->w.d_plan;
    A.tiles_off = plan_tiles_off(c->w.cap);
    A.tiles_cap = c->w.tiles_cap;
    A.hist_cap = c->w.hist_cap;
    A.trees_cap = c->w.trees_cap;
    A.leaf_target = c->leaf_target;
    A.max_treebits = c->max_treebits;
    A.max_oversample = c->max_oversample;
    A.chunk_max = c->chunk_max;
    A.chunk_max_root = c->chunk_max_root;
    A.min_tiles = c->min_tiles;
    A.kmax = (uint32_t)kmax;
    A.done_rp = done_rp;
    {
        Launch L(c, K_MISC, 0);
        hipLaunchKernelGGL(k_plan, dim3(1), dim3(kBigWG), 0, st, A, c->d_cnt);
    }
    const Plan P = plan_view(c->w.d_plan, c->w.cap);
    const uint32_t gl = std::max<uint32_t>(1, bound_large);
    {
        Launch L(c, K_TREE, 0);
        hipLaunchKernelGGL(k_tree, dim3(std::min(gl, kDevTreeWgs)), dim3(kBigWG), 0, st, P, B, seed, c->w.trees,
                           (int)from_arena);
    }
    const uint64_t tiles_bound = n / ((uint64_t)kWaves * kChunkMin) + gl;
    {
        Launch L(c, K_COUNT, 0);
        launch_count(c, B, P, dim3((uint32_t)std::min<uint64_t>(kDevCountWgs, div_up(tiles_bound, kXcds) * kXcds)),
                     kmax, from_arena, true);
    }
    {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(kDevScanWgs, div_up(c->w.hist_cap, kScanBlock));
        Launch L(c, K_SCAN, 0);
        hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kWG), 0, st, c->w.hist, P.hdr, c->w.part);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kWG), 0, st, c->w.part, P.hdr);
        hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kWG), 0, st, c->w.hist, P.hdr, c->w.part);
    }
    {
        Launch L(c, K_SCATTER, 0);
        const uint32_t wgs = (uint32_t)std::min<uint64_t>(c->scatter_wgs, div_up(tiles_bound * kWaves, kXcds) * kXcds);
        if (ident)
            hipLaunchKernelGGL(k_scatter<true>, dim3(wgs), dim3(kBigWG), ScatLds<true>::kBytes, st, P, B, c->w.hist);
        else
            hipLaunchKernelGGL(k_scatter<false>, dim3(wgs), dim3(kBigWG), ScatLds<false>::kBytes, st, P, B, c->w.hist);
    }
    {
        Launch L(c, K_CHILDREN, 0);
        hipLaunchKernelGGL(k_children, dim3(std::min(gl, kDevChildWgs)), dim3(kWG), 0, st, P, c->w.hist, c->w.trees,
                           sink);
    }
}

// Bound on the buckets of a level over segments of at most m strings (its
// treebits in *k): 4x their mean below 16 sample keys per splitter (large
// trees; keys repeated between two splitters can add to the spread of the
// sample spacings, 2.9x has been seen at 8), 1 + 8/sqrt(s) from s = 16 on
// (3x at 16, 2.5x at 32: small trees).  Below 4x, the bound shrinks with
// every level of at least 3 splitters.
uint64_t level_bucket_bound(const ps5_ctx* c, uint64_t m, int* k)
{
    *k = large_treebits((uint32_t)std::min<uint64_t>(m, 0xFFFFFFFFull), c->leaf_target, (int)c->max_treebits);
    const uint32_t nspl = (1u << *k) - 1;
    const uint32_t s = std::max<uint32_t>(2, std::min<uint32_t>(c->max_oversample, (kMaxSamples / nspl) & ~1u));
    const uint64_t quarters = s < 16 ? 16 : 4 + (uint64_t)std::ceil(32.0 / std::sqrt((double)s));   // the factor, in quarters
    return (m * quarters + (4ull << *k) - 1) / (4ull << *k);
}

// Levels a round needs when a segment may hold m strings, each level's
// buckets bounded by level_bucket_bound (blocks are sorted by the level that
// lists them, so no level of blocks alone is needed).  A round whose
// segments outgrow the bound leaves them pending, and the host finishes them.
uint32_t dev_levels_for(const ps5_ctx* c, uint64_t m)
{
    uint32_t L = 0;
    int k;
    while (m > kBlkT && L < 8) {
        m = level_bucket_bound(c, m, &k);
        ++L;
    }
    return std::max<uint32_t>(L, 1);
}

// Largest treebits a device-planned level `level` of a round gives its
// segments, when the round's segments hold at most m strings: the bound of
// dev_levels_for, level by level.  A segment above the bound only gets a
// smaller tree than it would choose, and the next level splits it further.
int dev_level_kmax(const ps5_ctx* c, uint64_t m, uint32_t level)
{
    int k;
    for (uint32_t l = 0; l < level; ++l) m = level_bucket_bound(c, m, &k);
    (void)level_bucket_bound(c, m, &k);
    return k;
}

// The host-planned levels of a round from the large list of counter cur_ci
// (in h_cnt / h_large): each level's large list is read back and planned on
// the host, and the blocks listed before a level are sorted in it
// (blk_done: blocks already sorted per size class).
int host_levels(ps5_ctx* c, const Bufs& B, uint32_t round, int cur_ci, uint32_t* blk_done, bool fused_root,
                bool* leaves_forked, uint32_t reset_mask, bool* after_blocks)
{
    *after_blocks = false;     // set when the stream's last packets are the last level's block kernels
    hipStream_t st = c->stream;
    const uint32_t depth = round * 8;
    const int rp = round & 1, np = rp ^ 1;
    const Sink nsink = make_sink(c, c->w.l_round[np], C_LROUND + np, np);
    int rc;
    for (uint32_t level = 0;; ++level) {
        if (level > 4096) return set_err(PS5_ERR_INTERNAL, "sample-sort levels do not terminate");
        const uint32_t nlarge = c->h_cnt[cur_ci];
        uint32_t nblk[kBlkClasses], nblk_all = 0;
        for (int cls = 0; cls < kBlkClasses; ++cls) {
            nblk[cls] = c->h_cnt[C_BLOCK + 2 * cls + rp] - blk_done[cls];
            nblk_all += nblk[cls];
        }
        Item* nx_large = c->w.l_lvl[level & 1];
        int nx_ci = C_LLVL + (level & 1);
        if (nx_ci == cur_ci) {      // (a round continued from device levels may start on either list)
            nx_large = c->w.l_lvl[(level & 1) ^ 1];
            nx_ci = C_LLVL + ((level & 1) ^ 1);
        }
        // counters to clear before anything of this level is enqueued (the
        // forked leaves included): the caller's (the next round's lists, at
        // level 0) and, when a large step runs, the next level's list (every
        // counter starts at zero in round 0)
        if (nlarge && (round > 0 || level > 0 || cur_ci != C_LROUND + rp)) reset_mask |= 1u << nx_ci;
        if (reset_mask) {
            Launch L(c, K_MISC, 0);
            hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, c->d_cnt, reset_mask);
            reset_mask = 0;
        }
        if (nlarge == 0 && nblk_all == 0) break;
        c->st_levels++;
        // the last level (blocks only): the round's other leaves are all
        // listed, and when they are many they run on the side stream
        // beside the blocks (not when profiling, whose events time the
        // main stream; for few leaves the fork costs more than it hides)
        const uint32_t nt = c->h_cnt[C_TINY + rp], nr = c->h_cnt[C_RUN + rp];
        if (nlarge == 0 && !c->profile && div_up(nt, kWaves) + nr >= kForkLeafGroups) {
            HIPCHK(hipEventRecord(c->ev_fork, st));
            HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
            launch_leaves(c, B, rp, nsink, depth, leaf_tiny_grid(nt), leaf_run_grid(nr), nt, nr, c->side);
            HIPCHK(hipEventRecord(c->ev_join, c->side));
            *leaves_forked = true;
        }
        const Sink sink = make_sink(c, nx_large, nx_ci, rp);
        const uint64_t seed = c->seed ^ ((uint64_t)round << 40) ^ ((uint64_t)level << 32);
        if (nlarge && (rc = large_step(c, B, nlarge, sink, seed, fused_root && round == 0 && level == 0)))
            return rc;
        // a level without large segments, in a round whose equal-key
        // leaves all end here (their keys hold a NUL), holds the round's
        // last work: its blocks finish every string themselves (tail mode)
        // and no further round is needed
        const bool tail = c->tail_blocks && nlarge == 0 && c->h_cnt[C_RUNCONT + rp] == 0;
        int bflags = 0;     // the first class waits for the level; the others may run beside it
        for (int cls = 0; cls < kBlkClasses; ++cls) {
            if (!nblk[cls]) continue;
            launch_block_class(c, B, rp, cls, blk_done[cls], nblk[cls], false, nblk[cls], nsink, depth, tail, bflags);
            if (!c->profile) bflags = hipExtAnyOrderLaunch;
            *after_blocks = nlarge == 0;
            sum_sizes(c, c->w.l_block[rp] + (uint64_t)cls * c->w.cap_block, blk_done[cls],
                      c->d_cnt + C_BLOCK + 2 * cls + rp, blk_done[cls] + nblk[cls], S_BLOCK);
            blk_done[cls] += nblk[cls];
        }
        HIPCHK(hipGetLastError());
        // blocks only append to the next round's lists: after a level
        // without large segments this round has no more levels
        if (nlarge == 0) break;
        uint64_t S = 0, mmax = 0;     // this level's strings and largest segment
        for (uint32_t i = 0; c->chain_levels && i < nlarge; ++i) {
            S += c->w.h_large[i].size;
            mmax = std::max<uint64_t>(mmax, c->w.h_large[i].size);
        }
        if (c->chain_levels && !c->profile && S >= kChainMin) {
            // the next level planned on the device at once, without reading
            // this level's large list back: its segments hold at most this
            // level's S strings and are no larger than its largest segment
            // (below kChainMin strings the bounded grids of a device level
            // cost more than the read-back they save; a profiled sort plans
            // every level on the host, which knows each launch's bytes)
            const int nn = (nx_ci - C_LLVL) ^ 1;     // the level list the device level fills
            Item* nn_large = c->w.l_lvl[nn];
            const int nn_ci = C_LLVL + nn;
            {
                Launch L(c, K_MISC, 0);
                hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, c->d_cnt, 1u << nn_ci);
            }
            ++level;
            c->st_levels++;
            const uint64_t seed2 = c->seed ^ ((uint64_t)round << 40) ^ ((uint64_t)level << 32);
            const int kmax = large_treebits((uint32_t)std::min<uint64_t>(mmax, 0xFFFFFFFFull), c->leaf_target,
                                            (int)c->max_treebits);
            dev_level(c, B, nx_large, nx_ci, (uint32_t)std::min<uint64_t>(c->w.cap_large, S / (kBlkT + 1) + 1), S,
                      make_sink(c, nn_large, nn_ci, rp), seed2, false, false, kmax, -1);
            HIPCHK(hipGetLastError());
            if ((rc = readback(c, nn_large, nn_ci))) return rc;
            cur_ci = nn_ci;
            continue;
        }
        if ((rc = readback(c, nx_large, nx_ci))) return rc;
        cur_ci = nx_ci;
    }
    return PS5_OK;
}

// Leaves and leaf boundaries of a round planned on the host (h_cnt holds
// the round's final counts), then the read-back of the next round's lists.
int host_round_end(ps5_ctx* c, const Bufs& B, uint32_t round, bool leaves_forked, bool after_blocks)
{
    hipStream_t st = c->stream;
    const uint32_t depth = round * 8;
    const int rp = round & 1, np = rp ^ 1;
    const Sink nsink = make_sink(c, c->w.l_round[np], C_LROUND + np, np);
    const uint32_t nt = c->h_cnt[C_TINY + rp], nr = c->h_cnt[C_RUN + rp];
    uint32_t nb = 0;
    for (int cls = 0; cls < kBlkClasses; ++cls) nb += c->h_cnt[C_BLOCK + 2 * cls + rp];
    c->st_leaves += nt + nb + nr;
    if (c->trace) {
        fprintf(stderr, "[ps5] round %u depth %u: tiny %u, blocks %u/%u/%u/%u, equal chunks %u, levels %llu\n",
                round, depth, nt, c->h_cnt[C_BLOCK + rp], c->h_cnt[C_BLOCK + 2 + rp], c->h_cnt[C_BLOCK + 4 + rp],
                c->h_cnt[C_BLOCK + 6 + rp], nr, (unsigned long long)c->st_levels);
    }
    if (leaves_forked) HIPCHK(hipStreamWaitEvent(st, c->ev_join, 0));
    // right after the last level's block kernels (the first of which waited
    // for everything before it) the leaves may start beside them; the leaf
    // boundaries below wait for all of them
    else launch_leaves(c, B, rp, nsink, depth, leaf_tiny_grid(nt), leaf_run_grid(nr), nt, nr, st,
                       after_blocks && !c->profile ? hipExtAnyOrderLaunch : 0);
    if (nt + nb + nr && B.lcp) launch_leafbound(c, B, rp, depth, div_up((uint64_t)nt + nb + nr, kWG), nt + nb + nr);
    HIPCHK(hipGetLastError());
    return readback(c, c->w.l_round[np], C_LROUND + np);
}

// One sort call's state, shared by its rounds.
struct SortRun {
    Bufs B;
    uint32_t n;
    uint64_t alen;
    bool fused_root;          // one root whose keys the first count pass extracts
    uint32_t root_kmax;       // largest treebits of a large root
    uint64_t root_max;        // largest root: no segment of any round is larger
    uint64_t root_sum;        // strings in large roots
};

// Round `round` of a sort, planned on the host or (dev) on the device; sets
// *done when no work is left for another round.
int sort_round(ps5_ctx* c, SortRun& R, uint32_t round, bool dev, bool* done)
{
    int rc;
    hipStream_t st = c->stream;
    Bufs& B = R.B;
    const uint32_t n = R.n;
    const uint64_t alen = R.alen;
    const bool fused_root = R.fused_root;
    const uint32_t root_kmax = R.root_kmax;
    const uint64_t root_max = R.root_max, root_sum = R.root_sum;
    uint32_t* lcp = B.lcp;
    *done = false;
    const uint32_t depth = round * 8;
    const int rp = round & 1, np = rp ^ 1;
    c->st_rounds++;
    uint32_t round_mask = 0;     // the lists of round r+1 are empty again (consumed in round r-1; set up empty in round 0)
    if (round > 0) {
        round_mask = (1u << (C_LROUND + np)) | (1u << (C_TINY + np)) | (1u << (C_RUN + np));
        for (int cls = 0; cls < kBlkClasses; ++cls)
            round_mask |= (1u << (C_BLOCK + 2 * cls + np)) | (1u << (C_BDONE + cls));
        round_mask |= 1u << (C_RUNCONT + np);
    }
    if (round_mask && dev) {     // (host-planned levels clear them with their first level's counters)
        Launch L(c, K_MISC, 0);
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, c->d_cnt, round_mask);
    }
    const Sink nsink = make_sink(c, c->w.l_round[np], C_LROUND + np, np);
    uint32_t blk_done[kBlkClasses] = {};
    bool leaves_forked = false, after_blocks = false;
    if (!dev) {
        if ((rc = host_levels(c, B, round, C_LROUND + rp, blk_done, fused_root, &leaves_forked, round_mask,
                              &after_blocks)))
            return rc;
        if ((rc = host_round_end(c, B, round, leaves_forked, after_blocks))) return rc;
    } else {
        // every level of the round without a host synchronisation.  The
        // host knows the round's lists (read back with the previous
        // round, or set up): S strings in nl0 large segments bound the
        // grids; a round without large segments is planned on the host
        const uint32_t nl0 = c->h_cnt[C_LROUND + rp];
        uint64_t S = 0, mmax = 0;
        if (round == 0) {
            S = root_sum;
            mmax = root_max;
        } else if (nl0 <= kPeek) {
            for (uint32_t i = 0; i < nl0; ++i) {
                S += c->w.h_large[i].size;
                mmax = std::max<uint64_t>(mmax, c->w.h_large[i].size);
            }
        } else {
            S = n;
            mmax = root_max;
        }
        const uint32_t L = nl0 ? dev_levels_for(c, mmax) : 0;
        if (L == 0) {
            if ((rc = host_levels(c, B, round, C_LROUND + rp, blk_done, false, &leaves_forked, 0u, &after_blocks)))
                return rc;
            if ((rc = host_round_end(c, B, round, leaves_forked, after_blocks))) return rc;
        } else {
            int in_ci = C_LROUND + rp;
            const Item* in_list = c->w.l_round[rp];
            const uint32_t deeper_large = (uint32_t)std::min<uint64_t>(c->w.cap_large, S / (kBlkT + 1) + 1);
            for (uint32_t level = 0; level < L; ++level) {
                c->st_levels++;
                const int out_ci = C_LLVL + (level & 1);
                if (!(round == 0 && level == 0)) {
                    Launch Lr(c, K_MISC, 0);
                    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, c->d_cnt, 1u << out_ci);
                }
                const Sink sink = make_sink(c, c->w.l_lvl[level & 1], out_ci, rp);
                const uint64_t seed = c->seed ^ ((uint64_t)round << 40) ^ ((uint64_t)level << 32);
                const bool rootlvl = round == 0 && level == 0;
                dev_level(c, B, in_list, in_ci, level == 0 ? nl0 : deeper_large, S, sink, seed,
                          fused_root && rootlvl, rootlvl, rootlvl ? (int)root_kmax : dev_level_kmax(c, mmax, level),
                          level > 0 ? rp : -1);
                int bflags = 0;     // the first class waits for the level; the others run beside it
                for (int cls = 0; cls < kBlkClasses; ++cls) {
                    // four workgroups for every wave slot of the GPU (32
                    // per CU): blocks differ in size up to 2x, and the
                    // dispatcher's later workgroups even out the load
                    const uint32_t minm = cls == 0 ? kTinyT + 1 : (128u << cls) + 1;
                    const uint64_t full = block_grid_full(c, cls);
                    const uint64_t known = level == 0 ? c->h_cnt[C_BLOCK + 2 * cls + rp] : 0;
                    const uint32_t g = (uint32_t)std::min<uint64_t>(full, known + S / minm + 1);
                    launch_block_class(c, B, rp, cls, 0, 0, true, g, nsink, depth, false, bflags);
                    if (!c->profile) bflags = hipExtAnyOrderLaunch;
                }
                HIPCHK(hipGetLastError());
                in_ci = out_ci;
                in_list = c->w.l_lvl[level & 1];
            }
            // leaves with persistent grids: they skip their work when
            // large segments remain (counter in_ci), which the host then
            // finishes
            const uint64_t nt0 = c->h_cnt[C_TINY + rp], nr0 = c->h_cnt[C_RUN + rp];
            const uint32_t gt = (uint32_t)std::min<uint64_t>(kDevTinyWgs, div_up(nt0 + S, kWaves));
            const uint32_t gr = (uint32_t)std::min<uint64_t>(kDevRunWgs, nr0 + S / kTinyT + 1);
            B.pending = c->d_cnt + in_ci;
            launch_leaves(c, B, rp, nsink, depth, gt, gr, nt0 + S, nr0 + S, st);
            if (lcp) {
                uint64_t known = nt0 + nr0;
                for (int cls = 0; cls < kBlkClasses; ++cls) known += c->h_cnt[C_BLOCK + 2 * cls + rp];
                launch_leafbound(c, B, rp, depth,
                                 (uint32_t)std::min<uint64_t>(kDevBoundWgs, div_up(known + S, kWG)), known + S);
            }
            B.pending = nullptr;
            HIPCHK(hipGetLastError());
            if ((rc = readback(c, c->w.l_round[np], C_LROUND + np))) return rc;
            if (c->h_cnt[in_ci] != 0) {
                // more levels than planned: read the remaining large list
                // back and finish the round on the host
                for (int cls = 0; cls < kBlkClasses; ++cls) blk_done[cls] = c->h_cnt[C_BLOCK + 2 * cls + rp];
                if ((rc = readback(c, in_list, in_ci))) return rc;
                if ((rc = host_levels(c, B, round, in_ci, blk_done, false, &leaves_forked, 0u, &after_blocks)))
                    return rc;
                if ((rc = host_round_end(c, B, round, leaves_forked, after_blocks))) return rc;
            } else {
                uint32_t nb = 0;
                for (int cls = 0; cls < kBlkClasses; ++cls) nb += c->h_cnt[C_BLOCK + 2 * cls + rp];
                c->st_leaves += c->h_cnt[C_TINY + rp] + nb + c->h_cnt[C_RUN + rp];
            }
        }
    }
    uint64_t pending = (uint64_t)c->h_cnt[C_LROUND + np] + c->h_cnt[C_TINY + np] + c->h_cnt[C_RUN + np];
    for (int cls = 0; cls < kBlkClasses; ++cls) pending += c->h_cnt[C_BLOCK + 2 * cls + np];
    if (pending == 0) { *done = true; return PS5_OK; }
    if (depth > alen + 8) return set_err(PS5_ERR_INTERNAL, "depth exceeds the arena: no progress");
    return PS5_OK;
}

// Round `round` for a captured graph: planned on the device with grids
// bounded by what the host knows before the sort (n strings, roots of at
// most root_max), no host synchronisation; L device levels, and a round that
// needs more is flagged (ERR_LEVELS) by its leaf kernels.
void graph_round(ps5_ctx* c, SortRun& R, uint32_t round, uint32_t L, uint32_t nl_root)
{
    hipStream_t st = c->stream;
    Bufs B = R.B;
    const uint64_t n = R.n;
    const uint32_t depth = round * 8;
    const int rp = round & 1, np = rp ^ 1;
    if (round > 0) {
        uint32_t mask = (1u << (C_LROUND + np)) | (1u << (C_TINY + np)) | (1u << (C_RUN + np));
        for (int cls = 0; cls < kBlkClasses; ++cls) mask |= (1u << (C_BLOCK + 2 * cls + np)) | (1u << (C_BDONE + cls));
        mask |= 1u << (C_RUNCONT + np);
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, c->d_cnt, mask);
    }
    const Sink nsink = make_sink(c, c->w.l_round[np], C_LROUND + np, np);
    const uint64_t S = round == 0 ? R.root_sum : n;
    const uint32_t any_large = (uint32_t)std::min<uint64_t>(c->w.cap_large, n / (kBlkT + 1) + 1);
    int in_ci = C_LROUND + rp;
    const Item* in_list = c->w.l_round[rp];
    for (uint32_t level = 0; level < L; ++level) {
        const int out_ci = C_LLVL + (level & 1);
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, c->d_cnt, 1u << out_ci);
        const Sink sink = make_sink(c, c->w.l_lvl[level & 1], out_ci, rp);
        const uint64_t seed = c->seed ^ ((uint64_t)round << 40) ^ ((uint64_t)level << 32);
        const bool rootlvl = round == 0 && level == 0;
        dev_level(c, B, in_list, in_ci, rootlvl ? nl_root : any_large, S, sink, seed, R.fused_root && rootlvl,
                  rootlvl, rootlvl ? (int)R.root_kmax : dev_level_kmax(c, R.root_max, level), level > 0 ? rp : -1);
        for (int cls = 0; cls < kBlkClasses; ++cls) {
            launch_block_class(c, B, rp, cls, 0, 0, true, block_grid_full(c, cls), nsink, depth);
        }
        in_ci = out_ci;
        in_list = c->w.l_lvl[level & 1];
    }
    B.pending = c->d_cnt + in_ci;
    B.pending_err = 1;
    launch_leaves(c, B, rp, nsink, depth, (uint32_t)std::min<uint64_t>(kDevTinyWgs, div_up(n, kWaves)),
                  (uint32_t)std::min<uint64_t>(kDevRunWgs, c->w.cap_run), n, n, st);
    if (B.lcp) launch_leafbound(c, B, rp, depth, (uint32_t)std::min<uint64_t>(kDevBoundWgs, div_up(n, kWG)), n);
}

// Round 0 keys the first classification pass does not extract: every
// position's (no fused root), or those of a batch's sets that start as tiny
// leaves or blocks (their lists and counts as setup_roots left them).
int init_root_keys(ps5_ctx* c, const Bufs& B, uint32_t n, const uint64_t* set_first, bool fused_root)
{
    hipStream_t st = c->stream;
    if (!fused_root) {
        Launch L(c, K_INIT, (uint64_t)n * 20);   // offs 4 R + 8 characters R + key 8 W
        hipLaunchKernelGGL(k_init, dim3(div_up(n, kWG)), dim3(kWG), 0, st, B, n);
    } else if (set_first) {
        Launch L(c, K_INIT, 0);
        const uint32_t nt = c->h_upcnt[C_TINY];
        if (nt) hipLaunchKernelGGL(k_init_items, dim3(std::min<uint32_t>(nt, 4096)), dim3(kWG), 0, st,
                                   c->w.l_tiny[0], nt, B);
        for (int cls = 0; cls < kBlkClasses; ++cls) {
            const uint32_t nb = c->h_upcnt[C_BLOCK + 2 * cls];
            if (nb) hipLaunchKernelGGL(k_init_items, dim3(std::min<uint32_t>(nb, 4096)), dim3(kWG), 0, st,
                                       c->w.l_block[0] + (uint64_t)cls * c->w.cap_block, nb, B);
        }
    }
    HIPCHK(hipGetLastError());
    return PS5_OK;
}

int run_sort(ps5_ctx* c, const uint8_t* arena, uint64_t alen, const uint32_t* offs, const uint64_t* offs64,
             uint32_t n, uint32_t* perm, uint32_t* lcp, const uint64_t* set_first, uint32_t nsets)
{
    int rc = ensure_capacity(c, n);
    if (rc) return rc;
    for (int i = 0; i < K_NUM; ++i) { c->k_ms[i] = 0; c->k_launch[i] = 0; c->k_bytes[i] = 0; }
    c->st_rounds = c->st_levels = c->st_segments = c->st_leaves = c->st_syncs = 0;
    hipStream_t st = c->stream;
    if (c->profile) HIPCHK(hipMemsetAsync(c->d_sums, 0, S_NUM * sizeof(unsigned long long), st));

    Bufs B;
    B.n = n;
    B.cnt = c->d_cnt;
    B.arena = arena;
    B.alen = alen;
    B.offs = offs;
    B.offs64 = offs64;
    B.key0 = c->w.key[0]; B.key1 = c->w.key[1];
    B.idx0 = c->w.idx[0]; B.idx1 = c->w.idx[1];
    B.perm = perm;
    B.lcp = lcp;
    B.bkey = c->w.bkey;
    B.bktrank = c->w.bktrank;
    B.pending = nullptr;
    B.pending_err = 0;
    B.leaf_keys = c->leaf_keys;

    // the roots (one, or a batch's sets) and their sizes
    uint32_t root_kmax = 1;                     // largest treebits of a large root
    uint64_t root_max = n;                      // largest root (no segment of any round is larger)
    uint64_t root_sum = n;                      // strings in large roots
    if (!set_first) {
        root_kmax = (uint32_t)large_treebits(n, c->leaf_target, (int)c->max_treebits);
    } else {
        root_max = 0;
        root_sum = 0;
        for (uint32_t s = 0; s < nsets; ++s) {
            const uint64_t m = set_first[s + 1] - set_first[s];
            root_max = std::max(root_max, m);
            if (m <= kBlkT) continue;
            root_sum += m;
            root_kmax = std::max<uint32_t>(root_kmax, (uint32_t)large_treebits((uint32_t)m, c->leaf_target,
                                                                               (int)c->max_treebits));
        }
    }
    // Roots larger than a block have their keys extracted by the first
    // classification pass itself (its level holds only roots, whose indices
    // are their positions); k_init covers the rest.  One such root alone is
    // known to the host without asking.
    const bool single_root = !set_first && n > kBlkT;
    const bool fused_root = single_root || (set_first && root_sum > 0);
    bool dev = c->planning == 1 ||
               (c->planning == 2 && !c->profile && !set_first && n < kAutoDevMax);   // (profiled sorts: host, whose
                                                                                     // launches know their bytes)
    // a single root planned on the host: the root level's plan fetch, the
    // sort's first kernel, also sets the counters and the root item up
    c->pending_root = single_root && !dev && !(c->graph_rounds && !c->profile) ? n : 0u;
    if (!c->pending_root && (rc = setup_roots(c, n, set_first, nsets))) return rc;
    if ((rc = init_root_keys(c, B, n, set_first, fused_root))) return rc;
    // device-planned levels: the plan buffers take their largest size now
    // (the levels' sizes are not known on the host)
    if ((dev || c->chain_levels) && (rc = grow_level(c, c->w.cap_large, c->w.cap + c->w.cap / 4 + 64))) return rc;
    if (single_root) {
        for (int i = 0; i < C_NCNT; ++i) c->h_cnt[i] = 0;
        c->h_cnt[C_LROUND] = 1;
        c->w.h_large[0] = Item{ 0u, n, 0u, F_IDENT };
    } else if (!dev) {
        if ((rc = readback(c, c->w.l_round[0], C_LROUND))) return rc;
    } else if (set_first) {     // the counts setup_roots uploaded
        memcpy(c->h_cnt, c->h_upcnt, C_NCNT * sizeof(uint32_t));
    } else {                    // one root, a tiny leaf or a block
        for (int i = 0; i < C_NCNT; ++i) c->h_cnt[i] = 0;
        c->h_cnt[n <= kTinyT ? C_TINY : C_BLOCK + 2 * block_class(n)] = 1;
    }

    SortRun R{ B, n, alen, fused_root, root_kmax, root_max, root_sum };
    uint32_t start = 0;
    if (c->graph_rounds && !c->profile) {
        // the first graph_rounds rounds as one captured graph, replayed while
        // the sort's inputs, buffers and shape settings are the same
        uint32_t nl_root = 0;
        for (uint32_t s = 0; set_first && s < nsets; ++s) nl_root += set_first[s + 1] - set_first[s] > kBlkT ? 1 : 0;
        if (!set_first) nl_root = n > kBlkT ? 1 : 0;
        std::vector<uint64_t> key = { (uint64_t)(uintptr_t)arena, alen, (uint64_t)(uintptr_t)offs,
                                      (uint64_t)(uintptr_t)offs64, n, (uint64_t)(uintptr_t)perm,
                                      (uint64_t)(uintptr_t)lcp, nsets, (uint64_t)(uintptr_t)c->w.key[0], c->w.cap,
                                      (uint64_t)(uintptr_t)c->w.trees, c->w.trees_cap, c->leaf_target,
                                      c->max_treebits, c->max_oversample, c->graph_rounds, c->seed,
                                      c->w_gen, c->w.fixed ? 1u : 0u, c->graph_levels, c->chunk_max,
                                      c->chunk_max_root, c->min_tiles, c->leaf_keys };
        for (uint32_t s = 0; set_first && s <= nsets; ++s) key.push_back(set_first[s]);
        if ((rc = grow_level(c, c->w.cap_large, c->w.cap + c->w.cap / 4 + 64))) return rc;
        if (!c->graph_exec || key != c->graph_key) {
            if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
            if (c->graph) (void)hipGraphDestroy(c->graph);
            c->graph_exec = nullptr;
            c->graph = nullptr;
            c->graph_key.clear();
            uint32_t L = dev_levels_for(c, root_max);
            if (c->graph_levels) L = std::min(L, c->graph_levels);
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (uint32_t r = 0; r < c->graph_rounds; ++r) graph_round(c, R, r, L, nl_root);
            hipLaunchKernelGGL(k_stage, dim3(1), dim3(kWG), 0, st, c->d_cnt, c->w.l_round[c->graph_rounds & 1],
                               C_LROUND + (int)(c->graph_rounds & 1), c->h_stage_dev, 0u);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(st, &g);
            if (ec != hipSuccess) return set_err(PS5_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
            c->graph = g;
            HIPCHK(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
            c->graph_key = key;
        }
        // the graph's k_stage publishes 0 last; the host marks the word first
        __atomic_store_n(reinterpret_cast<uint32_t*>(c->h_stage + kStageSeq), kGraphSeqPending, __ATOMIC_RELEASE);
        HIPCHK(hipGraphLaunch(c->graph_exec, st));
        if ((rc = wait_stage(c, 0u))) return rc;
        c->st_syncs++;
        c->st_rounds += c->graph_rounds;
        memcpy(c->h_cnt, c->h_stage, C_NCNT * sizeof(uint32_t));
        if (c->h_cnt[C_ERR] & ERR_LEVELS) {
            // a round needed more levels than the graph holds: sort again,
            // every level planned on the host
            if ((rc = setup_roots(c, n, set_first, nsets))) return rc;
            if ((rc = init_root_keys(c, B, n, set_first, fused_root))) return rc;
            if (single_root) {
                for (int i = 0; i < C_NCNT; ++i) c->h_cnt[i] = 0;
                c->h_cnt[C_LROUND] = 1;
                c->w.h_large[0] = Item{ 0u, n, 0u, F_IDENT };
            } else if ((rc = readback(c, c->w.l_round[0], C_LROUND))) {
                return rc;
            }
            dev = false;
        } else {
            if (c->h_cnt[C_ERR] & ERR_CAP) return set_err(PS5_ERR_CAPACITY, "work list capacity exceeded");
            if (c->h_cnt[C_ERR] & ERR_LCP) return set_err(PS5_ERR_INTERNAL, "equal keys across a leaf boundary");
            if (c->h_cnt[C_ERR] & ERR_BOUNDS) return set_err(PS5_ERR_INTERNAL, "device bounds check failed");
            const int rp = c->graph_rounds & 1;
            uint64_t pending = (uint64_t)c->h_cnt[C_LROUND + rp] + c->h_cnt[C_TINY + rp] + c->h_cnt[C_RUN + rp];
            for (int cls = 0; cls < kBlkClasses; ++cls) pending += c->h_cnt[C_BLOCK + 2 * cls + rp];
            if (pending == 0) return PS5_OK;
            // more rounds than captured: continue round by round
            if ((rc = readback(c, c->w.l_round[rp], C_LROUND + rp))) return rc;
            start = c->graph_rounds;
        }
    }
    for (uint32_t round = start;; ++round) {
        bool done = false;
        if ((rc = sort_round(c, R, round, dev, &done))) return rc;
        if (done) break;
    }
    if (c->profile) {
        HIPCHK(hipMemcpyAsync(c->h_sums, c->d_sums, S_NUM * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        // algorithmic bytes per string of the list-driven kernels
        c->k_bytes[K_BLOCK] += c->h_sums[S_BLOCK] * 20;     // key 8 + idx 4 R, perm 4 + lcp 4 W
        c->k_bytes[K_LEAF_TINY] += c->h_sums[S_TINY] * 20;
        c->k_bytes[K_LEAF_RUN] += c->h_sums[S_RUN] * 12;    // idx 4 R, perm 4 + lcp 4 W
        collect_events(c);
    }
    return PS5_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

const char* ps5_last_error(void) { return g_err.c_str(); }

int ps5_ctx_create(ps5_ctx** out, int device)
{
    if (!out) return set_err(PS5_ERR_ARG, "out is NULL");
    *out = nullptr;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipFuncSetAttribute((const void*)k_scatter<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               ScatLds<false>::kBytes));
    HIPCHK(hipFuncSetAttribute((const void*)k_scatter<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               ScatLds<true>::kBytes));
    for (const void* f : { (const void*)k_count<8, false>, (const void*)k_count<9, false>,
                           (const void*)k_count<10, false>, (const void*)k_count<11, false>,
                           (const void*)k_count<8, true>, (const void*)k_count<9, true>,
                           (const void*)k_count<10, true>, (const void*)k_count<11, true> })
        HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, count_lds_bytes(kMaxTreebits)));
    ps5_ctx* c = new ps5_ctx;
    c->device = device;
    c->trace = getenv("PS5_TRACE") != nullptr;
    if (const char* dl = getenv("PS5_DEV_LEVELS")) c->planning = std::min(std::max(atoi(dl), 0), 2);
    if (const char* gr = getenv("PS5_GRAPH_ROUNDS")) c->graph_rounds = std::min<uint32_t>(atoi(gr), kMaxGraphRounds);
    if (const char* gl = getenv("PS5_GRAPH_LEVELS")) c->graph_levels = (uint32_t)std::max(atoi(gl), 0);
    if (const char* tb = getenv("PS5_TAIL_BLOCKS")) c->tail_blocks = atoi(tb) != 0;
    if (const char* cl = getenv("PS5_CHAIN_LEVELS")) c->chain_levels = atoi(cl) != 0;
    int ncu = 0;
    hipError_t e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
    c->scatter_wgs = std::max<uint32_t>(kXcds, (uint32_t)std::max(ncu, 0) / kXcds * kXcds);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&c->d_cnt, C_NCNT * sizeof(uint32_t));
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&c->h_cnt), C_NCNT * sizeof(uint32_t));
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&c->h_upcnt), C_NCNT * sizeof(uint32_t));
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&c->h_stage), kStageHead + kPeek * sizeof(Item),
                          hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void**>(&c->h_stage_dev), c->h_stage, 0);
    if (e == hipSuccess) e = hipMalloc(&c->d_sums, S_NUM * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&c->h_sums), S_NUM * sizeof(unsigned long long));
    // the allocation-free entry point's pinned plan buffers, sized for the
    // scratch's largest sort
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&c->na_h_large), cap_large_of(kNoAllocMaxStrings) * sizeof(Item));
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&c->na_h_plan), plan_bytes_of(kNoAllocMaxStrings),
                          hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void**>(&c->na_h_plan_dev), c->na_h_plan, 0);
    if (e != hipSuccess) {
        ps5_ctx_destroy(c);
        return set_err(PS5_ERR_HIP, "context setup failed: %s", hipGetErrorString(e));
    }
    if (int rc = scratch_address(device, &c->na_scratch)) {
        ps5_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return PS5_OK;
}

void ps5_ctx_destroy(ps5_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_buffers(c);
    if (c->d_cnt) (void)hipFree(c->d_cnt);
    if (c->d_sums) (void)hipFree(c->d_sums);
    if (c->d_sets) (void)hipFree(c->d_sets);
    if (c->d_merge) (void)hipFree(c->d_merge);
    if (c->h_cnt) (void)hipHostFree(c->h_cnt);
    if (c->h_upcnt) (void)hipHostFree(c->h_upcnt);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->h_roots) (void)hipHostFree(c->h_roots);
    if (c->h_sums) (void)hipHostFree(c->h_sums);
    if (c->na_h_large) (void)hipHostFree(c->na_h_large);
    if (c->na_h_plan) (void)hipHostFree(c->na_h_plan);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    for (auto& u : c->ev_used) { (void)hipEventDestroy(u.second.first); (void)hipEventDestroy(u.second.second); }
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    delete c;
}

int ps5_ctx_reserve(ps5_ctx* c, uint64_t n)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    if (n >= 0xFFFFFF00ull) return set_err(PS5_ERR_ARG, "n too large (limit 2^32-256 strings)");
    HIPCHK(hipSetDevice(c->device));
    return ensure_capacity(c, n);
}

void* ps5_ctx_stream(ps5_ctx* c) { return c ? (void*)c->stream : nullptr; }

int ps5_sort_device(ps5_ctx* c, const uint8_t* d_arena, uint64_t arena_bytes, const uint32_t* d_offs,
                    uint64_t n, uint32_t* d_perm, uint32_t* d_lcp)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    if (n == 0) return PS5_OK;
    if (!d_arena || !d_offs || !d_perm) return set_err(PS5_ERR_ARG, "NULL device pointer");
    if (n >= 0xFFFFFF00ull) return set_err(PS5_ERR_ARG, "n too large (limit 2^32-256 strings)");
    if (arena_bytes == 0 || arena_bytes > 0xFFFFFFFFull)
        return set_err(PS5_ERR_ARG, "arena_bytes must be in [1, 2^32)");
    if (((uintptr_t)d_arena & 7u) != 0) return set_err(PS5_ERR_ARG, "arena must be 8-byte aligned");
    HIPCHK(hipSetDevice(c->device));
    int rc = run_sort(c, d_arena, arena_bytes, d_offs, nullptr, (uint32_t)n, d_perm, d_lcp, nullptr, 0);
    if (rc) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->side);
        if (c->profile) collect_events(c);
    }
    return rc;
}

int ps5_sort_device_noalloc(ps5_ctx* c, const uint8_t* d_arena, uint64_t arena_bytes, const uint32_t* d_offs,
                            uint64_t n, uint32_t* d_perm, uint32_t* d_lcp)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    if (n == 0) return PS5_OK;
    if (!d_arena || !d_offs || !d_perm) return set_err(PS5_ERR_ARG, "NULL device pointer");
    if (n > kNoAllocMaxStrings)
        return set_err(PS5_ERR_ARG, "n = %llu exceeds the allocation-free limit of %llu strings",
                       (unsigned long long)n, (unsigned long long)kNoAllocMaxStrings);
    if (arena_bytes == 0 || arena_bytes > 0xFFFFFFFFull)
        return set_err(PS5_ERR_ARG, "arena_bytes must be in [1, 2^32)");
    if (((uintptr_t)d_arena & 7u) != 0) return set_err(PS5_ERR_ARG, "arena must be 8-byte aligned");
    if (!c->na_scratch) return set_err(PS5_ERR_HIP, "the device scratch of this context is not resolved");
    HIPCHK(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(scratch_mutex(c->device));
    // the context's own buffers step aside for views into the scratch;
    // profiling (which creates events) and captured graphs (whose
    // instantiation allocates) are off for the call
    WorkBuffers own = scratch_buffers(c->na_scratch, c->na_h_large, c->na_h_plan, c->na_h_plan_dev);
    std::swap(c->w, own);
    const bool prof = c->profile;
    const uint32_t graph_rounds = c->graph_rounds;
    c->profile = false;
    c->graph_rounds = 0;
    int rc = run_sort(c, d_arena, arena_bytes, d_offs, nullptr, (uint32_t)n, d_perm, d_lcp, nullptr, 0);
    if (rc) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->side);
    }
    c->profile = prof;
    c->graph_rounds = graph_rounds;
    std::swap(c->w, own);
    return rc;
}

uint64_t ps5_noalloc_max_strings(void) { return kNoAllocMaxStrings; }

int ps5_merge_device(ps5_ctx* c, const uint8_t* d_arena, uint64_t arena_bytes, const uint32_t* d_offs,
                     const uint64_t* run_first, uint32_t nruns, const uint32_t* d_idx, const uint32_t* d_lcp,
                     uint32_t* d_idx_out, uint32_t* d_lcp_out)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    if (nruns == 0 || !run_first) return set_err(PS5_ERR_ARG, "no runs");
    const uint64_t n = run_first[nruns];
    if (run_first[0] != 0) return set_err(PS5_ERR_ARG, "run_first[0] must be 0");
    for (uint32_t r = 0; r < nruns; ++r)
        if (run_first[r + 1] < run_first[r]) return set_err(PS5_ERR_ARG, "run boundaries must be non-decreasing");
    if (n == 0) return PS5_OK;
    if (!d_arena || !d_offs || !d_idx || !d_lcp || !d_idx_out || !d_lcp_out)
        return set_err(PS5_ERR_ARG, "NULL device pointer");
    if (n >= 0xFFFFFF00ull) return set_err(PS5_ERR_ARG, "n too large (limit 2^32-256 strings)");
    if (arena_bytes == 0 || arena_bytes > 0xFFFFFFFFull) return set_err(PS5_ERR_ARG, "arena_bytes must be in [1, 2^32)");
    if (((uintptr_t)d_arena & 7u) != 0) return set_err(PS5_ERR_ARG, "arena must be 8-byte aligned");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    if (nruns == 1) {
        HIPCHK(hipMemcpyAsync(d_idx_out, d_idx, n * 4, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(d_lcp_out, d_lcp, n * 4, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        return PS5_OK;
    }
    int rc = ensure_capacity(c, n);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(c->d_cnt, 0, C_NCNT * sizeof(uint32_t), st));
    Bufs B{};
    B.n = (uint32_t)n;
    B.cnt = c->d_cnt;
    B.arena = d_arena;
    B.alen = arena_bytes;
    B.offs = d_offs;
    // ping-pong buffers: indices in idx[0]/idx[1], first 8 characters in
    // key[0]/key[1], LCPs in the two halves of bkey
    uint32_t* ti[2] = { c->w.idx[0], c->w.idx[1] };
    key64* tk[2] = { c->w.key[0], c->w.key[1] };
    uint32_t* tl[2] = { reinterpret_cast<uint32_t*>(c->w.bkey), reinterpret_cast<uint32_t*>(c->w.bkey) + c->w.cap };
    std::vector<uint64_t> bounds(run_first, run_first + nruns + 1);
    const uint32_t* si = d_idx;
    const uint32_t* sl = d_lcp;
    // the first round reads every string's first 8 characters from ik, keys
    // extracted in input order by one streaming pass (the previous round's
    // output carries them afterwards); key[1] is free until the second merge round writes it
    key64* ik = c->w.key[1];
    {
        Bufs BI = B;
        BI.key0 = ik;
        hipLaunchKernelGGL(k_init, dim3(div_up(n, kWG)), dim3(kWG), 0, st, BI, (uint32_t)n);
        HIPCHK(hipGetLastError());
    }
    const key64* sk = nullptr;
    int cur = 0;
    while (bounds.size() > 2) {
        const uint32_t nr = (uint32_t)bounds.size() - 1, np = (nr + 1) / 2;
        std::vector<uint64_t> plan(3 * (np + 1), 0);   // first[np+1], mid[np], tstart[np+1]
        uint64_t* f = plan.data();
        uint64_t* mdl = f + (np + 1);
        uint64_t* ts = mdl + (np + 1);
        for (uint32_t p = 0; p < np; ++p) {
            f[p] = bounds[2 * p];
            mdl[p] = bounds[2 * p + 1];
            const uint64_t e = 2 * p + 2 <= nr ? bounds[2 * p + 2] : bounds[2 * p + 1];
            ts[p + 1] = ts[p] + (e - f[p] + kMergeT - 1) / kMergeT;
        }
        f[np] = bounds[nr];
        const uint64_t tiles = ts[np];
        const bool last = np == 1;
        uint32_t* di = last ? d_idx_out : ti[cur];
        uint32_t* dl = last ? d_lcp_out : tl[cur];
        key64* dk = last ? nullptr : tk[cur];
        const uint64_t words = plan.size() + (tiles + 1) / 2;    // plan, then co[tiles] as uint32
        if (words > c->merge_cap) {
            if (c->d_merge) (void)hipFree(c->d_merge);
            c->d_merge = nullptr;
            c->merge_cap = 0;
            HIPCHK(hipMalloc(&c->d_merge, words * sizeof(uint64_t)));
            c->merge_cap = words;
        }
        HIPCHK(hipMemcpyAsync(c->d_merge, plan.data(), plan.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        if (tiles) {
            const uint64_t* pf = c->d_merge;
            const uint64_t* pm = pf + (np + 1);
            const uint64_t* pt = pm + (np + 1);
            uint32_t* co = reinterpret_cast<uint32_t*>(c->d_merge + plan.size());
            hipLaunchKernelGGL(k_merge_part, dim3(div_up(tiles, kWG)), dim3(kWG), 0, st, B, si, sk, ik, pf, pm, pt, np,
                               co);
            HIPCHK(hipGetLastError());
            hipLaunchKernelGGL(k_merge_tile, dim3(tiles), dim3(kWG), 0, st, B, si, sl, sk, ik, di, dl, dk, pf, pm, pt,
                               np, co);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(st));     // the plan buffer is rewritten next round
        std::vector<uint64_t> nb(np + 1);
        for (uint32_t p = 0; p <= np; ++p) nb[p] = f[p];
        bounds.swap(nb);
        si = di;
        sl = dl;
        sk = dk;
        cur ^= 1;
    }
    return readback(c, nullptr, 0);
}

int ps5_dist_chars_device(ps5_ctx* c, const uint8_t* d_arena, uint64_t arena_bytes, const uint32_t* d_offs,
                          uint64_t n, const uint32_t* d_perm, const uint32_t* d_lcp, uint8_t* d_cache)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    if (n < 2) return PS5_OK;
    if (!d_arena || !d_offs || !d_perm || !d_lcp || !d_cache) return set_err(PS5_ERR_ARG, "NULL device pointer");
    if (n >= 0xFFFFFF00ull) return set_err(PS5_ERR_ARG, "n too large (limit 2^32-256 strings)");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemsetAsync(c->d_cnt, 0, C_NCNT * sizeof(uint32_t), c->stream));
    hipLaunchKernelGGL(k_dist_chars, dim3(div_up(n, kWG)), dim3(kWG), 0, c->stream, d_arena, d_offs, d_perm, d_lcp,
                       (uint32_t)n, d_cache, c->d_cnt, arena_bytes);
    HIPCHK(hipGetLastError());
    int rc = readback(c, nullptr, 0);
    return rc;
}

int ps5_sort_batch_device(ps5_ctx* c, const uint8_t* d_arena, uint64_t arena_bytes, const uint64_t* d_offs,
                          const uint64_t* set_first, uint32_t nsets, uint32_t* d_perm, uint32_t* d_lcp)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    if (!set_first || nsets == 0) return set_err(PS5_ERR_ARG, "no sets");
    const uint64_t n = set_first[nsets];
    if (n == 0) return PS5_OK;
    if (!d_arena || !d_offs || !d_perm) return set_err(PS5_ERR_ARG, "NULL device pointer");
    if (n >= 0xFFFFFF00ull) return set_err(PS5_ERR_ARG, "n too large (limit 2^32-256 strings)");
    if (set_first[0] != 0) return set_err(PS5_ERR_ARG, "set_first[0] must be 0");
    if (arena_bytes == 0) return set_err(PS5_ERR_ARG, "empty arena");
    if (((uintptr_t)d_arena & 7u) != 0) return set_err(PS5_ERR_ARG, "arena must be 8-byte aligned");
    HIPCHK(hipSetDevice(c->device));
    int rc = run_sort(c, d_arena, arena_bytes, nullptr, d_offs, (uint32_t)n, d_perm, d_lcp, set_first, nsets);
    if (rc == PS5_OK) {
        if (c->sets_cap < (uint64_t)nsets + 1) {
            if (c->d_sets) (void)hipFree(c->d_sets);
            c->d_sets = nullptr;
            c->sets_cap = 0;
            HIPCHK(hipMalloc(&c->d_sets, ((uint64_t)nsets + 1) * sizeof(uint64_t)));
            c->sets_cap = (uint64_t)nsets + 1;
        }
        HIPCHK(hipMemcpyAsync(c->d_sets, set_first, ((uint64_t)nsets + 1) * sizeof(uint64_t), hipMemcpyHostToDevice,
                              c->stream));
        hipLaunchKernelGGL(k_localize, dim3(div_up(n, kLocTile)), dim3(kWG), 0, c->stream, d_perm, (uint32_t)n, c->d_sets,
                           nsets);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
    } else {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->side);
        if (c->profile) collect_events(c);
    }
    return rc;
}

int ps5_sort_device_timed(ps5_ctx* c, const uint8_t* d_arena, uint64_t arena_bytes, const uint32_t* d_offs,
                          uint64_t n, uint32_t* d_perm, uint32_t* d_lcp, double* ms)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    hipEvent_t a, b;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    (void)hipEventRecord(a, c->stream);
    int rc = ps5_sort_device(c, d_arena, arena_bytes, d_offs, n, d_perm, d_lcp);
    (void)hipEventRecord(b, c->stream);
    (void)hipEventSynchronize(b);
    float f = 0;
    (void)hipEventElapsedTime(&f, a, b);
    if (ms) *ms = f;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

int ps5_ctx_set_tuning(ps5_ctx* c, uint32_t leaf_target, uint32_t max_treebits, uint32_t max_oversample)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    if (leaf_target < kTinyT || leaf_target > kMaxLeafTarget)
        return set_err(PS5_ERR_ARG, "leaf_target must be in [64, 1024]");
    if (max_treebits < 1 || max_treebits > (uint32_t)kMaxTreebits)
        return set_err(PS5_ERR_ARG, "max_treebits must be in [1, 10]");
    if (max_oversample < 2 || max_oversample > 32 || (max_oversample & 1))
        return set_err(PS5_ERR_ARG, "max_oversample must be even and in [2, 32]");
    c->leaf_target = leaf_target;
    c->max_treebits = max_treebits;
    c->max_oversample = max_oversample;
    return PS5_OK;
}

int ps5_ctx_set_shape(ps5_ctx* c, uint32_t chunk_max, uint32_t chunk_max_root, uint32_t min_tiles, uint32_t leaf_keys)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    if (chunk_max < (uint32_t)kChunkMin || chunk_max > (uint32_t)kChunkMax)
        return set_err(PS5_ERR_ARG, "chunk_max must be in [%d, %d]", kChunkMin, kChunkMax);
    if (chunk_max_root < (uint32_t)kChunkMin || chunk_max_root > (uint32_t)kChunkMaxIdent)
        return set_err(PS5_ERR_ARG, "chunk_max_root must be in [%d, %d]", kChunkMin, kChunkMaxIdent);
    if (min_tiles < 1 || min_tiles > 65536) return set_err(PS5_ERR_ARG, "min_tiles must be in [1, 65536]");
    if (leaf_keys < 1 || leaf_keys > 8) return set_err(PS5_ERR_ARG, "leaf_keys must be in [1, 8]");
    c->chunk_max = chunk_max;
    c->chunk_max_root = chunk_max_root;
    c->min_tiles = min_tiles;
    c->leaf_keys = leaf_keys;
    return PS5_OK;
}

int ps5_ctx_get_shape(ps5_ctx* c, uint32_t* chunk_max, uint32_t* chunk_max_root, uint32_t* min_tiles,
                      uint32_t* leaf_keys)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    if (chunk_max) *chunk_max = c->chunk_max;
    if (chunk_max_root) *chunk_max_root = c->chunk_max_root;
    if (min_tiles) *min_tiles = c->min_tiles;
    if (leaf_keys) *leaf_keys = c->leaf_keys;
    return PS5_OK;
}

int ps5_ctx_get_tuning(ps5_ctx* c, uint32_t* leaf_target, uint32_t* max_treebits, uint32_t* max_oversample)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    if (leaf_target) *leaf_target = c->leaf_target;
    if (max_treebits) *max_treebits = c->max_treebits;
    if (max_oversample) *max_oversample = c->max_oversample;
    return PS5_OK;
}

int ps5_ctx_set_planning(ps5_ctx* c, int mode)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    if (mode < 0 || mode > 2) return set_err(PS5_ERR_ARG, "planning mode %d is not 0 (host), 1 (device) or 2 (auto)", mode);
    c->planning = mode;
    return PS5_OK;
}

int ps5_ctx_get_planning(ps5_ctx* c)
{
    return c ? c->planning : -1;
}

int ps5_ctx_set_graph_rounds(ps5_ctx* c, uint32_t rounds)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    if (rounds > kMaxGraphRounds) return set_err(PS5_ERR_ARG, "at most %u captured rounds", kMaxGraphRounds);
    if (rounds != c->graph_rounds) {
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
        if (c->graph) (void)hipGraphDestroy(c->graph);
        c->graph_exec = nullptr;
        c->graph = nullptr;
        c->graph_key.clear();
    }
    c->graph_rounds = rounds;
    return PS5_OK;
}

int ps5_ctx_get_graph_rounds(ps5_ctx* c)
{
    return c ? (int)c->graph_rounds : -1;
}

int ps5_ctx_set_profile(ps5_ctx* c, int on)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    c->profile = on != 0;
    return PS5_OK;
}

int ps5_ctx_kernel_count(ps5_ctx*) { return K_NUM; }

int ps5_ctx_kernel_stat(ps5_ctx* c, int i, const char** name, double* ms, uint64_t* launches, uint64_t* bytes)
{
    if (!c || i < 0 || i >= K_NUM) return set_err(PS5_ERR_ARG, "bad kernel index");
    if (name) *name = kKernelNames[i];
    if (ms) *ms = c->k_ms[i];
    if (launches) *launches = c->k_launch[i];
    if (bytes) *bytes = c->k_bytes[i];
    return PS5_OK;
}

int ps5_ctx_work_stats(ps5_ctx* c, uint64_t* rounds, uint64_t* levels, uint64_t* segments, uint64_t* leaves,
                       uint64_t* syncs)
{
    if (!c) return set_err(PS5_ERR_ARG, "ctx is NULL");
    if (rounds) *rounds = c->st_rounds;
    if (levels) *levels = c->st_levels;
    if (segments) *segments = c->st_segments;
    if (leaves) *leaves = c->st_leaves;
    if (syncs) *syncs = c->st_syncs;
    return PS5_OK;
}

int ps5_dev_malloc(void** p, uint64_t bytes)
{
    if (!p) return set_err(PS5_ERR_ARG, "p is NULL");
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess)
        return set_err(PS5_ERR_NOMEM, "hipMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
    return PS5_OK;
}

int ps5_dev_free(void* p)
{
    if (p) HIPCHK(hipFree(p));
    return PS5_OK;
}

int ps5_memcpy_h2d(void* dst, const void* src, uint64_t bytes)
{
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return PS5_OK;
}

int ps5_memcpy_d2h(void* dst, const void* src, uint64_t bytes)
{
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return PS5_OK;
}

int ps5_dev_copy(void* dst, const void* src, uint64_t bytes)
{
    if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
        return set_err(PS5_ERR_ARG, "ps5_dev_copy needs 16-byte aligned pointers and size");
    if (bytes == 0) return PS5_OK;
    const uint64_t n16 = bytes / 16;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n16 + kWG - 1) / kWG, 65536);
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(kWG), 0, nullptr, (const uint4*)src, (uint4*)dst, n16);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return PS5_OK;
}

int ps5_device_synchronize(void)
{
    HIPCHK(hipDeviceSynchronize());
    return PS5_OK;
}

}  // extern "C"

namespace ps5 {

int set_error(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int widen_offsets(ps5_ctx* c, const uint32_t* d_offs32, const uint64_t* d_first, const uint64_t* d_base,
                  uint32_t nsets, uint64_t n, uint64_t* d_offs64)
{
    if (!c || nsets == 0) return set_err(PS5_ERR_ARG, "widen_offsets: no context or no sets");
    if (n == 0) return PS5_OK;
    HIPCHK(hipSetDevice(c->device));
    hipLaunchKernelGGL(k_widen, dim3(div_up(n, kWG)), dim3(kWG), 0, c->stream, d_offs32, d_first, d_base, nsets, n,
                       d_offs64);
    HIPCHK(hipGetLastError());
    return PS5_OK;
}


int gather_pointers(ps5_ctx* c, const uint32_t* d_perm, const uint32_t* d_offs, uint64_t base, uint64_t n,
                    uint64_t* d_ptrs)
{
    if (!c) return set_err(PS5_ERR_ARG, "gather_pointers: no context");
    if (n == 0) return PS5_OK;
    HIPCHK(hipSetDevice(c->device));
    hipLaunchKernelGGL(k_gather_ptrs, dim3(div_up(n, kWG)), dim3(kWG), 0, c->stream, d_perm, d_offs, base, n, d_ptrs);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return PS5_OK;
}

int widen_lcps(ps5_ctx* c, const uint32_t* d_lcp, uint64_t n, uint64_t* d_wide)
{
    if (!c) return set_err(PS5_ERR_ARG, "widen_lcps: no context");
    if (n == 0) return PS5_OK;
    HIPCHK(hipSetDevice(c->device));
    hipLaunchKernelGGL(k_widen_lcp, dim3(div_up(n, kWG)), dim3(kWG), 0, c->stream, d_lcp, n, d_wide);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return PS5_OK;
}

}  // namespace ps5
