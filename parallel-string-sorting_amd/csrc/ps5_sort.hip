// ps5_sort.hip -- pS5 string sample sort on MI355X (gfx950), HIP kernels and
// the host orchestration behind the C ABI in include/ps5_mi355x.h.
//
// Reference algorithm (paths relative to the reference checkout):
//   SampleSortStep sample/count/distribute  src/parallel/bingmann-parallel_sample_sort.hpp:1264-1611
//   splitter tree classify                   src/sequential/bingmann-sample_sortBTC.hpp:45-72,297-399
//   bucket recursion rules                   src/parallel/bingmann-parallel_sample_sort.hpp:1492-1584
//   bucket-boundary LCPs                     src/parallel/bingmann-parallel_sample_sort.hpp:311-396
//   key extraction                           src/tools/stringset.hpp:201-231
// The GPU formulation keeps the reference's outputs but reorganises the work
// breadth-first: every round handles all segments at one 8-byte depth, and
// every pass is a kernel over a device-side work list.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ps5_kernels.hpp"
#include "../../include/ps5_mi355x.h"

using namespace ps5;

// ===========================================================================
// device helpers
// ===========================================================================

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// 8 characters of the arena at byte offset `off`, big-endian, zeroed from the
// first NUL on (the reference's get_char_uint64_simple, stringset.hpp:201-231).
// Two aligned 8-byte loads and a funnel shift; near the arena end the bytes
// are fetched one by one so nothing past arena[alen-1] is read.
__device__ __forceinline__ key64 load_key(const uint8_t* arena, uint64_t alen, uint64_t off)
{
    const uint64_t base = (uint64_t)(uintptr_t)arena;
    const uint64_t a = base + off;
    const uint64_t a8 = a & ~7ull;
    uint64_t lo, hi;
    if (a8 + 16 <= base + alen) {
        lo = *(const uint64_t*)(uintptr_t)a8;
        hi = *(const uint64_t*)(uintptr_t)(a8 + 8);
    } else {
        lo = 0;
        hi = 0;
        for (int i = 0; i < 16; ++i) {
            uint64_t p = a8 + i;
            if (p >= base && p < base + alen) {
                uint64_t c = *(const uint8_t*)(uintptr_t)p;
                if (i < 8) lo |= c << (8 * i);
                else hi |= c << (8 * (i - 8));
            }
        }
    }
    const uint32_t sh = (uint32_t)(a & 7) * 8;
    uint64_t v = sh ? ((lo >> sh) | (hi << (64 - sh))) : lo;
    v = __builtin_bswap64(v);
    // high bit of each byte of t is set iff that byte is non-zero (no carries)
    const uint64_t t = ((v & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full) | v;
    const uint64_t z = ~t & 0x8080808080808080ull;
    if (z) {
        const uint32_t p = (uint32_t)__clzll((long long)z);   // 8 * index of first NUL
        v = p ? (v & (~0ull << (64 - p))) : 0ull;
    }
    return v;
}

// number of leading characters the two keys share (lcpKeyType, .hpp:225-230)
__device__ __forceinline__ uint32_t lcp_key(key64 a, key64 b)
{
    key64 x = a ^ b;
    return x ? ((uint32_t)__clzll((long long)x) >> 3) : 8u;
}
// number of non-NUL characters of a key (lcpKeyDepth, .hpp:232-237)
__device__ __forceinline__ uint32_t key_depth(key64 a)
{
    return a ? 8u - ((uint32_t)__builtin_ctzll(a) >> 3) : 0u;
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

__device__ __forceinline__ key64 readlane64(key64 v, uint32_t l)
{
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)l);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)l);
    return ((key64)hi << 32) | lo;
}

// lanes holding the same 11-bit bucket id (all 64 lanes must execute this)
__device__ __forceinline__ uint64_t match_bucket(uint32_t b)
{
    uint64_t peers = __ballot(1);
#pragma unroll
    for (int i = 0; i < 11; ++i) {
        const bool bit = (b >> i) & 1u;
        const uint64_t bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
    }
    return peers;
}

// Append for every lane with pred: returns its slot (wave-aggregated atomic).
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t* counter)
{
    const uint64_t m = __ballot(pred);
    if (m == 0) return 0xFFFFFFFFu;
    const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, (int)leader, 64);
    return base + (uint32_t)__popcll(m & lanemask_lt());
}

__device__ __forceinline__ void flag_error(uint32_t* cnt, uint32_t e) { atomicOr(&cnt[C_ERR], e); }

// an item must describe a non-empty range inside [0, n)
__device__ __forceinline__ bool item_ok(const Item& it, const Bufs& B)
{
    return it.size != 0 && it.begin < B.n && it.size <= B.n - it.begin;
}

__device__ __forceinline__ void put_item(Item* list, uint32_t cap, uint32_t slot, Item it, uint32_t* cnt)
{
    if (slot < cap) list[slot] = it;
    else atomicOr(&cnt[C_ERR], ERR_CAP);
}

// append one next-round segment (keys already at the new depth, buffer 0)
__device__ __forceinline__ void wave_append_segment(bool pred, Item it, const Sink& s)
{
    const uint32_t m = it.size;
    const bool t = pred && m <= kTinyT;
    const bool sm = pred && m > kTinyT && m <= kSmallT;
    const bool md = pred && m > kSmallT && m <= kMedT;
    const bool lg = pred && m > kMedT;
    uint32_t slot;
    slot = wave_append(t, &s.cnt[s.c_tiny]);
    if (t) put_item(s.tiny, s.cap_tiny, slot, it, s.cnt);
    slot = wave_append(sm, &s.cnt[s.c_small]);
    if (sm) put_item(s.small, s.cap_small, slot, it, s.cnt);
    slot = wave_append(md, &s.cnt[s.c_med]);
    if (md) put_item(s.med, s.cap_med, slot, it, s.cnt);
    slot = wave_append(lg, &s.cnt[s.c_large]);
    if (lg) put_item(s.large, s.cap_large, slot, it, s.cnt);
}

// Exclusive block scan of one uint32 per thread; *total gets the sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total, uint32_t* s_tmp)
{
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = (uint32_t)__shfl_up((int)x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) s_tmp[w] = x;
    __syncthreads();
    uint32_t wpre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) {
        uint32_t t = s_tmp[i];
        wpre += (i < (int)w) ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return wpre + x - v;
}

// Bitonic sort of P (power of two) uint64 in LDS by the whole workgroup.
__device__ void bitonic_sort_u64(key64* s, int P)
{
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += kWG) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const key64 a = s[i], b = s[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) { s[i] = b; s[ixj] = a; }
                }
            }
            __syncthreads();
        }
    }
}

// Splitters from 2*nspl random keys of the segment (oversample factor 2,
// bingmann-sample_sort.hpp:57): sort the sample, take every second key
// (splitter i = sample 2i+1) and lay the tree out in level order so that
// descending i = 2i + (key > tree[i]) computes the lower bound.
__device__ void build_tree(const key64* K, uint32_t beg, uint32_t m, uint64_t seed, int k,
                           key64* s_samp, key64* s_tree, key64* s_spl)
{
    const int nspl = (1 << k) - 1;
    const int ss = 2 * nspl;
    int P = 2;
    while (P < ss) P <<= 1;
    for (int j = threadIdx.x; j < P; j += kWG) {
        key64 v = ~0ull;
        if (j < ss) {
            const uint64_t h = mix64(seed ^ ((uint64_t)beg << 21) ^ (uint64_t)j);
            const uint32_t pos = (uint32_t)(((h >> 32) * (uint64_t)m) >> 32);
            v = K[beg + pos];
        }
        s_samp[j] = v;
    }
    __syncthreads();
    bitonic_sort_u64(s_samp, P);
    for (int t = threadIdx.x + 1; t <= nspl; t += kWG) {
        const int l = 31 - __clz(t);
        const int j = t - (1 << l);
        const int inorder = ((2 * j + 1) << (k - l - 1)) - 1;
        s_tree[t] = s_samp[2 * inorder + 1];
    }
    for (int i = threadIdx.x; i < nspl; i += kWG) s_spl[i] = s_samp[2 * i + 1];
    __syncthreads();
}

// bucket of a key: b = 2 * lower_bound + [splitter == key]
// (ClassifyTreeSimple::find_bkt, bingmann-sample_sortBTC.hpp:45-72)
__device__ __forceinline__ uint32_t classify_key(key64 key, const key64* s_tree, const key64* s_spl,
                                                 int k, int nspl)
{
    uint32_t i = 1;
    for (int l = 0; l < k; ++l) i = 2 * i + (key > s_tree[i] ? 1u : 0u);
    i -= (1u << k);
    uint32_t b = 2 * i;
    if ((int)i < nspl && s_spl[i] == key) b += 1;
    return b;
}

// Emit the children buckets of a sample-sort step.  start(b) = bstart[b*stride],
// relative to seg.begin; the last bucket ends at seg.size.  Equal buckets
// (odd b) are sorted leaves, the rest by size: tiny / small leaves, medium /
// large segments of the next level (distribute_finished, .hpp:1492-1584).
__device__ void emit_children(const Item& seg, int nbkt, const uint32_t* bstart, uint32_t stride,
                              const Sink& s, uint32_t* s_tmp)
{
    constexpr int kPer = (kMaxBkt + kWG - 1) / kWG;   // 8
    const uint32_t dstbuf = (seg.flags & F_BUFB) ^ F_BUFB;
    uint32_t c_t = 0, c_s = 0, c_m = 0, c_l = 0, c_r = 0;
    uint32_t st[kPer], sz[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int b = threadIdx.x * kPer + j;
        st[j] = 0;
        sz[j] = 0;
        if (b < nbkt) {
            st[j] = bstart[(uint32_t)b * stride];
            const uint32_t en = (b + 1 < nbkt) ? bstart[(uint32_t)(b + 1) * stride] : seg.size;
            sz[j] = en - st[j];
            const uint32_t m = sz[j];
            if (m == 0) continue;
            if ((b & 1) && m > kTinyT) c_r += (m + kRunChunk - 1) / kRunChunk;
            else if (m <= kTinyT) c_t++;
            else if (m <= kSmallT) c_s++;
            else if (m <= kMedT) c_m++;
            else c_l++;
        }
    }
    uint32_t tt, ts, tm, tl, tr;
    uint32_t o_t = block_excl_scan(c_t, &tt, s_tmp);
    uint32_t o_s = block_excl_scan(c_s, &ts, s_tmp);
    uint32_t o_m = block_excl_scan(c_m, &tm, s_tmp);
    uint32_t o_l = block_excl_scan(c_l, &tl, s_tmp);
    uint32_t o_r = block_excl_scan(c_r, &tr, s_tmp);
    if (threadIdx.x == 0) {
        s_tmp[8] = tt ? atomicAdd(&s.cnt[s.c_tiny], tt) : 0;
        s_tmp[9] = ts ? atomicAdd(&s.cnt[s.c_small], ts) : 0;
        s_tmp[10] = tm ? atomicAdd(&s.cnt[s.c_med], tm) : 0;
        s_tmp[11] = tl ? atomicAdd(&s.cnt[s.c_large], tl) : 0;
        s_tmp[12] = tr ? atomicAdd(&s.cnt[s.c_run], tr) : 0;
    }
    __syncthreads();
    o_t += s_tmp[8];
    o_s += s_tmp[9];
    o_m += s_tmp[10];
    o_l += s_tmp[11];
    o_r += s_tmp[12];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int b = threadIdx.x * kPer + j;
        const uint32_t m = sz[j];
        if (b >= nbkt || m == 0) continue;
        Item it;
        it.begin = seg.begin + st[j];
        it.size = m;
        it.rbeg = seg.rbeg;
        it.flags = dstbuf;
        if ((b & 1) && m > kTinyT) {
            const uint32_t nch = (m + kRunChunk - 1) / kRunChunk;
            for (uint32_t c = 0; c < nch; ++c) {
                Item r = it;
                r.flags = dstbuf | F_SORTED | (c << 8);
                put_item(s.run, s.cap_run, o_r++, r, s.cnt);
            }
        } else if (m <= kTinyT) {
            put_item(s.tiny, s.cap_tiny, o_t++, it, s.cnt);
        } else if (m <= kSmallT) {
            put_item(s.small, s.cap_small, o_s++, it, s.cnt);
        } else if (m <= kMedT) {
            put_item(s.med, s.cap_med, o_m++, it, s.cnt);
        } else {
            put_item(s.large, s.cap_large, o_l++, it, s.cnt);
        }
    }
}

// ===========================================================================
// kernels
// ===========================================================================

// Round 0 keys: key0[i] = 8 characters of string i at depth 0, idx0[i] = i.
__global__ __launch_bounds__(kWG) void k_init(Bufs B, uint32_t n)
{
    const uint32_t i = blockIdx.x * kWG + threadIdx.x;
    if (i >= n) return;
    B.key0[i] = load_key(B.arena, B.alen, B.offs[i]);
    B.idx0[i] = i;
}

// Reset counters and place the root item.
__global__ void k_setup(uint32_t* cnt, Item* list, int c_index, uint32_t n)
{
    if (threadIdx.x < C_NCNT) cnt[threadIdx.x] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        Item it = { 0u, n, 0u, 0u };
        list[0] = it;
        cnt[c_index] = 1;
    }
}

__global__ void k_reset(uint32_t* cnt, int a, int b, int c, int d, int e)
{
    const int t = threadIdx.x;
    if (t == 0 && a >= 0) cnt[a] = 0;
    if (t == 1 && b >= 0) cnt[b] = 0;
    if (t == 2 && c >= 0) cnt[c] = 0;
    if (t == 3 && d >= 0) cnt[d] = 0;
    if (t == 4 && e >= 0) cnt[e] = 0;
}

// ---- multi-workgroup sample-sort step (segments > kMedT) -------------------

// Splitter tree of one large segment -> tree[0..1023] level order,
// tree[1024..2047] in order.
__global__ __launch_bounds__(kWG) void k_tree(Item seg, Bufs B, uint64_t seed, key64* tree)
{
    __shared__ key64 s_samp[2048];
    __shared__ key64 s_tree[1024];
    __shared__ key64 s_spl[1024];
    const key64* K = (seg.flags & F_BUFB) ? B.key1 : B.key0;
    build_tree(K, seg.begin, seg.size, seed, kMaxTreebits, s_samp, s_tree, s_spl);
    for (int i = threadIdx.x; i < 1024; i += kWG) {
        tree[i] = s_tree[i];
        tree[1024 + i] = s_spl[i];
    }
}

// Classify one tile of kTile keys.  Each wave ranks its own chunk of
// kWaveChunk keys stably (ballot matching, ranks follow input order) and the
// chunk counts go to hist[(b * T + t) * kWaves + w]: bucket-major, then tile,
// then wave -- the order of the reference's prefix sum (.hpp:1437-1460) with
// wave chunks as the parts -- so an exclusive scan of hist gives every chunk's
// base offset per bucket.  bktrank gets bucket and rank inside the chunk.
__global__ __launch_bounds__(kWG) void k_count(Item seg, Bufs B, const key64* tree, uint32_t T,
                                                uint32_t* hist)
{
    __shared__ key64 s_tree[1024];
    __shared__ key64 s_spl[1024];
    __shared__ uint32_t s_hist[kWaves][2048];
    const int k = kMaxTreebits, nspl = kMaxSpl, nbkt = kMaxBkt;
    for (int i = threadIdx.x; i < 1024; i += kWG) {
        s_tree[i] = tree[i];
        s_spl[i] = tree[1024 + i];
    }
    for (int i = threadIdx.x; i < kWaves * 2048; i += kWG) (&s_hist[0][0])[i] = 0;
    __syncthreads();
    const key64* K = (seg.flags & F_BUFB) ? B.key1 : B.key0;
    const uint32_t t = blockIdx.x;
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint32_t cb = t * kTile + w * kWaveChunk;
    const uint32_t ce = item_ok(seg, B) ? min(seg.size, cb + kWaveChunk) : 0u;
    for (uint32_t g = cb; g < ce; g += 64) {
        const uint32_t p = g + lane;
        const bool act = p < ce;
        uint32_t b = kNoBucket;
        if (act) b = classify_key(K[seg.begin + p], s_tree, s_spl, k, nspl);
        const uint64_t peers = match_bucket(b);
        const uint32_t r = (uint32_t)__popcll(peers & lanemask_lt());
        if (act) {
            const uint32_t old = s_hist[w][b];
            if (r == 0) s_hist[w][b] = old + (uint32_t)__popcll(peers);
            B.bktrank[seg.begin + p] = (b << kRankBits) | (old + r);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbkt * kWaves; i += kWG) {
        const uint32_t b = (uint32_t)i / kWaves, ww = (uint32_t)i % kWaves;
        hist[((uint32_t)b * T + t) * kWaves + ww] = s_hist[ww][b];
    }
}

// exclusive scan of hist[0..len): block sums
constexpr int kScanPer = 16;
constexpr int kScanBlock = kWG * kScanPer;   // 4096 entries per block

__global__ __launch_bounds__(kWG) void k_scan_reduce(const uint32_t* in, uint32_t len, uint32_t* part)
{
    __shared__ uint32_t s_tmp[16];
    const uint32_t base = blockIdx.x * kScanBlock;
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        const uint32_t i = base + j * kWG + threadIdx.x;
        if (i < len) sum += in[i];
    }
    uint32_t tot;
    block_excl_scan(sum, &tot, s_tmp);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// exclusive scan of the block sums (one workgroup, any count)
__global__ __launch_bounds__(kWG) void k_scan_top(uint32_t* part, uint32_t nparts)
{
    __shared__ uint32_t s_tmp[16];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nparts; base += kWG) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nparts ? part[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, &tot, s_tmp);
        if (i < nparts) part[i] = carry + ex;
        carry += tot;
    }
}

// exclusive scan inside each block, offset by its block prefix (in place)
__global__ __launch_bounds__(kWG) void k_scan_down(uint32_t* data, uint32_t len, const uint32_t* part)
{
    __shared__ uint32_t s_tmp[16];
    const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanPer;
    uint32_t v[kScanPer];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        const uint32_t i = base + j;
        v[j] = i < len ? data[i] : 0u;
        sum += v[j];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan(sum, &tot, s_tmp) + part[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        const uint32_t i = base + j;
        if (i < len) data[i] = run;
        run += v[j];
    }
}

// Move each key/index of a tile to seg.begin + base(bucket, tile, wave) + rank
// (distribute, .hpp:1464-1490, but stable: ranks follow input order).
__global__ __launch_bounds__(kWG) void k_scatter(Item seg, Bufs B, uint32_t T, const uint32_t* hscan)
{
    __shared__ uint32_t s_off[kWaves][2048];
    const uint32_t t = blockIdx.x;
    for (int i = threadIdx.x; i < kMaxBkt * kWaves; i += kWG) {
        const uint32_t b = (uint32_t)i / kWaves, ww = (uint32_t)i % kWaves;
        s_off[ww][b] = hscan[((uint32_t)b * T + t) * kWaves + ww];
    }
    __syncthreads();
    const int src = seg.flags & F_BUFB;
    const key64* K = src ? B.key1 : B.key0;
    const uint32_t* I = src ? B.idx1 : B.idx0;
    key64* KD = src ? B.key0 : B.key1;
    uint32_t* ID = src ? B.idx0 : B.idx1;
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint32_t cb = t * kTile + w * kWaveChunk;
    const uint32_t ce = item_ok(seg, B) ? min(seg.size, cb + kWaveChunk) : 0u;
    for (uint32_t p = cb + lane; p < ce; p += 64) {
        const uint32_t q = seg.begin + p;
        const uint32_t br = B.bktrank[q];
        const uint32_t b = br >> kRankBits;
        const uint32_t rel = b < (uint32_t)kMaxBkt ? s_off[w][b] + (br & kRankMask) : 0xFFFFFFFFu;
        if (rel >= seg.size) { flag_error(B.cnt, ERR_BOUNDS); continue; }
        KD[seg.begin + rel] = K[q];
        ID[seg.begin + rel] = I[q];
    }
}

__global__ __launch_bounds__(kWG) void k_children(Item seg, uint32_t T, const uint32_t* hscan, Sink s)
{
    __shared__ uint32_t s_tmp[16];
    emit_children(seg, kMaxBkt, hscan, T * kWaves, s, s_tmp);
}

// ---- one-workgroup sample-sort step (kSmallT < size <= kMedT) -----------------

__global__ __launch_bounds__(kWG) void k_medium(const Item* items, const uint32_t* cnt_items, Bufs B,
                                                 Sink s, uint64_t seed, int tb_shift)
{
    __shared__ key64 s_tree[1024];
    __shared__ key64 s_spl[1024];
    __shared__ uint32_t s_hist[kWaves][2048];
    __shared__ uint32_t s_tmp[16];
    if (blockIdx.x >= *cnt_items) return;
    const Item seg = items[blockIdx.x];
    if (!item_ok(seg, B) || seg.size > kMedT) {
        if (threadIdx.x == 0) flag_error(B.cnt, ERR_BOUNDS);
        return;
    }
    const uint32_t m = seg.size, beg = seg.begin;
    const int src = seg.flags & F_BUFB;
    const key64* K = src ? B.key1 : B.key0;
    const uint32_t* I = src ? B.idx1 : B.idx0;
    key64* KD = src ? B.key0 : B.key1;
    uint32_t* ID = src ? B.idx0 : B.idx1;
    const int k = treebits_for(m, tb_shift);
    const int nspl = (1 << k) - 1, nbkt = 2 * nspl + 1;

    build_tree(K, beg, m, seed, k, reinterpret_cast<key64*>(&s_hist[0][0]), s_tree, s_spl);
    for (int i = threadIdx.x; i < kWaves * 2048; i += kWG) (&s_hist[0][0])[i] = 0;
    __syncthreads();

    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint32_t chunk = (((m + kWaves - 1) / kWaves) + 63u) & ~63u;
    const uint32_t cb = w * chunk, ce = min(m, cb + chunk);
    for (uint32_t g = cb; g < cb + chunk; g += 64) {
        const uint32_t p = g + lane;
        const bool act = p < ce;
        uint32_t b = kNoBucket;
        if (act) b = classify_key(K[beg + p], s_tree, s_spl, k, nspl);
        const uint64_t peers = match_bucket(b);
        const uint32_t r = (uint32_t)__popcll(peers & lanemask_lt());
        if (act) {
            const uint32_t old = s_hist[w][b];
            if (r == 0) s_hist[w][b] = old + (uint32_t)__popcll(peers);
            B.bktrank[beg + p] = (b << kRankBits) | (old + r);
        }
    }
    __syncthreads();
    // bucket offsets (exclusive over buckets) and per-wave bases
    constexpr int kPer = 2048 / kWG;   // 8 buckets per thread
    uint32_t tot[kPer];
    uint32_t mysum = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int b = threadIdx.x * kPer + j;
        uint32_t t = 0;
        if (b < nbkt) {
#pragma unroll
            for (int ww = 0; ww < kWaves; ++ww) t += s_hist[ww][b];
        }
        tot[j] = t;
        mysum += t;
    }
    uint32_t all;
    uint32_t off = block_excl_scan(mysum, &all, s_tmp);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int b = threadIdx.x * kPer + j;
        if (b < nbkt) {
            uint32_t run = off;
#pragma unroll
            for (int ww = 0; ww < kWaves; ++ww) {
                const uint32_t c = s_hist[ww][b];
                s_hist[ww][b] = run;
                run += c;
            }
        }
        off += tot[j];
    }
    __syncthreads();
    for (uint32_t g = cb; g < cb + chunk; g += 64) {
        const uint32_t p = g + lane;
        if (p < ce) {
            const uint32_t br = B.bktrank[beg + p];
            const uint32_t b = br >> kRankBits;
            const uint32_t rel = b < (uint32_t)nbkt ? s_hist[w][b] + (br & kRankMask) : 0xFFFFFFFFu;
            if (rel >= m) { flag_error(B.cnt, ERR_BOUNDS); continue; }
            KD[beg + rel] = K[beg + p];
            ID[beg + rel] = I[beg + p];
        }
    }
    if (threadIdx.x == 0) atomicAdd(&s.cnt[C_EL_MED], m);
    __syncthreads();
    emit_children(seg, nbkt, &s_hist[0][0], 1, s, s_tmp);
}

// ---- leaves --------------------------------------------------------------------

// Finish one sorted position of a leaf at depth d (shared by all leaf kernels):
// LCP with the predecessor inside the leaf when keys differ, and the fate of
// the run of equal keys the position belongs to.
__device__ __forceinline__ void finish_position(const Bufs& B, uint32_t pos, uint32_t id, key64 k,
                                                bool has_prev, key64 pk, bool run_single,
                                                uint32_t depth)
{
    const bool eqp = has_prev && pk == k;
    if (has_prev && !eqp && B.lcp) B.lcp[pos] = depth + lcp_key(pk, k);
    const bool term = (k & 0xFFull) == 0;
    if (run_single || term) {
        B.perm[pos] = id;
        if (eqp && B.lcp) B.lcp[pos] = depth + key_depth(k);
    } else {
        B.idx0[pos] = id;
        if (id < B.n) B.key0[pos] = load_key(B.arena, B.alen, (uint64_t)B.offs[id] + depth + 8);
        else flag_error(B.cnt, ERR_BOUNDS);
    }
}

// One wave per leaf of at most 64 strings: rank sort with broadcast compares.
__global__ __launch_bounds__(kWG) void k_leaf_tiny(const Item* items, const uint32_t* cnt_items, Bufs B,
                                                   Sink nxt, uint32_t depth)
{
    __shared__ key64 s_k[kWaves][64];
    __shared__ uint32_t s_i[kWaves][64];
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint32_t it = blockIdx.x * kWaves + w;
    const bool valid = it < *cnt_items;
    Item L = { 0u, 0u, 0u, 0u };
    if (valid) L = items[it];
    L.begin = (uint32_t)__builtin_amdgcn_readfirstlane((int)L.begin);
    L.size = (uint32_t)__builtin_amdgcn_readfirstlane((int)L.size);
    L.flags = (uint32_t)__builtin_amdgcn_readfirstlane((int)L.flags);
    if (valid && (!item_ok(L, B) || L.size > kTinyT)) {
        if (lane == 0) flag_error(B.cnt, ERR_BOUNDS);
        L.size = 0;
    }
    const uint32_t m = L.size;
    const bool act = lane < m;
    const int src = L.flags & F_BUFB;
    key64 key = ~0ull;
    uint32_t id = 0;
    if (act) {
        key = src ? B.key1[L.begin + lane] : B.key0[L.begin + lane];
        id = src ? B.idx1[L.begin + lane] : B.idx0[L.begin + lane];
    }
    uint32_t rank = 0;
    for (uint32_t i = 0; i < m; ++i) {
        const key64 ki = readlane64(key, i);
        rank += (ki < key || (ki == key && i < lane)) ? 1u : 0u;
    }
    if (act) {
        s_k[w][rank] = key;
        s_i[w][rank] = id;
    }
    __syncthreads();
    key64 k = 0, pk = 0;
    if (act) {
        k = s_k[w][lane];
        id = s_i[w][lane];
        if (lane > 0) pk = s_k[w][lane - 1];
    }
    const bool eqp = act && lane > 0 && pk == k;
    const uint64_t starts = __ballot(act && !eqp);
    const uint64_t le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const uint64_t before = starts & le;
    const uint32_t rb = before ? 63u - (uint32_t)__clzll((long long)before) : 0u;
    const uint64_t after = starts & ~le;
    const uint32_t re = after ? (uint32_t)__ffsll((long long)after) - 1u : m;
    const uint32_t runlen = re - rb;
    if (act) {
        const uint32_t pos = L.begin + lane;
        finish_position(B, pos, id, k, lane > 0, pk, runlen == 1, depth);
        if (lane == 0 || lane == m - 1) B.bkey[pos] = k;
    }
    const bool nr = act && lane == rb && runlen >= 2 && (k & 0xFFull) != 0;
    Item ni = { L.begin + rb, runlen, L.begin + rb, 0u };
    wave_append_segment(nr, ni, nxt);
    if (lane == 0 && valid) atomicAdd(&nxt.cnt[C_EL_TINY], m);
}

// One workgroup per leaf of 65..1024 strings: bitonic sort of (key, position).
__global__ __launch_bounds__(kWG) void k_leaf_small(const Item* items, const uint32_t* cnt_items, Bufs B,
                                                    Sink nxt, uint32_t depth)
{
    __shared__ key64 s_k[kSmallT];
    __shared__ uint32_t s_p[kSmallT];
    __shared__ uint32_t s_rb[kSmallT];
    __shared__ uint32_t s_len[kSmallT];
    if (blockIdx.x >= *cnt_items) return;
    const Item L = items[blockIdx.x];
    if (!item_ok(L, B) || L.size > kSmallT) {
        if (threadIdx.x == 0) flag_error(B.cnt, ERR_BOUNDS);
        return;
    }
    const uint32_t m = L.size;
    const int src = L.flags & F_BUFB;
    const key64* K = src ? B.key1 : B.key0;
    const uint32_t* I = src ? B.idx1 : B.idx0;
    uint32_t P = 2;
    while (P < m) P <<= 1;
    for (uint32_t j = threadIdx.x; j < P; j += kWG) {
        s_k[j] = j < m ? K[L.begin + j] : ~0ull;
        s_p[j] = j;
    }
    __syncthreads();
    for (uint32_t kk = 2; kk <= P; kk <<= 1) {
        for (uint32_t jj = kk >> 1; jj > 0; jj >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += kWG) {
                const uint32_t ixj = i ^ jj;
                if (ixj > i) {
                    const key64 a = s_k[i], b = s_k[ixj];
                    const uint32_t pa = s_p[i], pb = s_p[ixj];
                    const bool gt = a > b || (a == b && pa > pb);
                    const bool up = (i & kk) == 0;
                    if (gt == up) {
                        s_k[i] = b; s_k[ixj] = a;
                        s_p[i] = pb; s_p[ixj] = pa;
                    }
                }
            }
            __syncthreads();
        }
    }
    // input indices in sorted order (read before any position is rewritten)
    for (uint32_t j = threadIdx.x; j < m; j += kWG) {
        s_len[j] = I[L.begin + s_p[j]];
        s_rb[j] = (j == 0 || s_k[j - 1] != s_k[j]) ? j : 0u;
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < m; j += kWG) s_p[j] = s_len[j];   // s_p := ids
    __syncthreads();
    // run begins: inclusive max scan
    for (uint32_t d = 1; d < m; d <<= 1) {
        uint32_t v[kSmallT / kWG];
        int c = 0;
        for (uint32_t j = threadIdx.x; j < m; j += kWG, ++c) v[c] = (j >= d) ? max(s_rb[j], s_rb[j - d]) : s_rb[j];
        __syncthreads();
        c = 0;
        for (uint32_t j = threadIdx.x; j < m; j += kWG, ++c) s_rb[j] = v[c];
        __syncthreads();
    }
    for (uint32_t j = threadIdx.x; j < m; j += kWG)
        if (j == m - 1 || s_k[j + 1] != s_k[j]) s_len[s_rb[j]] = j + 1 - s_rb[j];
    __syncthreads();
    for (uint32_t j0 = 0; j0 < m; j0 += kWG) {   // uniform loop for the wave appends
        const uint32_t j = j0 + threadIdx.x;
        const bool act = j < m;
        bool nr = false;
        Item ni = { 0u, 0u, 0u, 0u };
        if (act) {
            const key64 k = s_k[j];
            const uint32_t rb = s_rb[j], runlen = s_len[rb];
            const uint32_t pos = L.begin + j;
            finish_position(B, pos, s_p[j], k, j > 0, j > 0 ? s_k[j - 1] : 0ull, runlen == 1, depth);
            if (j == 0 || j == m - 1) B.bkey[pos] = k;
            nr = j == rb && runlen >= 2 && (k & 0xFFull) != 0;
            ni.begin = L.begin + rb;
            ni.size = runlen;
            ni.rbeg = L.begin + rb;
        }
        wave_append_segment(nr, ni, nxt);
    }
    if (threadIdx.x == 0) atomicAdd(&nxt.cnt[C_EL_SMALL], m);
}

// One workgroup per chunk of an equal-key leaf (all keys equal).
__global__ __launch_bounds__(kWG) void k_leaf_run(const Item* items, const uint32_t* cnt_items, Bufs B,
                                                  Sink nxt, uint32_t depth)
{
    if (blockIdx.x >= *cnt_items) return;
    const Item L = items[blockIdx.x];
    const uint32_t c = L.flags >> 8;
    if (!item_ok(L, B) || (uint64_t)c * kRunChunk >= L.size) {
        if (threadIdx.x == 0) flag_error(B.cnt, ERR_BOUNDS);
        return;
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
    for (uint32_t p = cb + threadIdx.x; p < ce; p += kWG) {
        const uint32_t id = I[p];
        if (single || term) {
            B.perm[p] = id;
            if (p > L.begin && B.lcp) B.lcp[p] = depth + key_depth(key);
        } else {
            if (src) B.idx0[p] = id;
            if (id < B.n) B.key0[p] = load_key(B.arena, B.alen, (uint64_t)B.offs[id] + depth + 8);
            else flag_error(B.cnt, ERR_BOUNDS);
        }
    }
    if (threadIdx.x < 64) {
        const bool nr = threadIdx.x == 0 && c == 0 && !term && !single;
        Item ni = { L.begin, L.size, L.begin, 0u };
        wave_append_segment(nr, ni, nxt);
        if (threadIdx.x == 0) {
            if (c == 0) B.bkey[L.begin] = key;
            if (ce == rend) B.bkey[rend - 1] = key;
            atomicAdd(&nxt.cnt[C_EL_RUN], ce - cb);
        }
    }
}

// LCP between neighbouring leaves of one round segment:
// depth + lcp_key(last key of the left leaf, first key of the right leaf)
// (sample_sort_lcp, .hpp:311-396).
__global__ __launch_bounds__(kWG) void k_leafbound(const Item* t, const Item* s, const Item* r,
                                                   uint32_t nt, uint32_t ns, uint32_t nr, Bufs B,
                                                   uint32_t* cnt, uint32_t depth)
{
    const uint32_t i = blockIdx.x * kWG + threadIdx.x;
    Item L;
    if (i < nt) L = t[i];
    else if (i < nt + ns) L = s[i - nt];
    else if (i < nt + ns + nr) {
        L = r[i - nt - ns];
        if ((L.flags >> 8) != 0) return;
    } else return;
    if (L.begin <= L.rbeg || !B.lcp) return;
    if (L.begin >= B.n) { flag_error(cnt, ERR_BOUNDS); return; }
    const key64 a = B.bkey[L.begin - 1], b = B.bkey[L.begin];
    if (a == b) { atomicOr(&cnt[C_ERR], ERR_LCP); return; }
    B.lcp[L.begin] = depth + lcp_key(a, b);
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
    K_INIT, K_TREE, K_COUNT, K_SCAN, K_SCATTER, K_CHILDREN, K_MEDIUM,
    K_LEAF_TINY, K_LEAF_SMALL, K_LEAF_RUN, K_LEAFBOUND, K_MISC, K_NUM
};
const char* kKernelNames[K_NUM] = {
    "k_init", "k_tree", "k_count", "k_scan", "k_scatter", "k_children", "k_medium",
    "k_leaf_tiny", "k_leaf_small", "k_leaf_run", "k_leafbound", "k_misc"
};

}  // namespace

struct ps5_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t cap = 0;
    key64* key[2] = { nullptr, nullptr };
    uint32_t* idx[2] = { nullptr, nullptr };
    key64* bkey = nullptr;
    uint32_t* bktrank = nullptr;
    // lists: level lists [2] x {large, med}; leaf lists [2] x {tiny, small, run}
    Item* l_large[2] = { nullptr, nullptr };
    Item* l_med[2] = { nullptr, nullptr };
    Item* l_tiny[2] = { nullptr, nullptr };
    Item* l_small[2] = { nullptr, nullptr };
    Item* l_run[2] = { nullptr, nullptr };
    uint32_t cap_large = 0, cap_med = 0, cap_tiny = 0, cap_small = 0, cap_run = 0;
    uint32_t* d_cnt = nullptr;
    uint32_t* h_cnt = nullptr;       // pinned
    Item* h_large = nullptr;         // pinned
    uint32_t h_large_cap = 0;
    uint32_t* hist = nullptr;
    uint64_t hist_cap = 0;
    uint32_t* part = nullptr;
    uint64_t part_cap = 0;
    key64* tree = nullptr;
    // tuning
    int tb_shift = 4;
    uint64_t seed = 0x5EED5EEDull;
    // profiling
    bool profile = false;
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
    auto f = [](void* p) { if (p) (void)hipFree(p); };
    for (int i = 0; i < 2; ++i) {
        f(c->key[i]); f(c->idx[i]); f(c->l_large[i]); f(c->l_med[i]);
        f(c->l_tiny[i]); f(c->l_small[i]); f(c->l_run[i]);
        c->key[i] = nullptr; c->idx[i] = nullptr; c->l_large[i] = nullptr; c->l_med[i] = nullptr;
        c->l_tiny[i] = nullptr; c->l_small[i] = nullptr; c->l_run[i] = nullptr;
    }
    f(c->bkey); f(c->bktrank); f(c->hist); f(c->part);
    c->bkey = nullptr; c->bktrank = nullptr; c->hist = nullptr; c->part = nullptr;
    c->cap = 0;
    c->hist_cap = 0;
    c->part_cap = 0;
}

int ensure_capacity(ps5_ctx* c, uint64_t n)
{
    if (n <= c->cap) return PS5_OK;
    free_buffers(c);
    uint64_t cap = std::max<uint64_t>(n, 1024);
    for (int i = 0; i < 2; ++i) {
        HIPCHK(hipMalloc(&c->key[i], cap * sizeof(key64)));
        HIPCHK(hipMalloc(&c->idx[i], cap * sizeof(uint32_t)));
    }
    HIPCHK(hipMalloc(&c->bkey, cap * sizeof(key64)));
    HIPCHK(hipMalloc(&c->bktrank, cap * sizeof(uint32_t)));
    c->cap_large = (uint32_t)(cap / kMedT + 16);
    c->cap_med = (uint32_t)(cap / kSmallT + 16);
    c->cap_tiny = (uint32_t)(cap + 16);
    c->cap_small = (uint32_t)(cap / kTinyT + 16);
    c->cap_run = (uint32_t)(cap + 16);
    for (int i = 0; i < 2; ++i) {
        HIPCHK(hipMalloc(&c->l_large[i], (uint64_t)c->cap_large * sizeof(Item)));
        HIPCHK(hipMalloc(&c->l_med[i], (uint64_t)c->cap_med * sizeof(Item)));
        HIPCHK(hipMalloc(&c->l_tiny[i], (uint64_t)c->cap_tiny * sizeof(Item)));
        HIPCHK(hipMalloc(&c->l_small[i], (uint64_t)c->cap_small * sizeof(Item)));
        HIPCHK(hipMalloc(&c->l_run[i], (uint64_t)c->cap_run * sizeof(Item)));
    }
    c->hist_cap = (uint64_t)kMaxBkt * kWaves * (cap / kTile + 2);
    HIPCHK(hipMalloc(&c->hist, c->hist_cap * sizeof(uint32_t)));
    c->part_cap = c->hist_cap / kScanBlock + 2;
    HIPCHK(hipMalloc(&c->part, c->part_cap * sizeof(uint32_t)));
    if (c->h_large_cap < c->cap_large) {
        if (c->h_large) (void)hipHostFree(c->h_large);
        c->h_large = nullptr;
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->h_large), (uint64_t)c->cap_large * sizeof(Item)));
        c->h_large_cap = c->cap_large;
    }
    c->cap = cap;
    return PS5_OK;
}

struct Launch {   // brackets a launch with events when profiling
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

Sink make_sink(ps5_ctx* c, int lvl, int leafpar)
{
    Sink s;
    s.large = c->l_large[lvl];
    s.med = c->l_med[lvl];
    s.tiny = c->l_tiny[leafpar];
    s.small = c->l_small[leafpar];
    s.run = c->l_run[leafpar];
    s.cnt = c->d_cnt;
    s.c_large = C_LARGE + lvl;
    s.c_med = C_MED + lvl;
    s.c_tiny = C_TINY + leafpar;
    s.c_small = C_SMALL + leafpar;
    s.c_run = C_RUN + leafpar;
    s.cap_large = c->cap_large;
    s.cap_med = c->cap_med;
    s.cap_tiny = c->cap_tiny;
    s.cap_small = c->cap_small;
    s.cap_run = c->cap_run;
    return s;
}

// copy counters (and up to the first kPeek large items of list `lvl`) to the host
int readback(ps5_ctx* c, int lvl, bool want_large)
{
    HIPCHK(hipMemcpyAsync(c->h_cnt, c->d_cnt, C_NCNT * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->st_syncs++;
    if (c->h_cnt[C_ERR] & ERR_CAP) return set_err(PS5_ERR_CAPACITY, "work list capacity exceeded");
    if (c->h_cnt[C_ERR] & ERR_LCP) return set_err(PS5_ERR_INTERNAL, "equal keys across a leaf boundary");
    if (c->h_cnt[C_ERR] & ERR_BOUNDS) return set_err(PS5_ERR_INTERNAL, "device bounds check failed");
    if (want_large) {
        const uint32_t nl = c->h_cnt[C_LARGE + lvl];
        if (nl > c->cap_large) return set_err(PS5_ERR_CAPACITY, "large list overflow");
        if (nl) {
            HIPCHK(hipMemcpyAsync(c->h_large, c->l_large[lvl], nl * sizeof(Item), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            c->st_syncs++;
        }
    }
    return PS5_OK;
}

uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

int run_sort(ps5_ctx* c, const uint8_t* arena, uint64_t alen, const uint32_t* offs, uint32_t n,
             uint32_t* perm, uint32_t* lcp)
{
    int rc = ensure_capacity(c, n);
    if (rc) return rc;
    for (int i = 0; i < K_NUM; ++i) { c->k_ms[i] = 0; c->k_launch[i] = 0; c->k_bytes[i] = 0; }
    c->st_rounds = c->st_levels = c->st_segments = c->st_leaves = c->st_syncs = 0;

    Bufs B;
    B.n = n;
    B.cnt = c->d_cnt;
    B.arena = arena;
    B.alen = alen;
    B.offs = offs;
    B.key0 = c->key[0]; B.key1 = c->key[1];
    B.idx0 = c->idx[0]; B.idx1 = c->idx[1];
    B.perm = perm;
    B.lcp = lcp;
    B.bkey = c->bkey;
    B.bktrank = c->bktrank;
    hipStream_t st = c->stream;

    {
        Launch L(c, K_INIT, (uint64_t)n * 24);
        hipLaunchKernelGGL(k_init, dim3(div_up(n, kWG)), dim3(kWG), 0, st, B, n);
    }
    {
        Item* list; int ci;
        if (n <= kTinyT) { list = c->l_tiny[0]; ci = C_TINY; }
        else if (n <= kSmallT) { list = c->l_small[0]; ci = C_SMALL; }
        else if (n <= kMedT) { list = c->l_med[0]; ci = C_MED; }
        else { list = c->l_large[0]; ci = C_LARGE; }
        Launch L(c, K_MISC, 0);
        hipLaunchKernelGGL(k_setup, dim3(1), dim3(64), 0, st, c->d_cnt, list, ci, n);
    }
    HIPCHK(hipGetLastError());
    if ((rc = readback(c, 0, true))) return rc;

    for (uint32_t round = 0;; ++round) {
        const uint32_t depth = round * 8;
        const int lp = round & 1;           // leaf list parity of this round
        c->st_rounds++;
        int lv = 0;
        for (uint32_t level = 0;; ++level) {   // levels
            if (level > 4096) return set_err(PS5_ERR_INTERNAL, "sample-sort levels do not terminate");
            const uint32_t nlarge = c->h_cnt[C_LARGE + lv];
            const uint32_t nmed = c->h_cnt[C_MED + lv];
            if (nlarge == 0 && nmed == 0) break;
            c->st_levels++;
            c->st_segments += nlarge + nmed;
            const int nx = lv ^ 1;
            {
                Launch L(c, K_MISC, 0);
                hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, c->d_cnt, C_LARGE + nx, C_MED + nx, -1, -1, -1);
            }
            Sink sink = make_sink(c, nx, lp);
            std::vector<Item> large(c->h_large, c->h_large + nlarge);
            for (const Item& seg : large) {
                const uint32_t T = div_up(seg.size, kTile);
                const uint64_t hl64 = (uint64_t)kMaxBkt * T * kWaves;
                if (hl64 > 0xFFFFFFFFull) return set_err(PS5_ERR_ARG, "segment too large");
                const uint32_t hl = (uint32_t)hl64;
                if (hl > c->hist_cap) return set_err(PS5_ERR_CAPACITY, "histogram capacity");
                const uint64_t seed = c->seed ^ ((uint64_t)round << 40);
                {
                    Launch L(c, K_TREE, 2046ull * 8);
                    hipLaunchKernelGGL(k_tree, dim3(1), dim3(kWG), 0, st, seg, B, seed, c->tree);
                }
                {
                    Launch L(c, K_COUNT, (uint64_t)seg.size * 12 + (uint64_t)hl * 4);
                    hipLaunchKernelGGL(k_count, dim3(T), dim3(kWG), 0, st, seg, B, c->tree, T, c->hist);
                }
                {
                    const uint32_t nb = div_up(hl, kScanBlock);
                    Launch L(c, K_SCAN, (uint64_t)hl * 12);
                    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kWG), 0, st, c->hist, hl, c->part);
                    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kWG), 0, st, c->part, nb);
                    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kWG), 0, st, c->hist, hl, c->part);
                }
                {
                    Launch L(c, K_SCATTER, (uint64_t)seg.size * 28 + (uint64_t)hl * 4);
                    hipLaunchKernelGGL(k_scatter, dim3(T), dim3(kWG), 0, st, seg, B, T, c->hist);
                }
                {
                    Launch L(c, K_CHILDREN, (uint64_t)kMaxBkt * 4);
                    hipLaunchKernelGGL(k_children, dim3(1), dim3(kWG), 0, st, seg, T, c->hist, sink);
                }
            }
            if (nmed) {
                Launch L(c, K_MEDIUM, 0);
                hipLaunchKernelGGL(k_medium, dim3(nmed), dim3(kWG), 0, st, c->l_med[lv], c->d_cnt + C_MED + lv,
                                   B, sink, c->seed ^ ((uint64_t)round << 40) ^ ((uint64_t)c->st_levels << 32),
                                   c->tb_shift);
            }
            HIPCHK(hipGetLastError());
            if ((rc = readback(c, nx, true))) return rc;
            lv = nx;
        }
        // leaves of this round
        const uint32_t nt = c->h_cnt[C_TINY + lp], ns = c->h_cnt[C_SMALL + lp], nr = c->h_cnt[C_RUN + lp];
        c->st_leaves += nt + ns + nr;
        const int np = lp ^ 1;
        {
            Launch L(c, K_MISC, 0);
            hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, c->d_cnt, C_LARGE + 0, C_MED + 0,
                               C_TINY + np, C_SMALL + np, C_RUN + np);
            hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, c->d_cnt, C_LARGE + 1, C_MED + 1, -1, -1, -1);
        }
        Sink nsink = make_sink(c, 0, np);
        if (nt) {
            Launch L(c, K_LEAF_TINY, 0);
            hipLaunchKernelGGL(k_leaf_tiny, dim3(div_up(nt, kWaves)), dim3(kWG), 0, st, c->l_tiny[lp],
                               c->d_cnt + C_TINY + lp, B, nsink, depth);
        }
        if (ns) {
            Launch L(c, K_LEAF_SMALL, 0);
            hipLaunchKernelGGL(k_leaf_small, dim3(ns), dim3(kWG), 0, st, c->l_small[lp],
                               c->d_cnt + C_SMALL + lp, B, nsink, depth);
        }
        if (nr) {
            Launch L(c, K_LEAF_RUN, 0);
            hipLaunchKernelGGL(k_leaf_run, dim3(nr), dim3(kWG), 0, st, c->l_run[lp],
                               c->d_cnt + C_RUN + lp, B, nsink, depth);
        }
        if (nt + ns + nr && lcp) {
            Launch L(c, K_LEAFBOUND, (uint64_t)(nt + ns + nr) * 36);
            hipLaunchKernelGGL(k_leafbound, dim3(div_up((uint64_t)nt + ns + nr, kWG)), dim3(kWG), 0, st,
                               c->l_tiny[lp], c->l_small[lp], c->l_run[lp], nt, ns, nr, B, c->d_cnt, depth);
        }
        HIPCHK(hipGetLastError());
        if ((rc = readback(c, 0, true))) return rc;
        if (c->h_cnt[C_LARGE] + c->h_cnt[C_MED] + c->h_cnt[C_TINY + np] + c->h_cnt[C_SMALL + np] +
                c->h_cnt[C_RUN + np] == 0)
            break;
        if (depth > alen + 8) return set_err(PS5_ERR_INTERNAL, "depth exceeds the arena: no progress");
    }
    // element-proportional byte accounting of the list-driven kernels
    const uint64_t el_med = c->h_cnt[C_EL_MED], el_tiny = c->h_cnt[C_EL_TINY];
    const uint64_t el_small = c->h_cnt[C_EL_SMALL], el_run = c->h_cnt[C_EL_RUN];
    c->k_bytes[K_MEDIUM] += el_med * 40;
    c->k_bytes[K_LEAF_TINY] += el_tiny * 20;
    c->k_bytes[K_LEAF_SMALL] += el_small * 20;
    c->k_bytes[K_LEAF_RUN] += el_run * 12;
    if (c->profile) collect_events(c);
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
    ps5_ctx* c = new ps5_ctx;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&c->d_cnt, C_NCNT * sizeof(uint32_t));
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&c->h_cnt), C_NCNT * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->tree, 2048 * sizeof(key64));
    if (e != hipSuccess) {
        ps5_ctx_destroy(c);
        return set_err(PS5_ERR_HIP, "context setup failed: %s", hipGetErrorString(e));
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
    if (c->tree) (void)hipFree(c->tree);
    if (c->h_cnt) (void)hipHostFree(c->h_cnt);
    if (c->h_large) (void)hipHostFree(c->h_large);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    for (auto& u : c->ev_used) { (void)hipEventDestroy(u.second.first); (void)hipEventDestroy(u.second.second); }
    if (c->stream) (void)hipStreamDestroy(c->stream);
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
    int rc = run_sort(c, d_arena, arena_bytes, d_offs, (uint32_t)n, d_perm, d_lcp);
    if (rc) {
        (void)hipStreamSynchronize(c->stream);
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
    if (e != hipSuccess) return set_err(PS5_ERR_NOMEM, "hipMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
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

int ps5_device_synchronize(void)
{
    HIPCHK(hipDeviceSynchronize());
    return PS5_OK;
}

}  // extern "C"
