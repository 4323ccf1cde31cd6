// ps5_device.hpp -- device helpers of the gfx950 pS5 kernels (wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ps5_kernels.hpp"

namespace ps5 {

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
// are fetched one by one so nothing outside arena[0, alen) is read.
// The arena is global memory: loads through address-space-1 pointers are
// global loads, which wait on the vector memory counter alone (a generic
// pointer would compile to flat loads that also wait for LDS operations).
typedef __attribute__((address_space(1))) const uint64_t g_u64;
typedef __attribute__((address_space(1))) const uint8_t g_u8;
__device__ __forceinline__ key64 load_key(const uint8_t* arena, uint64_t alen, uint64_t off)
{
    const uint64_t a8 = off & ~7ull;
    uint64_t lo, hi;
    if (a8 + 16 <= alen) {
        const g_u64* p = (const g_u64*)(arena + a8);
        lo = p[0];
        hi = p[1];
    } else {
        lo = 0;
        hi = 0;
        const g_u8* b = (const g_u8*)arena;
        for (int i = 0; i < 16; ++i) {
            const uint64_t q = a8 + i;
            if (q < alen) {
                const uint64_t c = b[q];
                if (i < 8) lo |= c << (8 * i);
                else hi |= c << (8 * (i - 8));
            }
        }
    }
    const uint32_t sh = (uint32_t)(off & 7) * 8;
    uint64_t v = sh ? ((lo >> sh) | (hi << (64 - sh))) : lo;
    v = __builtin_bswap64(v);
    // high bit of each byte of t is set iff that byte is non-zero (no carries)
    const uint64_t t = ((v & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full) | v;
    const uint64_t z = ~t & 0x8080808080808080ull;
    if (z) {
        const uint32_t p = (uint32_t)__clzll((long long)z);   // 8 * index of the first NUL
        v = p ? (v & (~0ull << (64 - p))) : 0ull;
    }
    return v;
}

// The keys at depths off, off+8, ..., off+8(KB-1) of one string: KB+1
// aligned 8-byte loads issued together (load_key near the arena end).  Keys
// past the string's NUL hold whatever follows it; callers use a key only
// while every key before it was NUL-free.
template <int KB>
__device__ __forceinline__ void load_keys(const uint8_t* arena, uint64_t alen, uint64_t off, key64 (&k)[KB],
                                          int nk = KB)
{
    const uint64_t a8 = off & ~7ull;
    if (a8 + 8 * (uint64_t)(nk + 1) > alen) {
#pragma unroll
        for (int i = 0; i < KB; ++i) k[i] = i < nk ? load_key(arena, alen, off + 8 * i) : 0ull;
        return;
    }
    const g_u64* p = (const g_u64*)(arena + a8);
    uint64_t w[KB + 1];
#pragma unroll
    for (int i = 0; i <= KB; ++i) w[i] = i <= nk ? p[i] : 0ull;
    const uint32_t sh = (uint32_t)(off & 7) * 8;
#pragma unroll
    for (int i = 0; i < KB; ++i) {
        uint64_t v = sh ? ((w[i] >> sh) | (w[i + 1] << (64 - sh))) : w[i];
        v = __builtin_bswap64(v);
        const uint64_t t = ((v & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full) | v;
        const uint64_t z = ~t & 0x8080808080808080ull;
        if (z) {
            const uint32_t q = (uint32_t)__clzll((long long)z);
            v = q ? (v & (~0ull << (64 - q))) : 0ull;
        }
        k[i] = v;
    }
}

// characters the two keys share (lcpKeyType, bingmann-parallel_sample_sort.hpp:225-230)
__device__ __forceinline__ uint32_t lcp_key(key64 a, key64 b)
{
    key64 x = a ^ b;
    return x ? ((uint32_t)__clzll((long long)x) >> 3) : 8u;
}
// non-NUL characters of a key (lcpKeyDepth, bingmann-parallel_sample_sort.hpp:232-237)
__device__ __forceinline__ uint32_t key_depth(key64 a)
{
    return a ? 8u - ((uint32_t)__builtin_ctzll(a) >> 3) : 0u;
}

// tid_x() through an opaque move: values derived from it are recomputed
// where they are used, instead of being hoisted out of the persistent loops
// of the kernels and holding registers across every iteration
__device__ __forceinline__ uint32_t tid_x()
{
    uint32_t t;
    asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"((uint32_t)threadIdx.x));
    return t;
}
__device__ __forceinline__ uint32_t lane_id() { return tid_x() & 63u; }
__device__ __forceinline__ uint32_t wave_id() { return (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid_x() >> 6)); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// (key, position) order: positions are distinct, so the order is total and
// equal keys keep their input order (stable).
__device__ __forceinline__ bool kp_less(key64 ka, uint32_t pa, key64 kb, uint32_t pb)
{
    return ka < kb || (ka == kb && pa < pb);
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup i runs on
// XCD i % 8), and every XCD has its own L2.  This bijection of [0, G) gives
// each XCD one contiguous range of work items, so neighbouring tiles -- whose
// histogram entries and bucket output ranges share cache lines -- meet in the
// same L2.
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t bid, uint32_t G)
{
    const uint32_t x = bid % kXcds, y = bid / kXcds, q = G / kXcds, r = G % kXcds;
    return x * q + min(x, r) + y;
}

// lanes holding the same 11-bit bucket id (all 64 lanes must execute this)
template <int NB = 11>
__device__ __forceinline__ uint64_t match_bucket(uint32_t b)
{
    const uint64_t act = __ballot(1);
    uint32_t lo = (uint32_t)act, hi = (uint32_t)(act >> 32);
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        // m: all ones if bit i of b is set, else zero; lanes agreeing with
        // this lane on bit i are ~(ballot ^ m), computed as (~ballot) ^ m
        const uint32_t m = (uint32_t)((int32_t)(b << (31 - i)) >> 31);
        const uint64_t nbal = ~(uint64_t)__ballot(m != 0u);
        lo &= (uint32_t)nbal ^ m;
        hi &= (uint32_t)(nbal >> 32) ^ m;
    }
    return ((uint64_t)hi << 32) | lo;
}

// The same lanes for NB-bit ids with fewer vector instructions (all 64 lanes
// must execute this): every lane ORs its bit (lanebit) into this wave's table
// of 64-bit lane masks at the id's low TB bits, reads the entry back -- the
// lanes agreeing on those bits -- and clears it, and ballots on the NB - TB
// high bits keep the lanes agreeing on the whole id.  tab holds 1 << TB
// entries, zero on entry and zero again on return.  The OR of the lanes' bits
// does not depend on the order the LDS applies them in, and a wave's LDS
// operations complete in issue order, so every read sees all of the ORs.
template <int NB, int TB>
__device__ __forceinline__ uint64_t match_bucket_lds(uint32_t b, uint64_t* tab, uint64_t lanebit)
{
    static_assert(TB >= 1 && TB <= NB, "table bits within the id");
    uint64_t* e = tab + (b & ((1u << TB) - 1));
    __hip_atomic_fetch_or(e, lanebit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    const uint64_t got = __hip_atomic_load(e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    __hip_atomic_store(e, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    uint32_t lo = (uint32_t)got, hi = (uint32_t)(got >> 32);
#pragma unroll
    for (int i = TB; i < NB; ++i) {
        const uint32_t m = (uint32_t)((int32_t)(b << (31 - i)) >> 31);
        const uint64_t nbal = ~(uint64_t)__ballot(m != 0u);
        lo &= (uint32_t)nbal ^ m;
        hi &= (uint32_t)(nbal >> 32) ^ m;
    }
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ void flag_error(uint32_t* cnt, uint32_t e) { atomicOr(&cnt[C_ERR], e); }

// byte offset of input string i in the arena
__device__ __forceinline__ uint64_t str_off(const Bufs& B, uint32_t i)
{
    return B.offs64 ? B.offs64[i] : (uint64_t)B.offs[i];
}

// an item must describe a non-empty range inside [0, n)
__device__ __forceinline__ bool item_ok(const Item& it, const Bufs& B)
{
    return it.size != 0 && it.begin < B.n && it.size <= B.n - it.begin;
}

__device__ __forceinline__ void put_item(Item* list, uint32_t cap, uint32_t slot, Item it, uint32_t* cnt)
{
    if (slot < cap) list[slot] = it;
    else flag_error(cnt, ERR_CAP);
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

// Append an unsorted segment (any size) to the list of its size class.
__device__ __forceinline__ void wave_append_segment(bool pred, Item it, const Sink& s)
{
    const uint32_t m = it.size;
    const bool t = pred && m <= kTinyT;
    const bool bl = pred && m > kTinyT && m <= kBlkT;
    const bool lg = pred && m > kBlkT;
    uint32_t slot;
    slot = wave_append(t, &s.cnt[s.c_tiny]);
    if (t) put_item(s.tiny, s.cap_tiny, slot, it, s.cnt);
#pragma unroll
    for (int c = 0; c < kBlkClasses; ++c) {
        const bool bc = bl && block_class(m) == c;
        slot = wave_append(bc, &s.cnt[s.c_block + 2 * c]);
        if (bc) put_item(s.block + (uint64_t)c * s.cap_block, s.cap_block, slot, it, s.cnt);
    }
    slot = wave_append(lg, &s.cnt[s.c_large]);
    if (lg) put_item(s.large, s.cap_large, slot, it, s.cnt);
}

// Exclusive block scan of one uint32 per thread (NW waves); *total = sum.
template <int NW>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total, uint32_t* s_tmp)
{
    const uint32_t lane = lane_id(), w = wave_id();
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
    for (int i = 0; i < NW; ++i) {
        uint32_t t = s_tmp[i];
        wpre += (i < (int)w) ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return wpre + x - v;
}

// Exclusive block scans of K uint32 per thread at once (NW waves): ex[k] is
// the sum of v[k] over the threads before this one, tot[k] the block total.
// Two barriers for all K (s_tmp: NW * K entries).
template <int NW, int K>
__device__ __forceinline__ void block_excl_scan_k(const uint32_t (&v)[K], uint32_t (&ex)[K], uint32_t (&tot)[K],
                                                  uint32_t* s_tmp)
{
    const uint32_t lane = lane_id(), w = wave_id();
    uint32_t x[K];
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] = v[k];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t y = (uint32_t)__shfl_up((int)x[k], d, 64);
            if (lane >= (uint32_t)d) x[k] += y;
        }
    }
    if (lane == 63) {
#pragma unroll
        for (int k = 0; k < K; ++k) s_tmp[w * K + k] = x[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        uint32_t wpre = 0, t = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            const uint32_t u = s_tmp[i * K + k];
            wpre += (i < (int)w) ? u : 0u;
            t += u;
        }
        ex[k] = wpre + x[k] - v[k];
        tot[k] = t;
    }
    __syncthreads();
}

// In-order index of level-order node t of a complete tree with k levels.
__device__ __forceinline__ int inorder_of(int t, int k)
{
    const int l = 31 - __clz(t);
    const int j = t - (1 << l);
    return ((2 * j + 1) << (k - l - 1)) - 1;
}

// Buckets of G keys: b = 2 * lower_bound + [splitter == key]
// (ClassifyTreeSimple::find_bkt, bingmann-sample_sortBTC.hpp:45-72).  The
// descent i = 2i + (key > tree[i]) over the level-order tree ends at the
// number of splitters below the key; the equality test reads that splitter
// from the sorted array (one load instead of tracking a candidate per level).
// The G descents advance level by level together, so their LDS reads are in
// flight at the same time.
template <int G>
__device__ __forceinline__ void classify_keys(const key64 (&key)[G], const key64* s_tree, const key64* s_sorted,
                                              int k, uint32_t (&b)[G])
{
    // the descent tracks byte offsets into s_tree (8 * node): each level is
    // one LDS read, one compare and one shift-add
    uint32_t a[G];
#pragma unroll
    for (int j = 0; j < G; ++j) a[j] = 8;
    const unsigned char* tb = reinterpret_cast<const unsigned char*>(s_tree);
    for (int l = 0; l < k; ++l) {
#pragma unroll
        for (int j = 0; j < G; ++j)
            a[j] = 2 * a[j] + (key[j] > *reinterpret_cast<const key64*>(tb + a[j]) ? 8u : 0u);
    }
    const uint32_t nspl = (1u << k) - 1;
#pragma unroll
    for (int j = 0; j < G; ++j) {
        const uint32_t x = (a[j] >> 3) - (1u << k);
        b[j] = 2 * x + ((x < nspl && s_sorted[x] == key[j]) ? 1u : 0u);
    }
}

}  // namespace ps5
