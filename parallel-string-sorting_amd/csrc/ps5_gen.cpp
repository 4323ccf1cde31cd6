// ps5_gen.cpp -- seeded synthetic string inputs (host memory).
//
// The reference's artificial inputs (src/tools/input.hpp:425-468) and the
// workloads this project measures.  PRNG: xoshiro256** seeded through
// splitmix64 (Blackman & Vigna, public domain algorithms).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/ps5_mi355x.h"

namespace {

struct SplitMix {
    uint64_t x;
    uint64_t next()
    {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
};

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed)
    {
        SplitMix sm{ seed };
        for (auto& v : s) v = sm.next();
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    uint64_t uniform(uint64_t lo, uint64_t hi) { return lo + (next() >> 11) % (hi - lo + 1); }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// LCGRandom, src/tools/lcgrandom.hpp:27-40
struct Lcg {
    uint64_t xn;
    uint64_t operator()()
    {
        xn = 0x27BB2EE687B0B0FDull * xn + 0xB504F32Dull;
        return xn;
    }
};

}  // namespace

extern "C" uint64_t ps5_gen_uniform(uint64_t seed, uint64_t arena_cap, uint32_t len_lo, uint32_t len_hi,
                                    const uint8_t* alphabet, uint32_t alpha_n, uint8_t* arena, uint32_t* offs,
                                    uint64_t offs_cap, uint64_t* arena_used)
{
    Xoshiro rng(seed);
    uint64_t pos = 0, n = 0;
    if (alpha_n == 0 || len_hi < len_lo) {
        if (arena_used) *arena_used = 0;
        return 0;
    }
    for (;;) {
        const uint64_t L = rng.uniform(len_lo, len_hi);
        if (pos + L + 1 > arena_cap || n >= offs_cap) break;
        offs[n++] = (uint32_t)pos;
        for (uint64_t j = 0; j < L; ++j) arena[pos + j] = alphabet[rng.uniform(0, alpha_n - 1)];
        arena[pos + L] = 0;
        pos += L + 1;
    }
    if (pos < arena_cap) memset(arena + pos, 0, arena_cap - pos);
    if (arena_used) *arena_used = pos;
    return n;
}

extern "C" int ps5_gen_windows(uint64_t seed, uint64_t text_len, uint32_t nseg, uint32_t seg_len,
                               double repeat_frac, uint64_t n, uint32_t win_len, uint8_t* arena, uint32_t* offs)
{
    if (text_len < (uint64_t)win_len || text_len < (uint64_t)seg_len || nseg == 0) return PS5_ERR_ARG;
    static const uint8_t acgt[4] = { 'A', 'C', 'G', 'T' };
    Xoshiro rng(seed);
    std::vector<uint8_t> text(text_len);
    for (uint64_t i = 0; i < text_len; ++i) text[i] = acgt[rng.next() >> 62];
    std::vector<uint8_t> segs((uint64_t)nseg * seg_len);
    for (uint32_t s = 0; s < nseg; ++s) {
        const uint64_t o = rng.uniform(0, text_len - seg_len);
        memcpy(&segs[(uint64_t)s * seg_len], &text[o], seg_len);
    }
    const uint64_t target = (uint64_t)std::ceil(repeat_frac * (double)text_len);
    uint64_t written = 0;
    while (written < target) {
        const uint64_t s = rng.uniform(0, nseg - 1);
        const uint64_t d = rng.uniform(0, text_len - seg_len);
        memcpy(&text[d], &segs[s * seg_len], seg_len);
        written += seg_len;
    }
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t o = rng.uniform(0, text_len - win_len);
        const uint64_t p = i * (uint64_t)(win_len + 1);
        offs[i] = (uint32_t)p;
        memcpy(arena + p, &text[o], win_len);
        arena[p + win_len] = 0;
    }
    return PS5_OK;
}

// generate_random, src/tools/input.hpp:425-468
extern "C" uint64_t ps5_gen_reference_random(uint64_t size, const uint8_t* letters, uint32_t k, uint8_t* arena,
                                             uint32_t* offs, uint64_t offs_cap)
{
    if (size == 0 || k == 0) return 0;
    Lcg rng{ 1234567 };
    uint64_t slen = 0, count = 0;
    for (uint64_t i = 0; i < size; ++i) {
        if (i == slen) {
            if (count < offs_cap) offs[count] = (uint32_t)i;
            count++;
            slen += (rng() % 3) + 16;
        }
        if (i + 1 == slen) arena[i] = 0;
        else arena[i] = letters[(rng() / 100) % k];
    }
    arena[size - 1] = 0;
    for (uint64_t i = size; i < size + 9; ++i) arena[i] = 0;
    return count;
}

extern "C" uint64_t ps5_gen_prefixed(uint64_t seed, uint64_t arena_cap, uint32_t nprefix, double zipf_s,
                                     uint8_t* arena, uint32_t* offs, uint64_t offs_cap, uint64_t* arena_used)
{
    Xoshiro rng(seed);
    if (nprefix == 0) {
        if (arena_used) *arena_used = 0;
        return 0;
    }
    std::vector<std::vector<uint8_t>> pre(nprefix);
    for (auto& p : pre) {
        p.resize(rng.uniform(8, 40));
        for (auto& c : p) c = (uint8_t)rng.uniform(33, 126);
    }
    std::vector<double> cdf(nprefix);
    double acc = 0;
    for (uint32_t r = 0; r < nprefix; ++r) {
        acc += 1.0 / std::pow((double)(r + 1), zipf_s);
        cdf[r] = acc;
    }
    uint64_t pos = 0, n = 0;
    for (;;) {
        const double u = rng.unit() * acc;
        uint32_t lo = 0, hi = nprefix - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (cdf[mid] < u) lo = mid + 1;
            else hi = mid;
        }
        const auto& p = pre[lo];
        const uint64_t L = rng.uniform(10, 60);
        if (pos + p.size() + L + 1 > arena_cap || n >= offs_cap) break;
        offs[n++] = (uint32_t)pos;
        memcpy(arena + pos, p.data(), p.size());
        pos += p.size();
        for (uint64_t j = 0; j < L; ++j) arena[pos + j] = (uint8_t)rng.uniform(33, 126);
        arena[pos + L] = 0;
        pos += L + 1;
    }
    if (pos < arena_cap) memset(arena + pos, 0, arena_cap - pos);
    if (arena_used) *arena_used = pos;
    return n;
}

// ---- batch workload -------------------------------------------------------

namespace {

struct BatchSet {
    uint64_t count;
    int alpha;          // 0: ACGT, 1: digits, 2: bytes 33..126
    uint64_t seed;
};

BatchSet batch_set(uint64_t seed, uint32_t k, uint64_t count_lo, uint64_t count_hi)
{
    SplitMix sm{ seed + k };
    const uint64_t s = sm.next();
    Xoshiro rng(s);
    const double u = rng.unit();
    const double lo = std::log((double)count_lo), hi = std::log((double)count_hi);
    uint64_t count = (uint64_t)std::floor(std::exp(lo + u * (hi - lo)));
    count = std::max(count_lo, std::min(count_hi, count));
    const int alpha = (int)rng.uniform(0, 2);
    return { count, alpha, s };
}

template <class F>
void parallel_sets(uint32_t nsets, F f)
{
    unsigned T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::atomic<uint32_t> next{ 0 };
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t)
        th.emplace_back([&] {
            for (uint32_t k; (k = next.fetch_add(1)) < nsets;) f(k);
        });
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" int ps5_gen_batch_plan(uint64_t seed, uint32_t nsets, uint64_t count_lo, uint64_t count_hi,
                                  uint32_t len_lo, uint32_t len_hi, uint64_t* counts, uint64_t* bytes)
{
    if (count_lo == 0 || count_hi < count_lo || len_hi < len_lo) return PS5_ERR_ARG;
    parallel_sets(nsets, [&](uint32_t k) {
        const BatchSet b = batch_set(seed, k, count_lo, count_hi);
        Xoshiro lens(b.seed ^ 0x5851F42D4C957F2Dull);
        uint64_t tot = 0;
        for (uint64_t i = 0; i < b.count; ++i) tot += lens.uniform(len_lo, len_hi) + 1;
        counts[k] = b.count;
        bytes[k] = (tot + 7) & ~7ull;      // every set starts 8-byte aligned
    });
    return PS5_OK;
}

extern "C" int ps5_gen_batch_fill(uint64_t seed, uint32_t nsets, uint64_t count_lo, uint64_t count_hi,
                                  uint32_t len_lo, uint32_t len_hi, uint8_t* arena, uint64_t* offs)
{
    if (count_lo == 0 || count_hi < count_lo || len_hi < len_lo) return PS5_ERR_ARG;
    std::vector<uint64_t> counts(nsets), bytes(nsets);
    ps5_gen_batch_plan(seed, nsets, count_lo, count_hi, len_lo, len_hi, counts.data(), bytes.data());
    std::vector<uint64_t> sfirst(nsets + 1, 0), bfirst(nsets + 1, 0);
    for (uint32_t k = 0; k < nsets; ++k) {
        sfirst[k + 1] = sfirst[k] + counts[k];
        bfirst[k + 1] = bfirst[k] + bytes[k];
    }
    static const uint8_t acgt[4] = { 'A', 'C', 'G', 'T' };
    parallel_sets(nsets, [&](uint32_t k) {
        const BatchSet b = batch_set(seed, k, count_lo, count_hi);
        Xoshiro lens(b.seed ^ 0x5851F42D4C957F2Dull), chars(b.seed ^ 0xC4A2C7E5ull);
        uint64_t pos = bfirst[k];
        for (uint64_t i = 0; i < b.count; ++i) {
            const uint64_t L = lens.uniform(len_lo, len_hi);
            offs[sfirst[k] + i] = pos;
            for (uint64_t j = 0; j < L; ++j) {
                uint8_t ch;
                if (b.alpha == 0) ch = acgt[chars.uniform(0, 3)];
                else if (b.alpha == 1) ch = (uint8_t)('0' + chars.uniform(0, 9));
                else ch = (uint8_t)chars.uniform(33, 126);
                arena[pos + j] = ch;
            }
            arena[pos + L] = 0;
            pos += L + 1;
        }
        if (pos < bfirst[k + 1]) memset(arena + pos, 0, bfirst[k + 1] - pos);
    });
    return PS5_OK;
}
