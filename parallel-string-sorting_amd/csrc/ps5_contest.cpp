// ps5_contest.cpp -- the reference's Contest entry points on host string pointers.
//
// Replaces (paths relative to the reference checkout):
//   parallel_sample_sortBTCUI      src/parallel/bingmann-parallel_sample_sort.cpp:31-37
//   parallel_sample_sortBTCUI_out  src/parallel/bingmann-parallel_sample_sort.cpp:44-62
//   parallel_sample_sort_lcp_base  src/parallel/bingmann-parallel_sample_sort.hpp:1733-1748
//   the LCP + character cache output of StringShadowLcpCacheOutPtr
//                                  src/tools/stringptr.hpp:860-969
// behind the signatures of src/tools/contest.hpp:88-91.
//
// The strings are staged to the device as one character arena plus uint32
// offsets.  When the pointers already lie in one span in increasing order --
// psstest's NUL-separated arena (src/psstest.cpp:479-501) or its suffixes of
// one text (src/psstest.cpp:502-518) -- that span is copied as it is;
// otherwise the strings are packed first.  The characters are never written.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/ps5_mi355x.h"
#include "ps5_internal.hpp"

namespace {

std::mutex g_mu;
ps5_ctx* g_ctx = nullptr;

struct Staging {
    uint8_t* d_arena = nullptr;
    uint64_t arena_cap = 0;
    uint32_t* d_offs = nullptr;
    uint32_t* d_perm = nullptr;
    uint32_t* d_lcp = nullptr;
    uint8_t* d_cache = nullptr;
    uint64_t* d_wide = nullptr;      // sorted pointers / widened LCPs, copied straight to the caller
    uint64_t n_cap = 0;
    uint32_t* d_out_idx = nullptr;   // merge outputs
    uint32_t* d_out_lcp = nullptr;
    uint64_t out_cap = 0;
} g_st;

// PS5_TRACE set: phase times of a host-entry sort on stderr
struct PhaseClock {
    bool on = getenv("PS5_TRACE") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void mark(const char* what)
    {
        if (!on) return;
        const auto now = std::chrono::steady_clock::now();
        fprintf(stderr, "[ps5] host %-10s %8.2f ms\n", what, std::chrono::duration<double, std::milli>(now - t).count());
        t = now;
    }
};

unsigned host_threads()
{
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    return std::min(t, 16u);
}

template <class F>
void parallel_for(uint64_t n, F f)
{
    const unsigned T = host_threads();
    if (n < (1u << 16) || T == 1) { f(0, n); return; }
    std::vector<std::thread> th;
    const uint64_t chunk = (n + T - 1) / T;
    for (unsigned t = 0; t < T; ++t) {
        const uint64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=] { f(b, e); });
    }
    for (auto& x : th) x.join();
}

int ensure_staging(uint64_t arena_bytes, uint64_t n)
{
    if (arena_bytes + 16 > g_st.arena_cap) {
        if (g_st.d_arena) (void)hipFree(g_st.d_arena);
        g_st.d_arena = nullptr;
        g_st.arena_cap = 0;
        void* p = nullptr;
        if (ps5_dev_malloc(&p, arena_bytes + 16)) return PS5_ERR_NOMEM;
        g_st.d_arena = (uint8_t*)p;
        g_st.arena_cap = arena_bytes + 16;
    }
    if (n > g_st.n_cap) {
        for (uint32_t** q : { &g_st.d_offs, &g_st.d_perm, &g_st.d_lcp }) {
            if (*q) (void)hipFree(*q);
            *q = nullptr;
        }
        g_st.n_cap = 0;
        if (g_st.d_cache) (void)hipFree(g_st.d_cache);
        g_st.d_cache = nullptr;
        if (g_st.d_wide) (void)hipFree(g_st.d_wide);
        g_st.d_wide = nullptr;
        for (uint32_t** q : { &g_st.d_offs, &g_st.d_perm, &g_st.d_lcp }) {
            void* p = nullptr;
            if (ps5_dev_malloc(&p, n * sizeof(uint32_t))) return PS5_ERR_NOMEM;
            *q = (uint32_t*)p;
        }
        void* p = nullptr;
        if (ps5_dev_malloc(&p, n)) return PS5_ERR_NOMEM;
        g_st.d_cache = (uint8_t*)p;
        if (ps5_dev_malloc(&p, n * sizeof(uint64_t))) return PS5_ERR_NOMEM;
        g_st.d_wide = (uint64_t*)p;
        g_st.n_cap = n;
    }
    return PS5_OK;
}

}  // namespace

extern "C" int ps5_sort_strings_cache(unsigned char** strings, size_t n, uintptr_t* lcp, unsigned char* cache,
                                      unsigned char** output)
{
    if (n == 0) return PS5_OK;
    if (!strings || (cache && !lcp)) return PS5_ERR_ARG;
    if (n >= 0xFFFFFF00ull) return PS5_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_mu);
    int rc;
    if (!g_ctx && (rc = ps5_ctx_create(&g_ctx, 0))) return rc;

    PhaseClock clk;
    // One span: pointers in non-decreasing order, each at most a page after
    // the one before.  That is psstest's arena (one NUL-separated string
    // after the other) and its --suffix mode (every position of one text).
    // Every byte between two consecutive starts then lies on one of their two
    // pages, so the whole span up to the last string's end is readable
    // (no string ends after the last one: a string that reaches past the next
    // start shares that string's end), and it is copied as it is, without a
    // pass over the characters here.
    constexpr uintptr_t kPage = 4096;
    const unsigned T = host_threads();
    std::vector<uint8_t> ok(T, 1);
    {
        std::vector<std::thread> th;
        const uint64_t chunk = (n + T - 1) / T;
        for (unsigned t = 0; t < T; ++t) {
            th.emplace_back([&, t] {
                const uint64_t b = std::max<uint64_t>(t * chunk, 1), e = std::min<uint64_t>(n, t * chunk + chunk);
                bool good = true;
                // (addresses compared as integers: the pointers need not share an allocation)
                for (uint64_t i = b; i < e; ++i) {
                    const uintptr_t a = (uintptr_t)strings[i - 1], c = (uintptr_t)strings[i];
                    good = good && c >= a && c - a <= kPage;
                }
                ok[t] = good;
            });
        }
        for (auto& x : th) x.join();
    }
    bool span = true;
    for (unsigned t = 0; t < T; ++t) span = span && ok[t];
    uint64_t alen = 0;
    if (span) {
        alen = (uint64_t)((uintptr_t)strings[n - 1] - (uintptr_t)strings[0]) + strlen((const char*)strings[n - 1]) + 1;
        span = alen <= 0xFFFFFFF0ull;
    }
    std::unique_ptr<uint32_t[]> offs(new uint32_t[n]);
    std::vector<uint8_t> packed;
    const uint8_t* host_arena;
    if (span) {
        host_arena = strings[0];
        parallel_for(n, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; ++i) offs[i] = (uint32_t)((uintptr_t)strings[i] - (uintptr_t)strings[0]);
        });
    } else {
        std::vector<uint64_t> len(n);
        parallel_for(n, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; ++i) len[i] = strlen((const char*)strings[i]);
        });
        uint64_t tot = 0;
        for (uint64_t i = 0; i < n; ++i) {
            if (tot > 0xFFFFFFF0ull) return PS5_ERR_ARG;
            offs[i] = (uint32_t)tot;
            tot += len[i] + 1;
        }
        alen = tot;
        if (alen > 0xFFFFFFF0ull) return PS5_ERR_ARG;
        packed.resize(alen);
        parallel_for(n, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; ++i) memcpy(&packed[offs[i]], strings[i], len[i] + 1);
        });
        host_arena = packed.data();
    }
    clk.mark("layout");
    if ((rc = ensure_staging(alen, n))) return rc;
    if ((rc = ps5_memcpy_h2d(g_st.d_arena, host_arena, alen))) return rc;
    if ((rc = ps5_memcpy_h2d(g_st.d_offs, offs.get(), n * sizeof(uint32_t)))) return rc;
    clk.mark("to device");
    if ((rc = ps5_sort_device(g_ctx, g_st.d_arena, alen, g_st.d_offs, n, g_st.d_perm, lcp ? g_st.d_lcp : nullptr)))
        return rc;
    clk.mark("sort");
    // LCPs widened on the device and copied straight into lcp[1..n)
    if (lcp && n > 1) {
        if ((rc = ps5::widen_lcps(g_ctx, g_st.d_lcp, n, g_st.d_wide))) return rc;
        if ((rc = ps5_memcpy_d2h(lcp + 1, g_st.d_wide + 1, (n - 1) * sizeof(uint64_t)))) return rc;
    }
    if (cache && n > 1) {
        if ((rc = ps5_dist_chars_device(g_ctx, g_st.d_arena, alen, g_st.d_offs, n, g_st.d_perm, g_st.d_lcp,
                                        g_st.d_cache)))
            return rc;
        if ((rc = ps5_memcpy_d2h(cache + 1, g_st.d_cache + 1, n - 1))) return rc;
    }
    clk.mark("to host");
    unsigned char** out = output ? output : strings;
    static_assert(sizeof(unsigned char*) == sizeof(uint64_t), "pointers are 64-bit");
    if (span) {     // the sorted pointers, computed on the device: base + offset of each sorted string
        if ((rc = ps5::gather_pointers(g_ctx, g_st.d_perm, g_st.d_offs, (uint64_t)(uintptr_t)strings[0], n,
                                       g_st.d_wide)))
            return rc;
        if ((rc = ps5_memcpy_d2h(out, g_st.d_wide, n * sizeof(uint64_t)))) return rc;
    } else {
        std::unique_ptr<uint32_t[]> perm(new uint32_t[n]);
        if ((rc = ps5_memcpy_d2h(perm.get(), g_st.d_perm, n * sizeof(uint32_t)))) return rc;
        if (out == strings) {
            std::vector<unsigned char*> orig(n);
            parallel_for(n, [&](uint64_t b, uint64_t e) { std::copy(strings + b, strings + e, orig.begin() + b); });
            parallel_for(n, [&](uint64_t b, uint64_t e) {
                for (uint64_t i = b; i < e; ++i) out[i] = orig[perm[i]];
            });
        } else {
            parallel_for(n, [&](uint64_t b, uint64_t e) {
                for (uint64_t i = b; i < e; ++i) out[i] = strings[perm[i]];
            });
        }
    }
    clk.mark("permute");
    return PS5_OK;
}

extern "C" int ps5_sort_strings(unsigned char** strings, size_t n, uintptr_t* lcp, unsigned char** output)
{
    return ps5_sort_strings_cache(strings, n, lcp, nullptr, output);
}

static void die_on(int rc, const char* who)
{
    if (rc == PS5_OK) return;
    fprintf(stderr, "%s: error %d: %s\n", who, rc, ps5_last_error());
    abort();
}

extern "C" void mi355x_ps5_sort(unsigned char** strings, size_t n)
{
    die_on(ps5_sort_strings(strings, n, nullptr, nullptr), "mi355x_ps5_sort");
}

extern "C" void mi355x_ps5_sort_lcp(unsigned char** strings, uintptr_t* lcp, size_t n)
{
    die_on(ps5_sort_strings(strings, n, lcp, nullptr), "mi355x_ps5_sort_lcp");
}

extern "C" void mi355x_ps5_sort_out(unsigned char** strings, size_t n)
{
    // a separate output array, copied back (parallel_sample_sortBTCUI_out)
    std::vector<unsigned char*> output(n);
    die_on(ps5_sort_strings(strings, n, nullptr, output.data()), "mi355x_ps5_sort_out");
    std::copy(output.begin(), output.end(), strings);
}

extern "C" void mi355x_ps5_sort_out_base(unsigned char** strings, unsigned char** output, size_t n)
{
    die_on(ps5_sort_strings(strings, n, nullptr, output), "mi355x_ps5_sort_out_base");
}

extern "C" void mi355x_ps5_sort_lcp_cache(unsigned char** strings, uintptr_t* lcp, unsigned char* cache, size_t n)
{
    die_on(ps5_sort_strings_cache(strings, n, lcp, cache, nullptr), "mi355x_ps5_sort_lcp_cache");
}

// ---- merge of sorted runs (host form) ---------------------------------------

extern "C" int ps5_lcp_merge_strings(unsigned char* const* const* runs, const uintptr_t* const* lcps,
                                     const size_t* sizes, unsigned nruns, unsigned char** output,
                                     uintptr_t* out_lcp)
{
    if (nruns == 0) return PS5_OK;
    if (!runs || !lcps || !sizes || !output) return PS5_ERR_ARG;
    std::vector<uint64_t> run_first(nruns + 1, 0);
    for (unsigned r = 0; r < nruns; ++r) {
        if (sizes[r] && (!runs[r] || !lcps[r])) return PS5_ERR_ARG;
        run_first[r + 1] = run_first[r] + sizes[r];
    }
    const uint64_t n = run_first[nruns];
    if (n == 0) return PS5_OK;
    if (n >= 0xFFFFFF00ull) return PS5_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_mu);
    int rc;
    if (!g_ctx && (rc = ps5_ctx_create(&g_ctx, 0))) return rc;
    // the strings of all runs, in run order, packed into one arena: string q
    // (q-th of the concatenation) gets index q
    std::vector<const unsigned char*> str(n);
    for (unsigned r = 0; r < nruns; ++r)
        parallel_for(sizes[r], [&](uint64_t b, uint64_t e) {
            std::copy(runs[r] + b, runs[r] + e, str.begin() + run_first[r] + b);
        });
    std::vector<uint64_t> len(n);
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) len[i] = strlen((const char*)str[i]);
    });
    std::vector<uint32_t> offs(n), idx(n), lcp(n, 0);
    uint64_t tot = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (tot > 0xFFFFFFF0ull) return PS5_ERR_ARG;
        offs[i] = (uint32_t)tot;
        tot += len[i] + 1;
    }
    if (tot > 0xFFFFFFF0ull) return PS5_ERR_ARG;
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) idx[i] = (uint32_t)i;
    });
    std::atomic<bool> lcp_ok{ true };
    for (unsigned r = 0; r < nruns; ++r)
        parallel_for(sizes[r], [&](uint64_t b, uint64_t e) {
            for (uint64_t i = std::max<uint64_t>(b, 1); i < e; ++i) {
                if (lcps[r][i] > 0xFFFFFFFFull) lcp_ok = false;
                lcp[run_first[r] + i] = (uint32_t)lcps[r][i];
            }
        });
    if (!lcp_ok) return PS5_ERR_ARG;
    std::vector<uint8_t> packed(tot);
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) memcpy(&packed[offs[i]], str[i], len[i] + 1);
    });
    if ((rc = ensure_staging(tot, n))) return rc;
    if (n > g_st.out_cap) {
        for (uint32_t** q : { &g_st.d_out_idx, &g_st.d_out_lcp }) {
            if (*q) (void)hipFree(*q);
            *q = nullptr;
        }
        g_st.out_cap = 0;
        for (uint32_t** q : { &g_st.d_out_idx, &g_st.d_out_lcp }) {
            void* p = nullptr;
            if (ps5_dev_malloc(&p, n * sizeof(uint32_t))) return PS5_ERR_NOMEM;
            *q = (uint32_t*)p;
        }
        g_st.out_cap = n;
    }
    if ((rc = ps5_memcpy_h2d(g_st.d_arena, packed.data(), tot))) return rc;
    if ((rc = ps5_memcpy_h2d(g_st.d_offs, offs.data(), n * 4))) return rc;
    if ((rc = ps5_memcpy_h2d(g_st.d_perm, idx.data(), n * 4))) return rc;
    if ((rc = ps5_memcpy_h2d(g_st.d_lcp, lcp.data(), n * 4))) return rc;
    if ((rc = ps5_merge_device(g_ctx, g_st.d_arena, tot, g_st.d_offs, run_first.data(), nruns, g_st.d_perm,
                               g_st.d_lcp, g_st.d_out_idx, g_st.d_out_lcp)))
        return rc;
    if ((rc = ps5_memcpy_d2h(idx.data(), g_st.d_out_idx, n * 4))) return rc;
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) output[i] = const_cast<unsigned char*>(str[idx[i]]);
    });
    if (out_lcp && n > 1) {     // widened on the device, copied straight into out_lcp[1..n)
        if ((rc = ps5::widen_lcps(g_ctx, g_st.d_out_lcp, n, g_st.d_wide))) return rc;
        if ((rc = ps5_memcpy_d2h(out_lcp + 1, g_st.d_wide + 1, (n - 1) * sizeof(uint64_t)))) return rc;
    }
    return PS5_OK;
}
