// ps5_tune.cpp -- autotuner for the sort's shape parameters.
//
// The output of a sort does not depend on how it is split up, only its speed
// does.  Seven parameters shape the work:
//   - tree depth: the number of strings the multi-workgroup levels aim to
//     leave per bucket (leaf_target: how many levels run and how large the
//     on-chip blocks are), the splitter-tree depth limit of one level
//     (max_treebits: fan-out per level, histogram size) and the oversampling
//     factor (max_oversample: splitter quality against tree-build cost);
//   - tile shapes: the keys per wave chunk of ordinary levels and of levels
//     of roots (the unit of stable ranks and of the scatter's LDS reorder:
//     longer chunks give longer write runs per bucket, shorter ones more
//     tiles), and the tiles a small level is spread over at least;
//   - key width: the keys (8 characters each) a leaf pass fetches per string
//     at a time (wide fetches pay on long shared prefixes, narrow ones save
//     bandwidth on short strings).
// The tuner times the caller's own input under candidate settings by
// coordinate descent -- one parameter at a time, two passes -- each candidate
// the best of two sorts, and leaves the fastest setting in the context.
#include <algorithm>
#include <chrono>
#include <functional>
#include <vector>

#include "../../include/ps5_mi355x.h"

namespace {

constexpr int kParams = 7;
struct Tuning {
    uint32_t v[kParams];     // leaf_target, max_treebits, max_oversample, chunk_max, chunk_max_root, min_tiles, leaf_keys
};

int apply(ps5_ctx* ctx, const Tuning& t)
{
    int rc = ps5_ctx_set_tuning(ctx, t.v[0], t.v[1], t.v[2]);
    return rc ? rc : ps5_ctx_set_shape(ctx, t.v[3], t.v[4], t.v[5], t.v[6]);
}

int autotune(ps5_ctx* ctx, const std::function<int()>& sort_once, uint32_t* leaf_target, uint32_t* max_treebits,
             uint32_t* max_oversample, double* best_ms)
{
    Tuning best{};
    int rc = ps5_ctx_get_tuning(ctx, &best.v[0], &best.v[1], &best.v[2]);
    if (!rc) rc = ps5_ctx_get_shape(ctx, &best.v[3], &best.v[4], &best.v[5], &best.v[6]);
    if (rc) return rc;
    auto time_of = [&](const Tuning& t, double* ms) -> int {
        int r = apply(ctx, t);
        if (r) return r;
        double m = 1e300;
        for (int rep = 0; rep < 2; ++rep) {
            const auto a = std::chrono::steady_clock::now();
            if ((r = sort_once())) return r;
            const auto b = std::chrono::steady_clock::now();
            m = std::min(m, std::chrono::duration<double, std::milli>(b - a).count());
        }
        *ms = m;
        return PS5_OK;
    };
    double best_t = 0;
    if ((rc = time_of(best, &best_t))) return rc;
    static const std::vector<uint32_t> cands[kParams] = {
        { 96, 128, 192, 256, 384, 512, 768, 1024 },     // leaf_target
        { 7, 8, 9, 10 },                                 // max_treebits
        { 4, 8, 16, 32 },                                // max_oversample
        { 4096, 6144, 8192, 9216 },                      // chunk_max
        { 6144, 9216, 12032 },                           // chunk_max_root
        { 256, 512, 1024, 2048 },                        // min_tiles
        { 1, 2, 4, 8 },                                  // leaf_keys
    };
    for (int pass = 0; pass < 2; ++pass) {
        for (int param = 0; param < kParams; ++param) {
            Tuning round_best = best;
            for (uint32_t c : cands[param]) {
                if (c == best.v[param]) continue;
                Tuning t = best;
                t.v[param] = c;
                double ms = 0;
                if ((rc = time_of(t, &ms))) return rc;
                if (ms < best_t) { best_t = ms; round_best = t; }
            }
            best = round_best;
        }
    }
    if ((rc = apply(ctx, best))) return rc;
    if (leaf_target) *leaf_target = best.v[0];
    if (max_treebits) *max_treebits = best.v[1];
    if (max_oversample) *max_oversample = best.v[2];
    if (best_ms) *best_ms = best_t;
    return PS5_OK;
}

}  // namespace

extern "C" int ps5_autotune_device(ps5_ctx* ctx, const uint8_t* d_arena, uint64_t arena_bytes, const uint32_t* d_offs,
                                   uint64_t n, uint32_t* d_perm, uint32_t* d_lcp, uint32_t* leaf_target,
                                   uint32_t* max_treebits, uint32_t* max_oversample, double* best_ms)
{
    if (!ctx) return PS5_ERR_ARG;
    return autotune(
        ctx, [&] { return ps5_sort_device(ctx, d_arena, arena_bytes, d_offs, n, d_perm, d_lcp); }, leaf_target,
        max_treebits, max_oversample, best_ms);
}

extern "C" int ps5_autotune_batch_device(ps5_ctx* ctx, const uint8_t* d_arena, uint64_t arena_bytes,
                                         const uint64_t* d_offs, const uint64_t* set_first, uint32_t nsets,
                                         uint32_t* d_perm, uint32_t* d_lcp, uint32_t* leaf_target,
                                         uint32_t* max_treebits, uint32_t* max_oversample, double* best_ms)
{
    if (!ctx) return PS5_ERR_ARG;
    return autotune(
        ctx,
        [&] { return ps5_sort_batch_device(ctx, d_arena, arena_bytes, d_offs, set_first, nsets, d_perm, d_lcp); },
        leaf_target, max_treebits, max_oversample, best_ms);
}
