// ps5_tune.cpp -- autotuner for the sort's shape parameters.
//
// The output of a sort does not depend on how it is split up, only its speed
// does.  Three parameters shape the work: the number of strings the
// multi-workgroup levels aim to leave per bucket (leaf_target: how many
// levels run and how large the on-chip blocks are), the splitter-tree depth
// limit of one level (max_treebits: fan-out per level, histogram size) and the
// oversampling factor (max_oversample: splitter quality against tree-build
// cost).  The tuner times the caller's own input under candidate settings by
// coordinate descent -- one parameter at a time, two passes -- each candidate
// the best of two sorts, and leaves the fastest setting in the context.
#include <algorithm>
#include <chrono>
#include <functional>
#include <vector>

#include "../../include/ps5_mi355x.h"

namespace {

struct Tuning {
    uint32_t leaf, bits, over;
};

int autotune(ps5_ctx* ctx, const std::function<int()>& sort_once, uint32_t* leaf_target, uint32_t* max_treebits,
             uint32_t* max_oversample, double* best_ms)
{
    Tuning best{};
    int rc = ps5_ctx_get_tuning(ctx, &best.leaf, &best.bits, &best.over);
    if (rc) return rc;
    auto time_of = [&](const Tuning& t, double* ms) -> int {
        int r = ps5_ctx_set_tuning(ctx, t.leaf, t.bits, t.over);
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
    static const uint32_t leafs[] = { 96, 128, 192, 256, 384, 512, 768, 1024 };
    static const uint32_t bits[] = { 7, 8, 9, 10 };
    static const uint32_t overs[] = { 4, 8, 16, 32 };
    for (int pass = 0; pass < 2; ++pass) {
        for (int param = 0; param < 3; ++param) {
            const uint32_t* cand = param == 0 ? leafs : param == 1 ? bits : overs;
            const int nc = param == 0 ? (int)(sizeof(leafs) / sizeof(leafs[0])) : 4;
            Tuning round_best = best;
            for (int i = 0; i < nc; ++i) {
                Tuning t = best;
                (param == 0 ? t.leaf : param == 1 ? t.bits : t.over) = cand[i];
                if (t.leaf == best.leaf && t.bits == best.bits && t.over == best.over) continue;
                double ms = 0;
                if ((rc = time_of(t, &ms))) return rc;
                if (ms < best_t) { best_t = ms; round_best = t; }
            }
            best = round_best;
        }
    }
    if ((rc = ps5_ctx_set_tuning(ctx, best.leaf, best.bits, best.over))) return rc;
    if (leaf_target) *leaf_target = best.leaf;
    if (max_treebits) *max_treebits = best.bits;
    if (max_oversample) *max_oversample = best.over;
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
