// ps5_serve.cpp -- continuous-batching sort service on one GPU.
//
// Requests (host string sets) are submitted from any thread.  A fixed number
// of slots each own a sort context: its HIP stream and device work buffers.
// A free slot takes every waiting request that fits under the HBM budget (in
// arrival order, at least one) and sorts them together as one batch
// (ps5_sort_batch_device: every request is its own set).  Slots run
// concurrently, so requests that arrive while others are on the GPU are
// admitted as soon as a slot and memory free up.  Each request's result
// depends only on its own strings.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ps5_mi355x.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id;
    const uint8_t* arena;
    uint64_t arena_bytes;
    const uint32_t* offs;
    uint64_t n;
    uint32_t* perm;
    uint32_t* lcp;
    uint64_t dev_bytes;      // device memory the request needs
    Clock::time_point t_submit, t_start, t_done;
    int status = PS5_OK;
    bool done = false;
};

}  // namespace

struct ps5_server {
    int device = 0;
    uint64_t budget = 0;              // device bytes for staged requests
    uint64_t max_batch = 0;           // strings per batch
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, Request*> reqs;
    uint64_t next_id = 1;
    uint64_t in_use = 0;              // device bytes held by running batches
    bool stop = false;
    std::vector<std::thread> workers;
    std::vector<double> latencies_ms;
    uint64_t batches = 0, completed = 0, bytes_done = 0;
    Clock::time_point t_first, t_last;
    bool any = false;
};

namespace {

// device bytes a request occupies while it runs: its characters, uint64
// offsets, results, and the sort context's work buffers (~48 B per string)
uint64_t request_bytes(uint64_t arena_bytes, uint64_t n) { return arena_bytes + 16 + n * (8 + 8 + 48) + 4096; }

void run_slot(ps5_server* s)
{
    if (hipSetDevice(s->device) != hipSuccess) return;
    ps5_ctx* ctx = nullptr;
    if (ps5_ctx_create(&ctx, s->device) != PS5_OK) return;
    hipStream_t st = (hipStream_t)ps5_ctx_stream(ctx);
    uint8_t* d_arena = nullptr;
    uint64_t* d_offs = nullptr;
    uint32_t *d_perm = nullptr, *d_lcp = nullptr;
    uint64_t cap_bytes = 0, cap_n = 0;
    std::vector<uint64_t> h_offs;
    std::vector<uint8_t> h_arena;
    std::vector<uint32_t> h_perm, h_lcp;
    for (;;) {
        std::vector<Request*> batch;
        uint64_t need = 0, nstr = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                return s->stop || (!s->queue.empty() &&
                                   (s->in_use == 0 || s->in_use + s->queue.front()->dev_bytes <= s->budget));
            });
            if (s->stop && s->queue.empty()) break;
            // admit in arrival order while the budget allows (the first one
            // always when nothing runs, so an oversized request still runs alone)
            while (!s->queue.empty()) {
                Request* r = s->queue.front();
                const bool fits = s->in_use + need + r->dev_bytes <= s->budget;
                if (!(fits || (batch.empty() && s->in_use == 0))) break;
                if (!batch.empty() && nstr + r->n > s->max_batch) break;
                batch.push_back(r);
                need += r->dev_bytes;
                nstr += r->n;
                s->queue.pop_front();
            }
            s->in_use += need;
        }
        const auto t_start = Clock::now();
        // stage the batch: one arena, uint64 offsets, set boundaries
        uint64_t abytes = 0;
        std::vector<uint64_t> set_first(batch.size() + 1, 0);
        for (size_t i = 0; i < batch.size(); ++i) {
            abytes += (batch[i]->arena_bytes + 7) & ~7ull;
            set_first[i + 1] = set_first[i] + batch[i]->n;
        }
        int rc = PS5_OK;
        if (abytes + 16 > cap_bytes || nstr > cap_n) {
            if (d_arena) (void)hipFree(d_arena);
            if (d_offs) (void)hipFree(d_offs);
            if (d_perm) (void)hipFree(d_perm);
            if (d_lcp) (void)hipFree(d_lcp);
            d_arena = nullptr; d_offs = nullptr; d_perm = nullptr; d_lcp = nullptr;
            cap_bytes = std::max(abytes + 16, cap_bytes);
            cap_n = std::max<uint64_t>(nstr, std::max<uint64_t>(cap_n, 1));
            if (hipMalloc(&d_arena, cap_bytes) != hipSuccess || hipMalloc(&d_offs, cap_n * 8) != hipSuccess ||
                hipMalloc(&d_perm, cap_n * 4) != hipSuccess || hipMalloc(&d_lcp, cap_n * 4) != hipSuccess)
                rc = PS5_ERR_NOMEM;
        }
        if (rc == PS5_OK) {
            h_arena.resize(abytes + 16);
            h_offs.resize(nstr);
            uint64_t pos = 0;
            for (size_t i = 0; i < batch.size(); ++i) {
                const Request* r = batch[i];
                memcpy(h_arena.data() + pos, r->arena, r->arena_bytes);
                const uint64_t pad = ((r->arena_bytes + 7) & ~7ull) - r->arena_bytes;
                memset(h_arena.data() + pos + r->arena_bytes, 0, pad);
                for (uint64_t j = 0; j < r->n; ++j) h_offs[set_first[i] + j] = pos + r->offs[j];
                pos += r->arena_bytes + pad;
            }
            memset(h_arena.data() + abytes, 0, 16);
            if (hipMemcpyAsync(d_arena, h_arena.data(), abytes + 16, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(d_offs, h_offs.data(), nstr * 8, hipMemcpyHostToDevice, st) != hipSuccess)
                rc = PS5_ERR_HIP;
        }
        if (rc == PS5_OK && nstr > 0)
            rc = ps5_sort_batch_device(ctx, d_arena, abytes + 16, d_offs, set_first.data(),
                                       (uint32_t)batch.size(), d_perm, d_lcp);
        if (rc == PS5_OK && nstr > 0) {
            h_perm.resize(nstr);
            h_lcp.resize(nstr);
            if (hipMemcpyAsync(h_perm.data(), d_perm, nstr * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(h_lcp.data(), d_lcp, nstr * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                rc = PS5_ERR_HIP;
        }
        const auto t_done = Clock::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (size_t i = 0; i < batch.size(); ++i) {
                Request* r = batch[i];
                if (rc == PS5_OK && r->n) {
                    memcpy(r->perm, h_perm.data() + set_first[i], r->n * 4);
                    if (r->lcp && r->n > 1) memcpy(r->lcp + 1, h_lcp.data() + set_first[i] + 1, (r->n - 1) * 4);
                }
                r->status = rc;
                r->t_start = t_start;
                r->t_done = t_done;
                r->done = true;
                s->latencies_ms.push_back(std::chrono::duration<double, std::milli>(t_done - r->t_submit).count());
                s->completed++;
                s->bytes_done += r->arena_bytes;
                s->t_last = t_done;
            }
            s->in_use -= need;
            s->batches++;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
    if (d_arena) (void)hipFree(d_arena);
    if (d_offs) (void)hipFree(d_offs);
    if (d_perm) (void)hipFree(d_perm);
    if (d_lcp) (void)hipFree(d_lcp);
    ps5_ctx_destroy(ctx);
}

}  // namespace

extern "C" {

int ps5_server_create(ps5_server** out, int device, uint64_t hbm_budget, uint32_t slots, uint64_t max_batch_strings)
{
    if (!out || slots == 0 || slots > 16 || hbm_budget == 0) return PS5_ERR_ARG;
    ps5_server* s = new ps5_server;
    s->device = device;
    s->budget = hbm_budget;
    s->max_batch = max_batch_strings ? max_batch_strings : (1ull << 31);
    for (uint32_t i = 0; i < slots; ++i) s->workers.emplace_back(run_slot, s);
    *out = s;
    return PS5_OK;
}

void ps5_server_destroy(ps5_server* s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (auto& kv : s->reqs) delete kv.second;
    delete s;
}

uint64_t ps5_server_submit(ps5_server* s, const uint8_t* arena, uint64_t arena_bytes, const uint32_t* offs,
                           uint64_t n, uint32_t* perm, uint32_t* lcp)
{
    if (!s || (n && (!arena || !offs || !perm))) return 0;
    Request* r = new Request;
    r->arena = arena;
    r->arena_bytes = arena_bytes;
    r->offs = offs;
    r->n = n;
    r->perm = perm;
    r->lcp = lcp;
    r->dev_bytes = request_bytes(arena_bytes, n);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        s->reqs[r->id] = r;
        s->queue.push_back(r);
        if (!s->any) { s->t_first = r->t_submit; s->any = true; }
    }
    s->cv_work.notify_all();
    return r->id;
}

int ps5_server_wait(ps5_server* s, uint64_t id, double* latency_ms)
{
    if (!s) return PS5_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return PS5_ERR_ARG;
    Request* r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    const int rc = r->status;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    s->reqs.erase(it);
    delete r;
    return rc;
}

int ps5_server_stats(ps5_server* s, double* p50_ms, double* p99_ms, uint64_t* completed, uint64_t* batches,
                     double* throughput_gbs)
{
    if (!s) return PS5_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> v = s->latencies_ms;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) { return v.empty() ? 0.0 : v[std::min(v.size() - 1, (size_t)(q * (v.size() - 1) + 0.5))]; };
    if (p50_ms) *p50_ms = pct(0.50);
    if (p99_ms) *p99_ms = pct(0.99);
    if (completed) *completed = s->completed;
    if (batches) *batches = s->batches;
    const double secs = s->any ? std::chrono::duration<double>(s->t_last - s->t_first).count() : 0.0;
    if (throughput_gbs) *throughput_gbs = secs > 0 ? (double)s->bytes_done / secs / 1e9 : 0.0;
    return PS5_OK;
}

}  // extern "C"
