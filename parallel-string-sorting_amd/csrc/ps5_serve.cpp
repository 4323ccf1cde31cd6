// ps5_serve.cpp -- continuous-batching sort service on one GPU.
//
// Requests (host string sets) are submitted from any thread.  A fixed number
// of slots each own a sort context: its HIP stream and device work buffers.
// Admission is decided on the device (ps5_admit.hpp): submit appends the
// request to a ring in mapped host memory, and a free slot launches k_admit,
// which takes every waiting request that fits under the HBM budget (in
// arrival order, at least one) from a ledger kept in device memory, and
// writes the claim to the slot's mailbox.  The slot sorts the claim as one
// batch (ps5_sort_batch_device: every request is its own set) and reports
// the bytes it releases.  Slots run concurrently, so requests that arrive
// while others are on the GPU are admitted as soon as a slot and memory free
// up.  Each request's result depends only on its own strings.  Requests are
// copied between their host arrays and the slot's device buffers directly;
// device buffers only grow.
//
// Failure isolation (the role of psstest's fork-per-run status
// classification, src/psstest.cpp:310-370): a malformed request is rejected
// at submit; a batch that fails is run again one request at a time, so only
// the request that fails on its own reports an error; a slot that cannot
// start leaves the others running, and when no slot is left every waiting and
// later request completes with an error instead of blocking its waiter.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/ps5_mi355x.h"
#include "ps5_admit.hpp"
#include "ps5_internal.hpp"

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
    bool inject_fault = false;    // test hook: fail this request's sort as a device error would
    std::string why;              // error message of a failed request
};

}  // namespace

struct ps5_server {
    int device = 0;
    uint64_t budget = 0;              // device bytes for staged requests
    uint64_t max_batch = 0;           // strings per batch
    uint64_t warm_strings = 0;        // per-slot buffers allocated up front
    uint64_t warm_bytes = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_space;
    std::unordered_map<uint64_t, Request*> reqs;
    uint64_t next_id = 1;
    // admission state (ps5_admit.hpp): the submission ring in mapped host
    // memory, the ledger in device memory, one claim mailbox per slot
    ps5::AdmitIn* ain = nullptr;          // host view
    ps5::AdmitIn* ain_dev = nullptr;      // device view
    ps5::AdmitSlot* aslot = nullptr;      // host view of kAdmitSlots mailboxes
    ps5::AdmitSlot* aslot_dev = nullptr;
    ps5::AdmitLedger* ledger = nullptr;   // device memory
    std::vector<Request*> by_seq;         // the request of sequence number q at q % kAdmitRing
    uint64_t published = 0;               // requests in the ring (its tail)
    uint64_t claimed = 0;                 // requests claimed by slots, as far as the host has seen
    uint64_t release_epoch = 0;           // batches that released their bytes
    uint64_t in_use_seen = 0;             // bytes in use after the latest claim (from the device)
    bool stop = false;
    uint32_t live_slots = 0;          // slots that started and have not stopped
    uint32_t starting = 0;            // slots still setting up
    std::string slot_error;           // why the last slot failed to start
    bool dead = false;                // no slot could start: every request fails at once
    std::unordered_set<uint64_t> fault_ids;   // test hook: ids whose sort fails as a device error would
    std::vector<std::thread> workers;
    std::vector<double> latencies_ms;
    uint64_t batches = 0, completed = 0, failed = 0, bytes_done = 0;
    Clock::time_point t_first, t_last;
    bool any = false;
};

namespace {

// device bytes a request occupies while it runs: its characters, uint32 and
// uint64 offsets, results, and the sort context's work buffers (~72 B per string)
uint64_t request_bytes(uint64_t arena_bytes, uint64_t n) { return arena_bytes + 16 + n * (4 + 8 + 8 + 72) + 4096; }

// device buffer that only grows, at least by half each time
template <class T>
int grow(T** p, uint64_t* cap, uint64_t need)
{
    if (need <= *cap) return PS5_OK;
    const uint64_t want = std::max<uint64_t>(need, *cap + *cap / 2);
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc(reinterpret_cast<void**>(p), want * sizeof(T)) != hipSuccess) return PS5_ERR_NOMEM;
    *cap = want;
    return PS5_OK;
}

struct SlotBuffers {
    uint8_t* arena = nullptr;
    uint64_t arena_cap = 0;
    uint32_t* offs32 = nullptr;
    uint64_t offs32_cap = 0;
    uint64_t* offs64 = nullptr;
    uint64_t offs64_cap = 0;
    uint32_t* perm = nullptr;
    uint64_t perm_cap = 0;
    uint32_t* lcp = nullptr;
    uint64_t lcp_cap = 0;
    uint64_t* first = nullptr;       // set boundaries (nsets + 1)
    uint64_t first_cap = 0;
    uint64_t* base = nullptr;        // arena position of every set
    uint64_t base_cap = 0;

    int reserve(uint64_t abytes, uint64_t n, uint64_t nsets)
    {
        int rc;
        if ((rc = grow(&arena, &arena_cap, abytes)) || (rc = grow(&offs32, &offs32_cap, n)) ||
            (rc = grow(&offs64, &offs64_cap, n)) || (rc = grow(&perm, &perm_cap, n)) ||
            (rc = grow(&lcp, &lcp_cap, n)) || (rc = grow(&first, &first_cap, nsets + 1)) ||
            (rc = grow(&base, &base_cap, nsets)))
            return rc;
        return PS5_OK;
    }
    void release()
    {
        for (void* q : { (void*)arena, (void*)offs32, (void*)offs64, (void*)perm, (void*)lcp, (void*)first,
                         (void*)base })
            if (q) (void)hipFree(q);
    }
};

// Complete a request (caller holds s->mu).
void complete_locked(ps5_server* s, Request* r, int rc, const std::string& why, Clock::time_point t_start,
                     Clock::time_point t_done)
{
    r->status = rc;
    r->why = why;
    r->t_start = t_start;
    r->t_done = t_done;
    r->done = true;
    s->latencies_ms.push_back(std::chrono::duration<double, std::milli>(t_done - r->t_submit).count());
    s->completed++;
    if (rc == PS5_OK) s->bytes_done += r->arena_bytes;
    else s->failed++;
    s->t_last = t_done;
}

// No slot is left to run requests: every waiting request completes with an
// error, and so will every later one (caller holds s->mu).
void fail_all_locked(ps5_server* s)
{
    s->dead = true;
    const auto now = Clock::now();
    const std::string why = "no serving slot could start: " + s->slot_error;
    for (uint64_t q = s->claimed; q < s->published; ++q)
        complete_locked(s, s->by_seq[q % ps5::kAdmitRing], PS5_ERR_HIP, why, now, now);
    s->claimed = s->published;
    s->cv_space.notify_all();
}

// Stage the requests of `batch` straight from their host arrays into the
// slot's device buffers (no host-side packing), sort them as one batch, copy
// each request's results straight back.  Returns the status of the whole.
int sort_batch(ps5_ctx* ctx, SlotBuffers& D, const std::vector<Request*>& batch, std::vector<uint64_t>& set_first,
               std::vector<uint64_t>& set_base, std::string* why)
{
    hipStream_t st = (hipStream_t)ps5_ctx_stream(ctx);
    const uint32_t nsets = (uint32_t)batch.size();
    set_first.assign(nsets + 1, 0);
    set_base.assign(nsets, 0);
    uint64_t abytes = 0;
    for (uint32_t i = 0; i < nsets; ++i) {
        set_base[i] = abytes;
        // every set starts 8-byte aligned and is followed by at least one
        // zero byte, so no key of one request reaches into the next
        abytes += (batch[i]->arena_bytes + 8) & ~7ull;
        set_first[i + 1] = set_first[i] + batch[i]->n;
    }
    const uint64_t nstr = set_first[nsets];
    for (const Request* r : batch) {
        if (r->inject_fault) {
            *why = "injected fault (test hook)";
            return PS5_ERR_INTERNAL;
        }
    }
    int rc = D.reserve(abytes + 16, std::max<uint64_t>(nstr, 1), nsets);
    if (rc) {
        *why = "device buffers for the batch could not be allocated";
        return rc;
    }
    hipError_t he = hipSuccess;
    auto ok = [&](hipError_t e) { if (e != hipSuccess && he == hipSuccess) he = e; };
    for (uint32_t i = 0; i < nsets; ++i) {
        const Request* r = batch[i];
        const uint64_t pad = ((r->arena_bytes + 8) & ~7ull) - r->arena_bytes;
        if (r->arena_bytes)
            ok(hipMemcpyAsync(D.arena + set_base[i], r->arena, r->arena_bytes, hipMemcpyHostToDevice, st));
        ok(hipMemsetAsync(D.arena + set_base[i] + r->arena_bytes, 0, pad, st));
        if (r->n) ok(hipMemcpyAsync(D.offs32 + set_first[i], r->offs, r->n * 4, hipMemcpyHostToDevice, st));
    }
    ok(hipMemsetAsync(D.arena + abytes, 0, 16, st));
    ok(hipMemcpyAsync(D.first, set_first.data(), (nsets + 1) * 8, hipMemcpyHostToDevice, st));
    ok(hipMemcpyAsync(D.base, set_base.data(), nsets * 8, hipMemcpyHostToDevice, st));
    if (he != hipSuccess) {
        (void)hipStreamSynchronize(st);
        *why = std::string("staging: ") + hipGetErrorString(he);
        return PS5_ERR_HIP;
    }
    if (nstr > 0) {
        if ((rc = ps5::widen_offsets(ctx, D.offs32, D.first, D.base, nsets, nstr, D.offs64)) ||
            (rc = ps5_sort_batch_device(ctx, D.arena, abytes + 16, D.offs64, set_first.data(), nsets, D.perm,
                                        D.lcp))) {
            *why = ps5_last_error();
            (void)hipStreamSynchronize(st);
            return rc;
        }
        for (uint32_t i = 0; i < nsets; ++i) {
            const Request* r = batch[i];
            if (r->n) ok(hipMemcpyAsync(r->perm, D.perm + set_first[i], r->n * 4, hipMemcpyDeviceToHost, st));
            if (r->lcp && r->n > 1)
                ok(hipMemcpyAsync(r->lcp + 1, D.lcp + set_first[i] + 1, (r->n - 1) * 4, hipMemcpyDeviceToHost, st));
        }
    }
    ok(hipStreamSynchronize(st));
    if (he != hipSuccess) {
        *why = std::string("results: ") + hipGetErrorString(he);
        return PS5_ERR_HIP;
    }
    return PS5_OK;
}

// One slot: claim a batch on the device (every waiting request that fits the
// budget, in arrival order), sort it; when the batch fails, run its requests
// again one by one so that only a request that fails alone reports an error.
void run_slot(ps5_server* s, uint32_t slot)
{
    ps5_ctx* ctx = nullptr;
    const int rc0 = ps5_ctx_create(&ctx, s->device);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->starting--;
        if (rc0 != PS5_OK) {
            s->slot_error = ps5_last_error();
            if (s->live_slots == 0 && s->starting == 0) fail_all_locked(s);
        } else {
            s->live_slots++;
        }
    }
    if (rc0 != PS5_OK) {
        s->cv_done.notify_all();
        s->cv_work.notify_all();
        return;
    }
    hipStream_t st = (hipStream_t)ps5_ctx_stream(ctx);
    volatile ps5::AdmitSlot* box = s->aslot + slot;
    SlotBuffers D;
    // warm start: buffers for one batch of the configured size, so that the
    // first requests do not pay for device allocations
    if (s->warm_strings) {
        (void)ps5_ctx_reserve(ctx, s->warm_strings);
        (void)D.reserve(s->warm_bytes + 16, s->warm_strings, 64);
    }
    std::vector<uint64_t> set_first, set_base;
    for (;;) {
        uint64_t epoch;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || s->published > s->claimed; });
            if (s->published == s->claimed) break;     // stopping, and nothing waits
            epoch = s->release_epoch;
        }
        // the claim, decided on the device
        int rc = ps5::admit_claim(s->ain_dev, s->ledger, s->aslot_dev + slot, st);
        if (rc == PS5_OK && hipStreamSynchronize(st) != hipSuccess) rc = PS5_ERR_HIP;
        if (rc != PS5_OK) {
            // the slot cannot claim: it stops, as a slot that could not start
            std::lock_guard<std::mutex> lk(s->mu);
            s->slot_error = ps5_last_error();
            s->live_slots--;
            if (s->live_slots == 0 && s->starting == 0) fail_all_locked(s);
            s->cv_done.notify_all();
            s->cv_work.notify_all();
            D.release();
            ps5_ctx_destroy(ctx);
            return;
        }
        if (box->busy) {
            std::this_thread::yield();
            continue;
        }
        const uint64_t first = box->first, count = box->count, bytes = box->bytes, in_use = box->in_use;
        if (count == 0) {
            // nothing waits (another slot claimed it), or the next request
            // does not fit until a running batch releases its bytes
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->release_epoch != epoch || s->claimed >= s->published; });
            continue;
        }
        std::vector<Request*> batch;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (uint64_t q = first; q < first + count; ++q) batch.push_back(s->by_seq[q % ps5::kAdmitRing]);
            s->claimed = std::max(s->claimed, first + count);
            s->in_use_seen = in_use;
        }
        s->cv_work.notify_all();
        s->cv_space.notify_all();
        const auto t_start = Clock::now();
        std::string why;
        rc = sort_batch(ctx, D, batch, set_first, set_base, &why);
        std::vector<int> rcs(batch.size(), rc);
        std::vector<std::string> whys(batch.size(), why);
        if (rc != PS5_OK && batch.size() > 1) {
            // isolate the failure: every request of the batch on its own
            for (size_t i = 0; i < batch.size(); ++i) {
                whys[i].clear();
                rcs[i] = sort_batch(ctx, D, { batch[i] }, set_first, set_base, &whys[i]);
            }
        }
        const auto t_done = Clock::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (size_t i = 0; i < batch.size(); ++i) complete_locked(s, batch[i], rcs[i], whys[i], t_start, t_done);
            // the batch's bytes go back to the budget (read by the next claim)
            __atomic_fetch_add(&s->ain->released, bytes, __ATOMIC_RELEASE);
            s->release_epoch++;
            s->batches++;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
    D.release();
    ps5_ctx_destroy(ctx);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->live_slots--;
    }
}

void free_admission(ps5_server* s)
{
    if (s->ain) (void)hipHostFree(s->ain);
    if (s->aslot) (void)hipHostFree(s->aslot);
    if (s->ledger) (void)hipFree(s->ledger);
    s->ain = nullptr;
    s->aslot = nullptr;
    s->ledger = nullptr;
}

}  // namespace

extern "C" {

int ps5_server_create(ps5_server** out, int device, uint64_t hbm_budget, uint32_t slots, uint64_t max_batch_strings)
{
    if (!out || slots == 0 || slots > 16 || hbm_budget == 0)
        return ps5::set_error(PS5_ERR_ARG, "ps5_server_create: need out, 1..16 slots and a budget");
    ps5_server* s = new ps5_server;
    s->device = device;
    s->budget = hbm_budget;
    s->max_batch = max_batch_strings ? max_batch_strings : (1ull << 31);
    // admission state: ring and mailboxes in mapped host memory, the ledger
    // in device memory (zeroed before any slot can claim)
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&s->ain), sizeof(ps5::AdmitIn),
                          hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void**>(&s->ain_dev), s->ain, 0);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&s->aslot), ps5::kAdmitSlots * sizeof(ps5::AdmitSlot),
                          hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void**>(&s->aslot_dev), s->aslot, 0);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->ledger), sizeof(ps5::AdmitLedger));
    if (e == hipSuccess) e = hipMemset(s->ledger, 0, sizeof(ps5::AdmitLedger));
    if (e != hipSuccess) {
        // no device to admit on: the server starts without slots, and every
        // request completes with this error (as when no slot can start)
        free_admission(s);
        s->slot_error = std::string("admission state: ") + hipGetErrorString(e);
        s->dead = true;
        *out = s;
        return PS5_OK;
    }
    memset(s->ain, 0, sizeof(ps5::AdmitIn));
    memset(s->aslot, 0, ps5::kAdmitSlots * sizeof(ps5::AdmitSlot));
    s->ain->budget = hbm_budget;
    s->ain->max_batch = std::min<uint64_t>(s->max_batch, ps5::kJoinMaxStrings);
    s->by_seq.assign(ps5::kAdmitRing, nullptr);
    if (max_batch_strings) {    // each slot starts with buffers for a full batch
        s->warm_strings = max_batch_strings;
        s->warm_bytes = std::min<uint64_t>(hbm_budget / slots, max_batch_strings * 64);
    }
    s->starting = slots;
    for (uint32_t i = 0; i < slots; ++i) s->workers.emplace_back(run_slot, s, i);
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
    free_admission(s);
    delete s;
}

uint64_t ps5_server_submit(ps5_server* s, const uint8_t* arena, uint64_t arena_bytes, const uint32_t* offs,
                           uint64_t n, uint32_t* perm, uint32_t* lcp)
{
    if (!s) { ps5::set_error(PS5_ERR_ARG, "ps5_server_submit: server is NULL"); return 0; }
    if (n && (!arena || !offs || !perm)) {
        ps5::set_error(PS5_ERR_ARG, "ps5_server_submit: NULL arena, offsets or result array");
        return 0;
    }
    if (n >= 0xFFFFFF00ull) { ps5::set_error(PS5_ERR_ARG, "ps5_server_submit: n too large"); return 0; }
    // a malformed request is rejected here, before it can share a batch:
    // every string must start inside the arena and the arena must end in a
    // NUL, so that every string ends inside it
    if (n) {
        if (arena_bytes == 0 || arena[arena_bytes - 1] != 0) {
            ps5::set_error(PS5_ERR_ARG, "ps5_server_submit: the arena's last byte must be a NUL");
            return 0;
        }
        uint32_t mx = 0;
        for (uint64_t i = 0; i < n; ++i) mx = std::max(mx, offs[i]);
        if ((uint64_t)mx >= arena_bytes) {
            ps5::set_error(PS5_ERR_ARG, "ps5_server_submit: a string offset lies outside the arena");
            return 0;
        }
    }
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
        std::unique_lock<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        r->inject_fault = s->fault_ids.erase(r->id) != 0;
        s->reqs[r->id] = r;
        if (!s->any) { s->t_first = r->t_submit; s->any = true; }
        // the ring holds kAdmitRing requests that no slot has claimed yet
        s->cv_space.wait(lk, [&] { return s->dead || s->published - s->claimed < ps5::kAdmitRing; });
        if (s->dead) {
            complete_locked(s, r, PS5_ERR_HIP, "no serving slot could start: " + s->slot_error, r->t_submit,
                            Clock::now());
        } else {
            // the entry first, then the tail that makes it visible to k_admit
            const uint64_t q = s->published++;
            s->by_seq[q % ps5::kAdmitRing] = r;
            s->ain->ring[q % ps5::kAdmitRing] = ps5::AdmitEntry{ r->dev_bytes, r->n };
            __atomic_store_n(&s->ain->tail, q + 1, __ATOMIC_RELEASE);
        }
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    return r->id;
}

int ps5_server_wait(ps5_server* s, uint64_t id, double* latency_ms)
{
    if (!s) return ps5::set_error(PS5_ERR_ARG, "ps5_server_wait: server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return ps5::set_error(PS5_ERR_ARG, "ps5_server_wait: unknown request id");
    Request* r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    const int rc = r->status;
    const std::string why = r->why;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    s->reqs.erase(it);
    delete r;
    lk.unlock();
    if (rc != PS5_OK) return ps5::set_error(rc, "request %llu failed: %s", (unsigned long long)id, why.c_str());
    return PS5_OK;
}

int ps5_server_stats(ps5_server* s, double* p50_ms, double* p99_ms, uint64_t* completed, uint64_t* batches,
                     double* throughput_gbs)
{
    if (!s) return ps5::set_error(PS5_ERR_ARG, "ps5_server_stats: server is NULL");
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

int ps5_server_health(ps5_server* s, uint32_t* live_slots, uint64_t* failed)
{
    if (!s) return ps5::set_error(PS5_ERR_ARG, "ps5_server_health: server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    if (live_slots) *live_slots = s->live_slots;
    if (failed) *failed = s->failed;
    return PS5_OK;
}

int ps5_server_ledger(ps5_server* s, uint64_t* admitted, uint64_t* in_use_bytes)
{
    if (!s) return ps5::set_error(PS5_ERR_ARG, "ps5_server_ledger: server is NULL");
    if (!s->ledger) return ps5::set_error(PS5_ERR_HIP, "ps5_server_ledger: no admission state (%s)", s->slot_error.c_str());
    ps5::AdmitLedger L;
    const hipError_t e = hipMemcpy(&L, s->ledger, sizeof(L), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return ps5::set_error(PS5_ERR_HIP, "ps5_server_ledger: %s", hipGetErrorString(e));
    const uint64_t released = __atomic_load_n(&s->ain->released, __ATOMIC_ACQUIRE);
    if (admitted) *admitted = L.head;
    if (in_use_bytes) *in_use_bytes = L.admitted - released;
    return PS5_OK;
}

int ps5_server_inject_fault(ps5_server* s, uint64_t id)
{
    if (!s || id == 0) return ps5::set_error(PS5_ERR_ARG, "ps5_server_inject_fault: bad server or id");
    std::lock_guard<std::mutex> lk(s->mu);
    if (id < s->next_id) return ps5::set_error(PS5_ERR_ARG, "ps5_server_inject_fault: request already submitted");
    s->fault_ids.insert(id);
    return PS5_OK;
}

}  // extern "C"
