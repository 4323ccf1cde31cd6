// ps5_admit.hpp -- admission state of the sort service, shared by its host
// side (ps5_serve.cpp) and its admission kernel (ps5_admit.hip).
//
// Submitted requests are appended to a ring in mapped, coherent host memory
// (AdmitIn): the device reads it directly.  The ledger of what has been
// admitted (next request, bytes admitted so far) lives in device memory and
// is changed only by k_admit, which a free slot launches on its own stream:
// under a device-side lock it takes the next waiting requests in arrival
// order while the HBM budget allows (bytes admitted minus bytes released),
// and writes the claim to that slot's mailbox (AdmitSlot, mapped memory).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace ps5 {

constexpr uint32_t kAdmitRing = 1u << 16;     // requests submitted and not yet admitted, at most
constexpr uint32_t kAdmitSlots = 16;
constexpr uint32_t kAdmitSpin = 1u << 20;     // lock attempts before k_admit gives up (claims nothing)
// A request joins a claim after its first only while the claim then holds at
// most this many strings: larger requests run as batches of their own, since
// batching them saves only the per-sort overhead (about 0.2 ms) and holds
// their results back until their batch mates are sorted and transferred.
constexpr uint64_t kJoinMaxStrings = 1u << 20;

struct AdmitEntry {
    uint64_t dev_bytes;      // device memory the request holds while it runs
    uint64_t n;              // its strings
};

// host-written, device-read (mapped, coherent)
struct AdmitIn {
    uint64_t tail;           // requests submitted so far (sequence numbers 0 .. tail-1)
    uint64_t released;       // bytes of finished batches so far
    uint64_t budget;         // HBM bytes that admitted requests may hold at once
    uint64_t max_batch;      // strings a claim may hold when a request joins it after its first
    AdmitEntry ring[kAdmitRing];   // request seq at seq % kAdmitRing
};

// one slot's claim, written by k_admit (mapped, coherent)
struct AdmitSlot {
    uint64_t first;          // first sequence number claimed
    uint64_t count;          // requests claimed (0: none waiting, or none fits the budget)
    uint64_t bytes;          // their device bytes, released when the batch is done
    uint64_t in_use;         // bytes admitted and not released after this claim
    uint32_t busy;           // the lock was not obtained: nothing claimed, try again
    uint32_t pad;
};

// device memory: the ledger, changed only by k_admit
struct AdmitLedger {
    uint64_t head;           // requests admitted so far
    uint64_t admitted;       // bytes admitted so far
    uint32_t lock;
    uint32_t pad;
};

// Launch k_admit for one slot on stream st (its claim lands in *slot_dev
// when the stream reaches it).
int admit_claim(const AdmitIn* in_dev, AdmitLedger* ledger, AdmitSlot* slot_dev, hipStream_t st);

}  // namespace ps5
