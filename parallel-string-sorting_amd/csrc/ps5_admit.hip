// ps5_admit.hip -- admission kernel of the sort service (see ps5_admit.hpp).
#include <hip/hip_runtime.h>

#include "ps5_admit.hpp"
#include "ps5_internal.hpp"

namespace ps5 {
namespace {

__device__ __forceinline__ uint64_t load_sys(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ void store_sys(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// One wave claims the next waiting requests for one slot.  Lane 0 takes the
// ledger's lock (bounded: after kAdmitSpin attempts the claim reports busy);
// then the wave reads 64 ring entries at a time, and prefix sums over the
// lanes give, for every entry, the bytes and strings of the claim if it ends
// there.  An entry is admitted when all before it are and
//   - the bytes in use plus the claim fit the budget, or it is the first
//     entry while nothing is in use (an oversized request runs alone), and
//   - it is the claim's first entry or the claim's strings fit max_batch,
// which is the arrival-order rule a host loop would apply one entry at a time.
__global__ __launch_bounds__(64) void k_admit(const AdmitIn* in, AdmitLedger* L, AdmitSlot* out)
{
    const uint32_t lane = threadIdx.x;
    uint32_t got = 0;
    if (lane == 0) {
        for (uint32_t spin = 0; spin < kAdmitSpin; ++spin) {
            uint32_t expect = 0;
            if (__hip_atomic_compare_exchange_strong(&L->lock, &expect, 1u, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)) {
                got = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    got = __shfl(got, 0, 64);
    if (!got) {
        if (lane == 0) {
            store_sys(&out->count, 0);
            __hip_atomic_store(&out->busy, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        return;
    }
    uint64_t head = 0, admitted = 0;
    if (lane == 0) {
        head = __hip_atomic_load(&L->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        admitted = __hip_atomic_load(&L->admitted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    head = __shfl(head, 0, 64);
    admitted = __shfl(admitted, 0, 64);
    const uint64_t tail = load_sys(&in->tail);
    const uint64_t released = load_sys(&in->released);
    const uint64_t budget = load_sys(&in->budget), max_batch = load_sys(&in->max_batch);
    const uint64_t in_use = admitted - released;
    uint64_t cnt = 0, need = 0, nstr = 0;
    for (;;) {
        const uint64_t q = head + cnt + lane;
        const bool valid = q < tail && cnt + lane < kAdmitRing;
        uint64_t b = 0, m = 0;
        if (valid) {
            const AdmitEntry* e = &in->ring[q % kAdmitRing];
            b = load_sys(&e->dev_bytes);
            m = load_sys(&e->n);
        }
        uint64_t pb = b, pm = m;     // inclusive prefix sums over the lanes
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint64_t ub = __shfl_up(pb, d, 64), um = __shfl_up(pm, d, 64);
            if (lane >= d) {
                pb += ub;
                pm += um;
            }
        }
        const bool first = cnt + lane == 0;
        const bool ok = valid && (in_use + need + pb <= budget || (first && in_use == 0)) &&
                        (first || nstr + pm <= max_batch);
        const uint64_t okm = __ballot(ok);
        const uint32_t k = ~okm ? (uint32_t)__builtin_ctzll(~okm) : 64u;     // leading admitted lanes
        if (k > 0) {
            need += __shfl(pb, k - 1, 64);
            nstr += __shfl(pm, k - 1, 64);
            cnt += k;
        }
        if (k < 64) break;
    }
    if (lane == 0) {
        __hip_atomic_store(&L->head, head + cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&L->admitted, admitted + need, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        store_sys(&out->first, head);
        store_sys(&out->count, cnt);
        store_sys(&out->bytes, need);
        store_sys(&out->in_use, in_use + need);
        __hip_atomic_store(&out->busy, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&L->lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace

int admit_claim(const AdmitIn* in_dev, AdmitLedger* ledger, AdmitSlot* slot_dev, hipStream_t st)
{
    hipLaunchKernelGGL(k_admit, dim3(1), dim3(64), 0, st, in_dev, ledger, slot_dev);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(PS5_ERR_HIP, "k_admit: %s", hipGetErrorString(e));
    return PS5_OK;
}

}  // namespace ps5
