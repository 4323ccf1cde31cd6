// ps5_internal.hpp -- library-internal helpers shared by the C ABI sources
// (not part of the exported interface).
#pragma once
#include <cstdint>

#include "../../include/ps5_mi355x.h"

namespace ps5 {

// Record the calling thread's last error (ps5_last_error) and return code.
int set_error(int code, const char* fmt, ...);

// Absolute uint64 offsets of a batch staged set by set: for string i of set s
// (set_first[s] <= i < set_first[s+1]), offs64[i] = set_base[s] + offs32[i].
// d_first / d_base are device arrays of nsets+1 / nsets entries.  Enqueued on
// the context's stream (no synchronisation).
int widen_offsets(ps5_ctx* c, const uint32_t* d_offs32, const uint64_t* d_first, const uint64_t* d_base,
                  uint32_t nsets, uint64_t n, uint64_t* d_offs64);

// Host-entry results computed on the device, for a direct copy into the
// caller's arrays: ptrs[i] = base + offs[perm[i]] (the sorted string
// pointers of a caller whose strings lie in one span starting at host
// address base), and wide[i] = lcp[i] (uintptr_t LCPs).  Both run on the
// context's stream and synchronise it.
int gather_pointers(ps5_ctx* c, const uint32_t* d_perm, const uint32_t* d_offs, uint64_t base, uint64_t n,
                    uint64_t* d_ptrs);
int widen_lcps(ps5_ctx* c, const uint32_t* d_lcp, uint64_t n, uint64_t* d_wide);

}  // namespace ps5
