// Host-side helpers shared by the C ABI translation units: error state,
// status mapping, device queries and the capacity primes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/mxco.h"

namespace mxco {

// Records `msg` as the calling thread's last error and returns `st`.
mxco_status fail(mxco_status st, const std::string& msg);
// Clears the last error and returns MXCO_OK.
mxco_status ok();
// Maps a HIP error to MXCO_OUT_OF_MEMORY / MXCO_RUNTIME_ERROR with context.
mxco_status hip_fail(hipError_t e, const char* what);

#define MXCO_HIP(call, what)                                   \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return ::mxco::hip_fail(e_, what); \
    } while (0)

// Compute units of the current device (cached per device).
int cu_count();

// Smallest member >= n of the capacity prime table (include/utility/prime.h:15),
// or 0 when n exceeds its largest member 17,177,758,133.
uint64_t prime_lower_bound(uint64_t n);

// MakeValidExtent (detail/extent/extent.inl:21-63).  0 = invalid (the
// reference throws std::invalid_argument).
uint64_t valid_extent(uint64_t requested, uint32_t bucket, uint32_t probing);

// Largest batch one launch accepts: the reference's CheckedBatchSize guard
// (open_addressing_impl.h:534-545) with its stride taken as CUs x 1024.
uint64_t max_batch();

// Persistent-grid size for a launch of `threads` threads in blocks of 256.
unsigned grid_for(uint64_t threads, unsigned blocks_per_cu);

// Integer tuning knob from the environment (dflt when unset or empty).
int env_int(const char* name, int dflt);

}  // namespace mxco
