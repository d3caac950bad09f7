// Batched StaticMultimap reads over many int64 -> int64 tables in one pass
// (the serving engine's multimap ticks; no reference counterpart as such:
// each read's result is StaticMultimap::Find / Count / Retrieve's).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mxco.h"

namespace mxco {

// One read request on one table.  op: MXCO_REQ_MM_FIND (out = smallest
// values), MXCO_REQ_MM_COUNT, MXCO_REQ_MM_RETRIEVE (out = probe keys, out2 =
// values, out_capacity).  On return: result (Count / Retrieve matches) and
// status (MXCO_LENGTH_ERROR for a Retrieve whose matches exceed its capacity:
// nothing written for it, as StaticMultimap::Retrieve).
struct MmRead {
    mxco_mm* table;
    uint32_t op;
    const void* keys;
    uint64_t n;
    void* out;
    void* out2;
    uint64_t out_capacity;
    uint64_t result;
    mxco_status status;
};

// Device and pinned buffers of the batched reads, grown by doubling; an
// outgrown buffer is kept until destroy (hipFree waits for the device).
struct MmBatchWs;
MmBatchWs* mm_batch_ws_create();
// Sizes the buffers for ticks of up to `reads` requests and `keys` queries
// (larger ticks grow them).
mxco_status mm_batch_ws_reserve(MmBatchWs* ws, uint32_t reads, uint64_t keys);
void mm_batch_ws_destroy(MmBatchWs* ws);

// Runs reads[0..n) on stream s: one upload, a walk pass (Find outputs, match
// counts), one scan, a gather pass for the Retrieves, one read-back; then a
// segmented sort for Retrieves with segments longer than the in-place limit.
// Synchronous.  Every table must be an int64 -> int64 StaticMultimap.
mxco_status mm_batch_reads(MmRead* reads, uint32_t n, MmBatchWs* ws, hipStream_t s);

}  // namespace mxco
