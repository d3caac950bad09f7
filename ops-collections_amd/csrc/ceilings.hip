// Measured bandwidth ceilings of the access patterns the containers use
// (no counterpart in the reference; a measurement entry point like
// mxco_oa_probe_stats):
//   - random reads of `access` bytes (32 or 128) at aligned offsets of a
//     buffer: 128 bytes over a table far larger than the caches is the
//     probe walk's pattern (one line per window); 32 bytes over a 128 MiB
//     buffer is the Bloom filter's block read, served by the MALL;
//   - a streaming copy, the pattern of the region build and the fill, with
//     4 workgroups per CU (a sweep of 4-32 workgroups per CU and 1-8 loads
//     in flight per lane read 4.7-5.5 TB/s; 4 per CU was the fastest, and
//     hipMemcpyAsync device-to-device reads 4.9).
// Reads are spread over lane groups of access/16 lanes (one 16-byte load
// each), eight accesses in flight per group, offsets from a multiplicative
// hash of the access index masked to the largest power-of-two count of units
// that fits the buffer (no division, so the address arithmetic stays a few
// VALU instructions per access); each lane folds what it reads into one word
// written once, so no load is dead.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mxco_common.h"

using namespace mxco;

namespace {

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

template <int LANES>
__global__ __launch_bounds__(256) void random_read_kernel(const uint8_t* buf, uint64_t unit_mask, uint64_t n,
                                                          uint64_t* sink)
{
    constexpr int U = 8;
    const uint32_t lane = threadIdx.x % LANES;
    const uint64_t group = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / LANES;
    const uint64_t groups = ((uint64_t)gridDim.x * blockDim.x) / LANES;
    uint64_t acc = 0;
    for (uint64_t i0 = group; i0 < n; i0 += U * groups) {
        u64x2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t i = i0 + (uint64_t)u * groups;
            if (i < n) {
                const uint64_t unit = ((i * 0x9E3779B97F4A7C15ull) >> 17) & unit_mask;
                v[u] = *(const u64x2*)(buf + unit * (16 * LANES) + lane * 16);
            } else {
                v[u] = u64x2{0, 0};
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc ^= v[u].x + v[u].y;
    }
    sink[(uint64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

__global__ __launch_bounds__(256) void stream_copy_kernel(const u64x2* src, u64x2* dst, uint64_t n16)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" {

MXCO_API mxco_status mxco_diag_random_read(const void* buf, uint64_t bytes, uint32_t access, uint64_t n,
                                           uint64_t* sink, uint64_t sink_words, hipStream_t stream)
{
    if (!buf || !sink) return fail(MXCO_INVALID_ARGUMENT, "random_read: null buffer");
    if (access != 16 && access != 32 && access != 64 && access != 128 && access != 256)
        return fail(MXCO_INVALID_ARGUMENT, "random_read: access must be 16, 32, 64, 128 or 256 bytes");
    if (((uintptr_t)buf % access) != 0 || bytes < access)
        return fail(MXCO_INVALID_ARGUMENT, "random_read: buffer not aligned to the access or too small");
    const unsigned grid = (unsigned)cu_count() * 8;
    if (sink_words < (uint64_t)grid * 256) return fail(MXCO_INVALID_ARGUMENT, "random_read: sink too small");
    uint64_t units = 1;  // the largest power of two of whole units in the buffer
    while (units * 2 <= bytes / access) units *= 2;
    const uint64_t mask = units - 1;
    const uint8_t* b = (const uint8_t*)buf;
    switch (access) {
    case 16: random_read_kernel<1><<<grid, 256, 0, stream>>>(b, mask, n, sink); break;
    case 32: random_read_kernel<2><<<grid, 256, 0, stream>>>(b, mask, n, sink); break;
    case 64: random_read_kernel<4><<<grid, 256, 0, stream>>>(b, mask, n, sink); break;
    case 128: random_read_kernel<8><<<grid, 256, 0, stream>>>(b, mask, n, sink); break;
    default: random_read_kernel<16><<<grid, 256, 0, stream>>>(b, mask, n, sink); break;
    }
    MXCO_HIP(hipGetLastError(), "random_read");
    return ok();
}

MXCO_API mxco_status mxco_diag_stream_copy(const void* src, void* dst, uint64_t bytes, hipStream_t stream)
{
    if (!src || !dst || bytes % 16 || ((uintptr_t)src | (uintptr_t)dst) % 16)
        return fail(MXCO_INVALID_ARGUMENT, "stream_copy: 16-byte aligned buffers and size");
    stream_copy_kernel<<<grid_for(bytes / 16, 4), 256, 0, stream>>>((const u64x2*)src, (u64x2*)dst, bytes / 16);
    MXCO_HIP(hipGetLastError(), "stream_copy");
    return ok();
}

}  // extern "C"
