// StaticMultiset (include/static_multiset.h:36-575) for gfx950.
//
// The reference's contract (static_multiset.h:27-34 and its tests) fixes
// results, not layout:
//   - Insert keeps, in input order, the earliest selected non-sentinel keys
//     that fit the remaining capacity (every duplicate takes one element);
//     it returns the selected keys not kept plus the selected sentinel keys;
//   - RetrieveAll writes the kept keys in accepted insertion order;
//   - Contains / Find / CountEach / CountEachOuter / Count read exact
//     multiplicities; Retrieve writes every match in query order (the query
//     key as probe and as match), or fails with std::length_error and writes
//     nothing.
//
// Layout here: the accepted keys in insertion order (capacity keys), and a
// power-of-two table of T >= 2 * capacity slots kept as two arrays, keys and
// 64-bit multiplicities, probed linearly from murmurhash3_fmix32/64(key).  A
// slot's multiplicity starts at 1, so the thread whose CAS claims the slot
// needs no counter atomic and every later copy adds one.  Keys are written
// once (no erase), so a non-empty key read without an atomic is current.
#include <string.h>

#include <algorithm>
#include <hipcub/hipcub.hpp>
#include <new>
#include <string>
#include <vector>

#include "mxco_common.h"
#include "mxco/detail/hash.h"

using namespace mxco;

struct mxco_ms {
    uint32_t ks;
    uint64_t capacity;
    uint64_t size;
    uint64_t tcap;  // T, a power of two
    uint8_t* ordered;
    uint8_t* tkeys;
    unsigned long long* tcounts;
    uint64_t empty;
    uint64_t* host;  // pinned, 4 words
    void* buf[3];    // scratch: ranks / counts, scan temp, totals
    size_t buf_size[3];
    // outgrown scratch, freed by destroy (hipFree waits for the whole device)
    std::vector<void*> retired;
};

namespace {

template <int KS>
__device__ __forceinline__ uint64_t ld_key(const uint8_t* p, uint64_t i)
{
    if constexpr (KS == 4) return ((const uint32_t*)p)[i];
    else return ((const uint64_t*)p)[i];
}

template <int KS>
__device__ __forceinline__ void st_key(uint8_t* p, uint64_t i, uint64_t v)
{
    if constexpr (KS == 4) ((uint32_t*)p)[i] = (uint32_t)v;
    else ((uint64_t*)p)[i] = v;
}

// Probe start: the container's Hasher, murmurhash3_fmix32 / _fmix64 with
// seed 0 (static_multiset.h:45), masked to the table size.
template <int KS>
__device__ __forceinline__ uint64_t ms_start(uint64_t key, uint64_t mask)
{
    if constexpr (KS == 4) return fmix32((uint32_t)key) & mask;
    else return fmix64(key) & mask;
}

__device__ __forceinline__ bool selected(const uint8_t* stencil, uint32_t esz, uint32_t pred, uint64_t i)
{
    if (!stencil) return true;
    int64_t s;
    if (esz == 1) s = (int8_t)stencil[i];
    else if (esz == 4) s = ((const int32_t*)stencil)[i];
    else s = ((const int64_t*)stencil)[i];
    switch (pred) {
    case MXCO_PRED_NONZERO: return s != 0;
    case MXCO_PRED_ODD: return (s & 1) != 0;
    case MXCO_PRED_EVEN: return (s & 1) == 0;
    default: return true;
    }
}

template <int KS>
__global__ __launch_bounds__(256) void ms_clear_kernel(uint8_t* tkeys, unsigned long long* tcounts, uint64_t tcap,
                                                       uint64_t empty)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < tcap; i += stride) {
        st_key<KS>(tkeys, i, empty);
        tcounts[i] = 1;
    }
}

// flags[i] = selected and not the sentinel; stats[0] += selected sentinels.
template <int KS>
__global__ __launch_bounds__(256) void ms_select_kernel(const uint8_t* keys, const uint8_t* stencil, uint32_t esz,
                                                        uint32_t pred, uint64_t n, uint64_t empty, uint64_t* flags,
                                                        unsigned long long* stats)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const bool sel = selected(stencil, esz, pred, i);
        const bool sentinel = ld_key<KS>(keys, i) == empty;
        flags[i] = (sel && !sentinel) ? 1 : 0;
        bad += (sel && sentinel) ? 1 : 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) bad += __shfl_xor(bad, o);
    if (__lane_id() == 0 && bad) atomicAdd(stats, bad);
}

// ranks = exclusive prefix sums of the flags (n + 1 entries, so flag i =
// ranks[i+1] - ranks[i]).  Kept inputs go to ordered[size + rank] and into
// the table.
template <int KS>
__global__ __launch_bounds__(256) void ms_insert_kernel(const uint8_t* keys, uint64_t n, const uint64_t* ranks,
                                                        uint64_t accepted, uint8_t* ordered, uint64_t size,
                                                        uint8_t* tkeys, unsigned long long* tcounts, uint64_t mask,
                                                        uint64_t empty)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t r = ranks[i];
        if (ranks[i + 1] == r || r >= accepted) continue;
        const uint64_t key = ld_key<KS>(keys, i);
        st_key<KS>(ordered, size + r, key);
        uint64_t slot = ms_start<KS>(key, mask);
        // T >= 2 * capacity > distinct keys: an empty slot or the key is always reached
        for (uint64_t visited = 0; visited <= mask; ++visited) {
            uint64_t cur = ld_key<KS>(tkeys, slot);
            if (cur == empty) {
                if constexpr (KS == 4)
                    cur = atomicCAS((unsigned int*)tkeys + slot, (unsigned int)empty, (unsigned int)key);
                else
                    cur = atomicCAS((unsigned long long*)tkeys + slot, (unsigned long long)empty,
                                    (unsigned long long)key);
                if (cur == empty) break;  // claimed: the slot's multiplicity already reads 1
            }
            if (cur == key) {
                atomicAdd(tcounts + slot, 1ull);
                break;
            }
            slot = (slot + 1) & mask;
        }
    }
}

// Multiplicity of a key (0 for the sentinel and for absent keys).
template <int KS>
__device__ __forceinline__ uint64_t ms_lookup(uint64_t key, const uint8_t* tkeys,
                                              const unsigned long long* tcounts, uint64_t mask, uint64_t empty)
{
    if (key == empty) return 0;
    uint64_t slot = ms_start<KS>(key, mask);
    for (uint64_t visited = 0; visited <= mask; ++visited) {
        const uint64_t cur = ld_key<KS>(tkeys, slot);
        if (cur == key) return tcounts[slot];
        if (cur == empty) return 0;
        slot = (slot + 1) & mask;
    }
    return 0;
}

// mode: MXCO_MS_CONTAINS (u8), MXCO_MS_FIND (key), MXCO_MS_COUNT_EACH (u64),
// MXCO_MS_COUNT_EACH_OUTER (u64, zero becomes one); 4: total into *total.
template <int KS, int MODE>
__global__ __launch_bounds__(256) void ms_query_kernel(const uint8_t* keys, const uint8_t* stencil, uint32_t esz,
                                                       uint32_t pred, uint64_t n, const uint8_t* tkeys,
                                                       const unsigned long long* tcounts, uint64_t mask,
                                                       uint64_t empty, uint8_t* out, unsigned long long* total)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long sum = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t key = ld_key<KS>(keys, i);
        const uint64_t c = selected(stencil, esz, pred, i) ? ms_lookup<KS>(key, tkeys, tcounts, mask, empty) : 0;
        if constexpr (MODE == MXCO_MS_CONTAINS) out[i] = c ? 1 : 0;
        else if constexpr (MODE == MXCO_MS_FIND) st_key<KS>(out, i, c ? key : empty);
        else if constexpr (MODE == MXCO_MS_COUNT_EACH) ((uint64_t*)out)[i] = c;
        else if constexpr (MODE == MXCO_MS_COUNT_EACH_OUTER) ((uint64_t*)out)[i] = c ? c : 1;
        else sum += c;
    }
    if constexpr (MODE == 4) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
        if (__lane_id() == 0 && sum) atomicAdd(total, sum);
    }
}

// Output position p belongs to the query i with offsets[i] <= p < offsets[i+1]
// (offsets: n + 1 exclusive prefix sums of the multiplicities).
template <int KS>
__global__ __launch_bounds__(256) void ms_retrieve_kernel(const uint8_t* keys, uint64_t n, const uint64_t* offsets,
                                                          uint64_t total, uint8_t* probe_out, uint8_t* match_out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += stride) {
        uint64_t lo = 0, hi = n;  // last i in [0, n) with offsets[i] <= p
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) / 2;
            if (offsets[mid] <= p) lo = mid;
            else hi = mid;
        }
        const uint64_t key = ld_key<KS>(keys, lo);
        st_key<KS>(probe_out, p, key);
        st_key<KS>(match_out, p, key);
    }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

mxco_status reserve(mxco_ms* m, int which, size_t bytes)
{
    if (m->buf_size[which] >= bytes) return MXCO_OK;
    // at least double, so the retired buffers together stay below the live one
    const size_t want = std::max<size_t>({bytes, 2 * m->buf_size[which], 4096});
    void* p = nullptr;
    MXCO_HIP(hipMalloc(&p, want), "StaticMultiset: scratch");
    if (m->buf[which]) m->retired.push_back(m->buf[which]);
    m->buf[which] = p;
    m->buf_size[which] = want;
    return MXCO_OK;
}

// out[0..n] = exclusive prefix sums of in[0..n-1] (in[n] is read but does
// not change the result); the total out[n] is also returned in *total.
mxco_status exclusive_sum(mxco_ms* m, const uint64_t* in, uint64_t* out, uint64_t n, hipStream_t s, uint64_t* total)
{
    size_t temp = 0;
    MXCO_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, in, out, (int64_t)n + 1, s), "StaticMultiset: scan");
    if (mxco_status st = reserve(m, 2, temp)) return st;
    MXCO_HIP(hipcub::DeviceScan::ExclusiveSum(m->buf[2], temp, in, out, (int64_t)n + 1, s), "StaticMultiset: scan");
    MXCO_HIP(hipMemcpyAsync(m->host, out + n, 8, hipMemcpyDeviceToHost, s), "StaticMultiset: scan total");
    MXCO_HIP(hipStreamSynchronize(s), "StaticMultiset: scan");
    *total = m->host[0];
    return MXCO_OK;
}

mxco_status check_ms(const mxco_ms* m)
{
    if (!m || !m->ordered) return fail(MXCO_LOGIC_ERROR, "StaticMultiset: moved-from object");
    return MXCO_OK;
}

mxco_status set_pred(const mxco_stencil* st, const uint8_t** data, uint32_t* esz, uint32_t* pred)
{
    *data = nullptr;
    *esz = 1;
    *pred = MXCO_PRED_ALWAYS;
    if (!st) return MXCO_OK;
    if (!st->data) return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: null pointer for nonempty operation");
    if (st->elem_size != 1 && st->elem_size != 4 && st->elem_size != 8)
        return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: stencil elements must be 1, 4 or 8 bytes");
    *data = (const uint8_t*)st->data;
    *esz = st->elem_size;
    *pred = st->pred;
    return MXCO_OK;
}

mxco_status clear_table(mxco_ms* m, hipStream_t s)
{
    const unsigned grid = grid_for(m->tcap, 8);
    if (m->ks == 4) ms_clear_kernel<4><<<grid, 256, 0, s>>>(m->tkeys, m->tcounts, m->tcap, m->empty);
    else ms_clear_kernel<8><<<grid, 256, 0, s>>>(m->tkeys, m->tcounts, m->tcap, m->empty);
    MXCO_HIP(hipGetLastError(), "StaticMultiset: clear");
    MXCO_HIP(hipStreamSynchronize(s), "StaticMultiset: clear");
    m->size = 0;
    return MXCO_OK;
}

// Total multiplicity of n queries (after the reference's overflow guard).
mxco_status count_total(mxco_ms* m, const void* keys, uint64_t n, hipStream_t s, uint64_t* count)
{
    if (m->size && n > UINT64_MAX / m->size)
        return fail(MXCO_OVERFLOW_ERROR, "StaticMultiset: potential count overflow");
    if (mxco_status e = reserve(m, 1, 64)) return e;
    unsigned long long* total = (unsigned long long*)m->buf[1];
    MXCO_HIP(hipMemsetAsync(total, 0, 8, s), "StaticMultiset: count");
    const unsigned grid = grid_for(n, 8);
    const uint64_t mask = m->tcap - 1;
    if (m->ks == 4)
        ms_query_kernel<4, 4><<<grid, 256, 0, s>>>((const uint8_t*)keys, nullptr, 1, 0, n, m->tkeys, m->tcounts, mask,
                                                   m->empty, nullptr, total);
    else
        ms_query_kernel<8, 4><<<grid, 256, 0, s>>>((const uint8_t*)keys, nullptr, 1, 0, n, m->tkeys, m->tcounts, mask,
                                                   m->empty, nullptr, total);
    MXCO_HIP(hipGetLastError(), "StaticMultiset: count");
    MXCO_HIP(hipMemcpyAsync(m->host, total, 8, hipMemcpyDeviceToHost, s), "StaticMultiset: count");
    MXCO_HIP(hipStreamSynchronize(s), "StaticMultiset: count");
    *count = m->host[0];
    return MXCO_OK;
}

}  // namespace

extern "C" {

MXCO_API mxco_status mxco_ms_create(uint32_t key_size, uint64_t capacity, const void* empty_key, hipStream_t stream,
                                    mxco_ms** out)
{
    if (!out || !empty_key) return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: null argument");
    *out = nullptr;
    if (key_size != 4 && key_size != 8) return fail(MXCO_UNSUPPORTED, "StaticMultiset supports int32_t and int64_t keys");
    if (capacity == 0 || capacity > (1ull << 40)) return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: invalid capacity");
    mxco_ms* m = new (std::nothrow) mxco_ms();
    if (!m) return fail(MXCO_OUT_OF_MEMORY, "StaticMultiset: host allocation");
    m->ks = key_size;
    m->capacity = capacity;
    m->tcap = 64;
    while (m->tcap < 2 * capacity) m->tcap *= 2;
    memcpy(&m->empty, empty_key, key_size);
    hipError_t e = hipMalloc((void**)&m->ordered, capacity * key_size);
    if (e == hipSuccess) e = hipMalloc((void**)&m->tkeys, m->tcap * key_size);
    if (e == hipSuccess) e = hipMalloc((void**)&m->tcounts, m->tcap * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipHostMalloc((void**)&m->host, 4 * sizeof(uint64_t), 0);
    mxco_status st = e == hipSuccess ? MXCO_OK
                                     : fail(e == hipErrorOutOfMemory ? MXCO_OUT_OF_MEMORY : MXCO_RUNTIME_ERROR,
                                            "StaticMultiset: allocation");
    if (st == MXCO_OK) st = clear_table(m, stream);
    if (st != MXCO_OK) {
        (void)mxco_ms_destroy(m);
        return st;
    }
    *out = m;
    return ok();
}

MXCO_API mxco_status mxco_ms_destroy(mxco_ms* m)
{
    if (!m) return ok();
    if (m->ordered) (void)hipFree(m->ordered);
    if (m->tkeys) (void)hipFree(m->tkeys);
    if (m->tcounts) (void)hipFree(m->tcounts);
    if (m->host) (void)hipHostFree(m->host);
    for (void* b : m->buf)
        if (b) (void)hipFree(b);
    for (void* b : m->retired) (void)hipFree(b);
    delete m;
    return ok();
}

MXCO_API mxco_status mxco_ms_capacity(const mxco_ms* m, uint64_t* capacity)
{
    if (!capacity) return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: null output");
    *capacity = m && m->ordered ? m->capacity : 0;
    return ok();
}

MXCO_API mxco_status mxco_ms_size(mxco_ms* m, hipStream_t stream, uint64_t* size)
{
    if (mxco_status e = check_ms(m)) return e;
    if (!size) return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: null output");
    MXCO_HIP(hipStreamSynchronize(stream), "StaticMultiset: size");
    *size = m->size;
    return ok();
}

MXCO_API mxco_status mxco_ms_clear(mxco_ms* m, hipStream_t stream)
{
    if (mxco_status e = check_ms(m)) return e;
    if (mxco_status e = clear_table(m, stream)) return e;
    return ok();
}

MXCO_API mxco_status mxco_ms_insert(mxco_ms* m, const void* keys, const mxco_stencil* stencil, uint64_t n,
                                    hipStream_t stream, uint64_t* num_failed)
{
    if (mxco_status e = check_ms(m)) return e;
    if (!num_failed) return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: null output");
    *num_failed = n;
    if (!n || !keys) return ok();  // a null keys pointer fails every input, changing nothing
    const uint8_t* sd;
    uint32_t esz, pred;
    if (mxco_status e = set_pred(stencil, &sd, &esz, &pred)) return e;
    if (mxco_status e = reserve(m, 0, 2 * (n + 1) * sizeof(uint64_t))) return e;
    if (mxco_status e = reserve(m, 1, 64)) return e;
    uint64_t* flags = (uint64_t*)m->buf[0];
    uint64_t* ranks = flags + (n + 1);
    unsigned long long* stats = (unsigned long long*)m->buf[1];
    MXCO_HIP(hipMemsetAsync(stats, 0, 8, stream), "StaticMultiset: insert");
    const unsigned grid = grid_for(n, 8);
    if (m->ks == 4)
        ms_select_kernel<4><<<grid, 256, 0, stream>>>((const uint8_t*)keys, sd, esz, pred, n, m->empty, flags, stats);
    else
        ms_select_kernel<8><<<grid, 256, 0, stream>>>((const uint8_t*)keys, sd, esz, pred, n, m->empty, flags, stats);
    MXCO_HIP(hipGetLastError(), "StaticMultiset: select");
    MXCO_HIP(hipMemcpyAsync(m->host + 1, stats, 8, hipMemcpyDeviceToHost, stream), "StaticMultiset: insert");
    uint64_t valid = 0;
    if (mxco_status e = exclusive_sum(m, flags, ranks, n, stream, &valid)) return e;
    const uint64_t sentinels = m->host[1];
    const uint64_t accepted = std::min(valid, m->capacity - m->size);
    if (accepted) {
        const uint64_t mask = m->tcap - 1;
        if (m->ks == 4)
            ms_insert_kernel<4><<<grid, 256, 0, stream>>>((const uint8_t*)keys, n, ranks, accepted, m->ordered, m->size,
                                                          m->tkeys, m->tcounts, mask, m->empty);
        else
            ms_insert_kernel<8><<<grid, 256, 0, stream>>>((const uint8_t*)keys, n, ranks, accepted, m->ordered, m->size,
                                                          m->tkeys, m->tcounts, mask, m->empty);
        MXCO_HIP(hipGetLastError(), "StaticMultiset: insert");
        MXCO_HIP(hipStreamSynchronize(stream), "StaticMultiset: insert");
        m->size += accepted;
    }
    *num_failed = valid - accepted + sentinels;
    return ok();
}

MXCO_API mxco_status mxco_ms_query(mxco_ms* m, int mode, const void* keys, const mxco_stencil* stencil, void* out,
                                   uint64_t n, hipStream_t stream)
{
    if (mxco_status e = check_ms(m)) return e;
    if (mode < MXCO_MS_CONTAINS || mode > MXCO_MS_COUNT_EACH_OUTER)
        return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: bad query mode");
    if (!n) return ok();
    if (!keys || !out) return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: null pointer for nonempty operation");
    const uint8_t* sd;
    uint32_t esz, pred;
    if (mxco_status e = set_pred(stencil, &sd, &esz, &pred)) return e;
    const unsigned grid = grid_for(n, 8);
    const uint64_t mask = m->tcap - 1;
#define MSQ(K, M)                                                                                                \
    ms_query_kernel<K, M><<<grid, 256, 0, stream>>>((const uint8_t*)keys, sd, esz, pred, n, m->tkeys, m->tcounts, \
                                                    mask, m->empty, (uint8_t*)out, nullptr)
#define MSQK(K)                                              \
    switch (mode) {                                          \
    case MXCO_MS_CONTAINS: MSQ(K, MXCO_MS_CONTAINS); break;  \
    case MXCO_MS_FIND: MSQ(K, MXCO_MS_FIND); break;          \
    case MXCO_MS_COUNT_EACH: MSQ(K, MXCO_MS_COUNT_EACH); break; \
    default: MSQ(K, MXCO_MS_COUNT_EACH_OUTER); break;        \
    }
    if (m->ks == 4) {
        MSQK(4)
    } else {
        MSQK(8)
    }
#undef MSQK
#undef MSQ
    MXCO_HIP(hipGetLastError(), "StaticMultiset: query");
    MXCO_HIP(hipStreamSynchronize(stream), "StaticMultiset: query");
    return ok();
}

MXCO_API mxco_status mxco_ms_count(mxco_ms* m, const void* keys, uint64_t n, hipStream_t stream, uint64_t* count)
{
    if (mxco_status e = check_ms(m)) return e;
    if (!count) return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: null output");
    *count = 0;
    if (!n) return ok();
    if (!keys) return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: null pointer for nonempty operation");
    if (mxco_status e = count_total(m, keys, n, stream, count)) return e;
    return ok();
}

MXCO_API mxco_status mxco_ms_retrieve(mxco_ms* m, const void* keys, uint64_t n, void* probe_out, void* match_out,
                                      uint64_t output_capacity, hipStream_t stream, uint64_t* num_out)
{
    if (mxco_status e = check_ms(m)) return e;
    if (!num_out) return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: null output");
    *num_out = 0;
    if (!n) return ok();
    if (!keys) return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: null pointer for nonempty operation");
    if (m->size && n > UINT64_MAX / m->size)
        return fail(MXCO_OVERFLOW_ERROR, "StaticMultiset: potential count overflow");
    if (mxco_status e = reserve(m, 0, 2 * (n + 1) * sizeof(uint64_t))) return e;
    uint64_t* counts = (uint64_t*)m->buf[0];
    uint64_t* offsets = counts + (n + 1);
    if (mxco_status e = mxco_ms_query(m, MXCO_MS_COUNT_EACH, keys, nullptr, counts, n, stream)) return e;
    uint64_t total = 0;
    if (mxco_status e = exclusive_sum(m, counts, offsets, n, stream, &total)) return e;
    if (total > output_capacity)
        return fail(MXCO_LENGTH_ERROR, "StaticMultiset: output capacity too small");
    if (!total) return ok();
    if (!probe_out || !match_out)
        return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: null pointer for nonempty operation");
    const unsigned grid = grid_for(total, 8);
    if (m->ks == 4)
        ms_retrieve_kernel<4><<<grid, 256, 0, stream>>>((const uint8_t*)keys, n, offsets, total, (uint8_t*)probe_out,
                                                        (uint8_t*)match_out);
    else
        ms_retrieve_kernel<8><<<grid, 256, 0, stream>>>((const uint8_t*)keys, n, offsets, total, (uint8_t*)probe_out,
                                                        (uint8_t*)match_out);
    MXCO_HIP(hipGetLastError(), "StaticMultiset: retrieve");
    MXCO_HIP(hipStreamSynchronize(stream), "StaticMultiset: retrieve");
    *num_out = total;
    return ok();
}

MXCO_API mxco_status mxco_ms_retrieve_all(mxco_ms* m, void* out, uint64_t output_capacity, hipStream_t stream,
                                          uint64_t* num_out)
{
    if (mxco_status e = check_ms(m)) return e;
    if (!num_out) return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: null output");
    *num_out = 0;
    if (m->size > output_capacity) return fail(MXCO_LENGTH_ERROR, "StaticMultiset: output capacity too small");
    if (!m->size) return ok();
    if (!out) return fail(MXCO_INVALID_ARGUMENT, "StaticMultiset: null pointer for nonempty operation");
    MXCO_HIP(hipMemcpyAsync(out, m->ordered, m->size * m->ks, hipMemcpyDeviceToDevice, stream),
             "StaticMultiset: retrieve all");
    MXCO_HIP(hipStreamSynchronize(stream), "StaticMultiset: retrieve all");
    *num_out = m->size;
    return ok();
}

}  // extern "C"
