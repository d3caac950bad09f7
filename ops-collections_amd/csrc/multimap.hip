// StaticMultimap (include/static_multimap.h:88-556) for gfx950.
//
// The reference's public contract (docs/static_multimap_API文档和使用示例.md:21-45)
// fixes results, not layout:
//   - Insert admits the input-order prefix of the selected pairs that fits
//     the remaining capacity, after checking every selected pair for the
//     reserved key/value (std::invalid_argument, nothing modified);
//   - Find returns the minimum value of the key, Contains a bool, Count the
//     total number of matches (repeated queries count again);
//   - Retrieve writes every match grouped by query index with the values of
//     each query ascending; RetrieveAll writes the stored multiset.
//
// Layout here: a power-of-two table of P >= 2 * capacity Pair<K,V> slots
// (load <= 1/2), linear probing from fmix64(key).  Every pair takes its own
// slot (first empty slot on its walk, 64-bit CAS on the key word), so all
// duplicates of a key lie on one contiguous run that ends at an empty slot.
// A walk reads the rest of a 128-byte line per round trip.  Scans and the
// per-query value sort use hipCUB.
#include <string.h>

#include <algorithm>
#include <hipcub/hipcub.hpp>
#include <new>
#include <string>
#include <vector>

#include "multimap_batch.h"
#include "mxco_common.h"
#include "mxco/detail/oa_kernels.h"

using namespace mxco;

struct mxco_mm {
    uint32_t ks, vs, ss;
    uint64_t capacity;
    uint64_t size;
    uint64_t slots_pow2;  // P
    uint8_t* table;
    uint64_t empty_key, empty_val;
    // scratch, grown on demand
    uint32_t* status;
    unsigned long long* counters;  // failure counters of the lane-group insert (oa_kernel)
    uint64_t* host;  // pinned, 4 words
    void* buf[6];
    size_t buf_size[6];
    // Outgrown scratch buffers, freed by destroy: hipFree waits for the whole
    // device, so freeing them while other streams run (a server's map ticks)
    // would stall those streams' requests.
    std::vector<void*> retired;
};

namespace {

template <int KS>
__device__ __forceinline__ uint64_t sext(uint64_t k)
{
    if constexpr (KS == 4) return (uint64_t)(int64_t)(int32_t)(uint32_t)k;
    else return k;
}

template <int KS, int VS>
__device__ __forceinline__ uint64_t mm_start(uint64_t key, uint64_t mask)
{
    return fmix64(sext<KS>(key)) & mask;
}

template <int VS>
__device__ __forceinline__ int64_t sval(uint64_t v)
{
    if constexpr (VS == 4) return (int64_t)(int32_t)(uint32_t)v;
    else return (int64_t)v;
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

// Selection flags and reserved-value check (the reference's Validate kernel).
template <int KS, int VS>
__global__ __launch_bounds__(256) void mm_validate_kernel(const uint8_t* pairs, const uint8_t* stencil, uint32_t esz,
                                                          uint32_t pred, uint64_t n, uint64_t ek, uint64_t ev,
                                                          uint64_t* flags, uint32_t* status)
{
    constexpr int SS = Lay<KS, VS>::SS;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const bool sel = selected(stencil, esz, pred, i);
        if (flags) flags[i] = sel ? 1 : 0;
        if (sel) {
            const uint64_t k = load_bits<KS>(pairs + i * SS);
            const uint64_t v = load_bits<VS>(pairs + i * SS + (SS == 8 ? 4 : 8));
            bad |= (k == ek) | (v == ev);
        }
    }
    if (__ballot(bad) != 0 && __lane_id() == 0) atomicOr(status, 1u);
}

// Insert: input i goes in when it is selected and its selection rank is
// below `accepted` (rank = i for unconditional inserts).
template <int KS, int VS>
__global__ __launch_bounds__(256) void mm_insert_kernel(uint8_t* table, uint64_t mask, const uint8_t* pairs,
                                                        const uint8_t* stencil, uint32_t esz, uint32_t pred,
                                                        const uint64_t* ranks, uint64_t n, uint64_t accepted,
                                                        uint64_t ek, uint64_t ev, uint32_t* status)
{
    using L = Lay<KS, VS>;
    constexpr int SS = L::SS;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (ranks) {
            if (!selected(stencil, esz, pred, i) || ranks[i] >= accepted) continue;
        } else if (i >= accepted) {
            continue;
        }
        const uint64_t key = load_bits<KS>(pairs + i * SS);
        const uint64_t val = load_bits<VS>(pairs + i * SS + (SS == 8 ? 4 : 8));
        uint64_t pos = mm_start<KS, VS>(key, mask);
        bool done = false;
        for (uint64_t scanned = 0; scanned <= mask && !done; ++scanned, pos = (pos + 1) & mask) {
            uint8_t* slot = table + pos * SS;
            if (load_bits<KS>(slot) != ek) continue;
            if constexpr (SS == 8) {
                const uint64_t expect = (ek & 0xffffffffull) | (ev << 32);
                const uint64_t want = (key & 0xffffffffull) | (val << 32);
                done = atomicCAS((unsigned long long*)slot, expect, want) == expect;
            } else {
                if (cas_bits<KS>(slot, ek, key) == ek) {
                    store_bits<VS>(slot + 8, val);
                    done = true;
                }
            }
        }
        if (!done) atomicOr(status, 2u);
    }
}

// The multimap's walk as the map probe kernel's policy: linear probing from
// fmix64 of the sign-extended key (bucket 1 over the power-of-two table), and
// no key compares equal, so every pair takes the first empty slot of its walk.
struct MmFns {
    template <int KS>
    static __device__ __forceinline__ uint32_t hash1(const OaArgs&, uint64_t key)
    {
        return (uint32_t)fmix64(sext<KS>(key));
    }
    template <int KS>
    static __device__ __forceinline__ uint32_t hash2(const OaArgs&, uint64_t)
    {
        return 0u;
    }
    template <int KS>
    static __device__ __forceinline__ bool equal(const OaArgs&, uint64_t, uint64_t)
    {
        return false;
    }
    template <int KS>
    static __device__ __forceinline__ bool sentinel(uint64_t key, uint64_t sentinel_key)
    {
        return key == sentinel_key;
    }
};

// Status bit 2 when the lane-group insert left a pair out (its walk found no
// empty slot).
__global__ void mm_counters_status_kernel(const unsigned long long* counters, uint32_t* status)
{
    if (__ballot(counters[2 * threadIdx.x] != 0) && threadIdx.x == 0) atomicOr(status, 2u);
}

enum { Q_CONTAINS = 0, Q_FIND = 1, Q_COUNT = 2, Q_GATHER = 3 };

// Retrieve's gather sorts a query's values in place (insertion into its
// output segment) when it has at most this many; larger segments go to a
// scratch copy and a segmented sort.
constexpr uint64_t kInlineSort = 64;

// One query's walk: the run of occupied slots from the key's start slot, one
// line-remainder window per round trip.  Returns the matches; Find keeps the
// smallest value in *best (any = a match was seen), Contains stops at the
// first match, Gather writes the matches to positions wpos, wpos + 1, ... of
// the output (inserted into sorted place within [seg_off, seg_off + seg_len)
// when seg_len <= kInlineSort, unsorted otherwise).
template <int KS, int VS, int MODE>
__device__ __forceinline__ uint64_t mm_walk(const uint8_t* table, uint64_t mask, uint64_t key, uint64_t ek,
                                            int64_t& best, bool& any, uint64_t wpos, uint64_t seg_off,
                                            uint64_t seg_len, uint8_t* probe_out, uint8_t* value_out)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    constexpr int SS = L::SS;
    constexpr int SPL = 128 / SS;
    uint64_t matches = 0;
    uint64_t pos = mm_start<KS, VS>(key, mask);
    uint64_t left = mask + 1;
    bool stop = false;
    while (!stop && left) {
        uint32_t wlen = SPL - (uint32_t)(pos & (SPL - 1));
        if (wlen > left) wlen = (uint32_t)left;
        if ((mask + 1) - pos < wlen) wlen = (uint32_t)((mask + 1) - pos);
        Raw s[SPL];
#pragma unroll
        for (int k = 0; k < SPL; ++k)
            if ((uint32_t)k < wlen) s[k] = *(const Raw*)(table + (pos + k) * SS);
#pragma unroll
        for (int k = 0; k < SPL; ++k) {
            if ((uint32_t)k >= wlen || stop) continue;
            const uint64_t sk = L::key(s[k]);
            if (sk == ek) {
                stop = true;
            } else if (sk == key) {
                const int64_t v = sval<VS>(L::val(s[k]));
                if (MODE == Q_CONTAINS) stop = true;
                if (MODE == Q_GATHER) {
                    store_bits<KS>(probe_out + wpos * KS, key);
                    if (seg_len > kInlineSort) {
                        store_bits<VS>(value_out + wpos * VS, (uint64_t)v);
                    } else {
                        // insertion into the sorted prefix of the segment (ascending, signed)
                        uint64_t p = wpos;
                        while (p > seg_off && sval<VS>(load_bits<VS>(value_out + (p - 1) * VS)) > v) {
                            store_bits<VS>(value_out + p * VS, load_bits<VS>(value_out + (p - 1) * VS));
                            --p;
                        }
                        store_bits<VS>(value_out + p * VS, (uint64_t)v);
                    }
                    ++wpos;
                }
                best = (!any || v < best) ? v : best;
                any = true;
                ++matches;
            }
        }
        left -= wlen;
        pos = (pos + wlen) & mask;
    }
    return matches;
}

// One lane per query.
template <int KS, int VS, int MODE>
__global__ __launch_bounds__(256) void mm_query_kernel(const uint8_t* table, uint64_t mask, const uint8_t* keys,
                                                       const uint8_t* stencil, uint32_t esz, uint32_t pred,
                                                       uint64_t n, uint64_t ek, uint64_t ev, uint8_t* out,
                                                       uint64_t* counts, const uint64_t* offsets, uint8_t* probe_out,
                                                       uint8_t* value_out, unsigned long long* total,
                                                       uint64_t out_capacity, unsigned long long* clear)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long local = 0;
    // Count clears the word the following gather counts long segments in
    if (MODE == Q_COUNT && clear && blockIdx.x == 0 && threadIdx.x == 0) *clear = 0;
    if constexpr (MODE == Q_GATHER) {
        // total[1] = the number of matches, for the host; a gather that would
        // overflow the output writes nothing (the host reports the error)
        const uint64_t matches_all = offsets[n];
        if (blockIdx.x == 0 && threadIdx.x == 0) total[1] = matches_all;
        if (matches_all > out_capacity) return;
    }
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t key = load_bits<KS>(keys + i * KS);
        const bool sel = selected(stencil, esz, pred, i);
        uint64_t matches = 0;
        int64_t best = 0;
        bool any = false;
        const uint64_t seg_off = (MODE == Q_GATHER) ? offsets[i] : 0;
        const uint64_t seg_len = (MODE == Q_GATHER) ? offsets[i + 1] - offsets[i] : 0;
        if (MODE == Q_GATHER && seg_len > kInlineSort) {
            // listed for the segmented sort: begins in counts[0..n],
            // ends in counts[n + 1..2n + 1]
            const unsigned long long j = atomicAdd(total, 1ull);
            counts[j] = seg_off;
            counts[n + 1 + j] = seg_off + seg_len;
        }
        if (sel && key != ek)
            matches = mm_walk<KS, VS, MODE>(table, mask, key, ek, best, any, seg_off, seg_off, seg_len, probe_out,
                                            value_out);
        if constexpr (MODE == Q_CONTAINS) out[i] = any ? 1 : 0;
        if constexpr (MODE == Q_FIND) store_bits<VS>(out + i * VS, any ? (uint64_t)best : ev);
        if constexpr (MODE == Q_COUNT) {
            if (counts) counts[i] = matches;
            local += matches;
        }
    }
    if constexpr (MODE == Q_COUNT) {
        if (total) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o);
            if (__lane_id() == 0 && local) atomicAdd(total, local);
        }
    }
}

template <int KS, int VS>
__global__ __launch_bounds__(256) void mm_clear_kernel(uint8_t* table, uint64_t slots, uint64_t ek, uint64_t ev)
{
    constexpr int SS = Lay<KS, VS>::SS;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < slots; i += stride) {
        if constexpr (SS == 8) {
            *(uint64_t*)(table + i * 8) = (ek & 0xffffffffull) | (ev << 32);
        } else {
            u64x2 v;
            v.x = KS == 4 ? (ek & 0xffffffffull) : ek;
            v.y = VS == 4 ? (ev & 0xffffffffull) : ev;
            *(u64x2*)(table + i * 16) = v;
        }
    }
}

template <int KS, int VS>
__global__ __launch_bounds__(256) void mm_occupied_kernel(const uint8_t* table, uint64_t slots, uint64_t ek,
                                                          uint64_t* flags)
{
    constexpr int SS = Lay<KS, VS>::SS;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < slots; i += stride)
        flags[i] = load_bits<KS>(table + i * SS) != ek;
}

template <int KS, int VS>
__global__ __launch_bounds__(256) void mm_export_kernel(const uint8_t* table, uint64_t slots, uint64_t ek,
                                                        const uint64_t* offsets, uint8_t* keys_out,
                                                        uint8_t* values_out)
{
    constexpr int SS = Lay<KS, VS>::SS;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < slots; i += stride) {
        const uint64_t k = load_bits<KS>(table + i * SS);
        if (k == ek) continue;
        const uint64_t o = offsets[i];
        store_bits<KS>(keys_out + o * KS, k);
        store_bits<VS>(values_out + o * VS, load_bits<VS>(table + i * SS + (SS == 8 ? 4 : 8)));
    }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

mxco_status reserve(mxco_mm* m, int which, size_t bytes)
{
    if (m->buf_size[which] >= bytes) return MXCO_OK;
    // at least double: a buffer is replaced O(log) times, and the retired
    // ones together stay below the live one
    size_t want = std::max<size_t>({bytes, 2 * m->buf_size[which], 4096});
    void* p = nullptr;
    MXCO_HIP(hipMalloc(&p, want), "StaticMultimap: scratch");
    // the old buffer is kept until destroy (see `retired`)
    if (m->buf[which]) m->retired.push_back(m->buf[which]);
    m->buf[which] = p;
    m->buf_size[which] = want;
    return MXCO_OK;
}

template <typename T>
T* sbuf(mxco_mm* m, int which)
{
    return (T*)m->buf[which];
}

// out[0..n] = exclusive prefix sums of in[0..n-1]; out[n] (the total) is also
// returned in *total.  in[n] is read but does not affect the result.
mxco_status exclusive_sum(mxco_mm* m, const uint64_t* in, uint64_t* out, uint64_t n, hipStream_t s, uint64_t* total)
{
    size_t temp = 0;
    MXCO_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, in, out, (int64_t)n + 1, s), "scan");
    if (mxco_status st = reserve(m, 3, temp)) return st;
    MXCO_HIP(hipcub::DeviceScan::ExclusiveSum(m->buf[3], temp, in, out, (int64_t)n + 1, s), "scan");
    MXCO_HIP(hipMemcpyAsync(m->host, out + n, 8, hipMemcpyDeviceToHost, s), "scan total");
    MXCO_HIP(hipStreamSynchronize(s), "scan");
    *total = m->host[0];
    return MXCO_OK;
}

#define MM_DISPATCH(m, CALL)                              \
    do {                                                  \
        if ((m)->ks == 4 && (m)->vs == 4) { CALL(4, 4); } \
        else if ((m)->ks == 4) { CALL(4, 8); }            \
        else if ((m)->vs == 4) { CALL(8, 4); }            \
        else { CALL(8, 8); }                              \
    } while (0)

mxco_status check_mm(const mxco_mm* m)
{
    if (!m || !m->table) return fail(MXCO_LOGIC_ERROR, "StaticMultimap: moved-from object");
    return MXCO_OK;
}

mxco_status clear_table(mxco_mm* m, hipStream_t s)
{
    const unsigned grid = grid_for(m->slots_pow2, 8);
#define CLR(K, V) mm_clear_kernel<K, V><<<grid, 256, 0, s>>>(m->table, m->slots_pow2, m->empty_key, m->empty_val)
    MM_DISPATCH(m, CLR);
#undef CLR
    MXCO_HIP(hipGetLastError(), "StaticMultimap: clear");
    MXCO_HIP(hipStreamSynchronize(s), "StaticMultimap: clear");
    m->size = 0;
    return MXCO_OK;
}

mxco_status set_pred(const mxco_stencil* st, const uint8_t** data, uint32_t* esz, uint32_t* pred)
{
    *data = nullptr;
    *esz = 1;
    *pred = MXCO_PRED_ALWAYS;
    if (!st) return MXCO_OK;
    if (!st->data) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: null nonempty buffer");
    if (st->elem_size != 1 && st->elem_size != 4 && st->elem_size != 8)
        return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: stencil elements must be 1, 4 or 8 bytes");
    *data = (const uint8_t*)st->data;
    *esz = st->elem_size;
    *pred = st->pred;
    return MXCO_OK;
}

mxco_status query(mxco_mm* m, int mode, const void* keys, const mxco_stencil* st, void* out, uint64_t n,
                  uint64_t* counts, const uint64_t* offsets, void* probe_out, void* value_out,
                  unsigned long long* total, hipStream_t s, uint64_t out_capacity = 0,
                  unsigned long long* clear = nullptr)
{
    const uint8_t* sd;
    uint32_t esz, pred;
    if (mxco_status e = set_pred(st, &sd, &esz, &pred)) return e;
    const unsigned grid = grid_for(n, 16);
    const uint64_t mask = m->slots_pow2 - 1;
#define QL(K, V)                                                                                                  \
    switch (mode) {                                                                                               \
    case Q_CONTAINS:                                                                                              \
        mm_query_kernel<K, V, Q_CONTAINS><<<grid, 256, 0, s>>>(m->table, mask, (const uint8_t*)keys, sd, esz,   \
                                                               pred, n, m->empty_key, m->empty_val,               \
                                                               (uint8_t*)out, nullptr, nullptr, nullptr, nullptr, \
                                                               nullptr, 0, nullptr);                              \
        break;                                                                                                    \
    case Q_FIND:                                                                                                  \
        mm_query_kernel<K, V, Q_FIND><<<grid, 256, 0, s>>>(m->table, mask, (const uint8_t*)keys, sd, esz, pred,  \
                                                           n, m->empty_key, m->empty_val, (uint8_t*)out, nullptr, \
                                                           nullptr, nullptr, nullptr, nullptr, 0, nullptr);       \
        break;                                                                                                    \
    case Q_COUNT:                                                                                                 \
        mm_query_kernel<K, V, Q_COUNT><<<grid, 256, 0, s>>>(m->table, mask, (const uint8_t*)keys, sd, esz, pred, \
                                                            n, m->empty_key, m->empty_val, nullptr, counts,       \
                                                            nullptr, nullptr, nullptr, total, 0, clear);          \
        break;                                                                                                    \
    default:                                                                                                      \
        mm_query_kernel<K, V, Q_GATHER><<<grid, 256, 0, s>>>(m->table, mask, (const uint8_t*)keys, sd, esz,      \
                                                             pred, n, m->empty_key, m->empty_val, (uint8_t*)out,  \
                                                             counts, offsets, (uint8_t*)probe_out,               \
                                                             (uint8_t*)value_out, total, out_capacity, nullptr);  \
    }
    MM_DISPATCH(m, QL);
#undef QL
    MXCO_HIP(hipGetLastError(), "StaticMultimap: query");
    return MXCO_OK;
}

}  // namespace

extern "C" {

MXCO_API mxco_status mxco_mm_create(uint32_t key_size, uint32_t value_size, uint64_t capacity, const void* empty_key,
                                    const void* empty_value, hipStream_t stream, mxco_mm** out)
{
    if (!out || !empty_key || !empty_value) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: null argument");
    *out = nullptr;
    if ((key_size != 4 && key_size != 8) || (value_size != 4 && value_size != 8))
        return fail(MXCO_UNSUPPORTED, "StaticMultimap supports signed I32/I64 keys and values");
    if (capacity == 0) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: capacity must be positive");
    if (capacity > (1ull << 40)) return fail(MXCO_LENGTH_ERROR, "StaticMultimap: capacity too large");
    mxco_mm* m = new (std::nothrow) mxco_mm();
    if (!m) return fail(MXCO_OUT_OF_MEMORY, "StaticMultimap: host allocation");
    m->ks = key_size;
    m->vs = value_size;
    m->ss = (key_size == 4 && value_size == 4) ? 8 : 16;
    m->capacity = capacity;
    uint64_t p = 64;
    while (p < 2 * capacity) p <<= 1;
    m->slots_pow2 = p;
    memcpy(&m->empty_key, empty_key, key_size);
    memcpy(&m->empty_val, empty_value, value_size);
    hipError_t e = hipMalloc((void**)&m->table, p * m->ss);
    if (e == hipSuccess) e = hipMalloc((void**)&m->status, 4 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&m->counters, 2 * kCounterSlots * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipHostMalloc((void**)&m->host, 4 * sizeof(uint64_t), 0);
    mxco_status st = e == hipSuccess ? clear_table(m, stream) : hip_fail(e, "StaticMultimap: allocation");
    if (st != MXCO_OK) {
        if (m->table) (void)hipFree(m->table);
        if (m->status) (void)hipFree(m->status);
        if (m->host) (void)hipHostFree(m->host);
        delete m;
        return st;
    }
    *out = m;
    return ok();
}

MXCO_API mxco_status mxco_mm_destroy(mxco_mm* m)
{
    if (!m) return ok();
    if (m->table) (void)hipFree(m->table);
    if (m->status) (void)hipFree(m->status);
    if (m->counters) (void)hipFree(m->counters);
    if (m->host) (void)hipHostFree(m->host);
    for (int i = 0; i < 6; ++i)
        if (m->buf[i]) (void)hipFree(m->buf[i]);
    for (void* p : m->retired) (void)hipFree(p);
    delete m;
    return ok();
}

MXCO_API mxco_status mxco_mm_capacity(const mxco_mm* m, uint64_t* capacity)
{
    if (!capacity) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: null output");
    *capacity = m && m->table ? m->capacity : 0;
    return ok();
}

MXCO_API mxco_status mxco_mm_size(mxco_mm* m, hipStream_t stream, uint64_t* size)
{
    if (mxco_status s = check_mm(m)) return s;
    if (!size) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: null output");
    MXCO_HIP(hipStreamSynchronize(stream), "StaticMultimap: size");
    *size = m->size;
    return ok();
}

MXCO_API mxco_status mxco_mm_clear(mxco_mm* m, hipStream_t stream)
{
    if (mxco_status s = check_mm(m)) return s;
    if (mxco_status s = clear_table(m, stream)) return s;
    return ok();
}

MXCO_API mxco_status mxco_mm_insert(mxco_mm* m, const void* pairs, const mxco_stencil* stencil, uint64_t n,
                                    hipStream_t s, uint64_t* num_failed)
{
    if (mxco_status e = check_mm(m)) return e;
    if (num_failed) *num_failed = 0;
    if (n == 0) return ok();
    if (!pairs) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: null nonempty buffer");
    const uint8_t* sd;
    uint32_t esz, pred;
    if (mxco_status e = set_pred(stencil, &sd, &esz, &pred)) return e;
    const bool conditional = sd != nullptr && pred != MXCO_PRED_ALWAYS;
    if (!conditional) sd = nullptr;
    uint64_t* flags = nullptr;
    if (conditional) {
        if (mxco_status e = reserve(m, 0, (n + 1) * sizeof(uint64_t))) return e;
        flags = sbuf<uint64_t>(m, 0);
    }
    MXCO_HIP(hipMemsetAsync(m->status, 0, sizeof(uint32_t), s), "StaticMultimap: clear status");
    const unsigned grid = grid_for(n, 8);
#define VAL(K, V)                                                                                                  \
    mm_validate_kernel<K, V><<<grid, 256, 0, s>>>((const uint8_t*)pairs, sd, esz, pred, n, m->empty_key,          \
                                                  m->empty_val, flags, m->status)
    MM_DISPATCH(m, VAL);
#undef VAL
    MXCO_HIP(hipGetLastError(), "StaticMultimap: validate");
    uint64_t selected = n;
    uint64_t* ranks = nullptr;
    if (conditional) {
        if (mxco_status e = reserve(m, 4, (n + 1) * sizeof(uint64_t))) return e;
        ranks = sbuf<uint64_t>(m, 4);
        if (mxco_status e = exclusive_sum(m, flags, ranks, n, s, &selected)) return e;
    }
    MXCO_HIP(hipMemcpyAsync(m->host + 1, m->status, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "status");
    MXCO_HIP(hipStreamSynchronize(s), "StaticMultimap: validate");
    if (((uint32_t*)(m->host + 1))[0] & 1u)
        return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: selected input contains a reserved key or value");
    const uint64_t room = m->capacity - m->size;
    const uint64_t accepted = selected < room ? selected : room;
    if (accepted && !ranks && m->slots_pow2 <= (1ull << 31)) {
        // the first `accepted` pairs: the map probe kernel with lane groups of 4
        // (one 128-byte line per window), a walk that never finds its key
        OaArgs a;
        memset(&a, 0, sizeof(a));
        a.table = m->table;
        a.in = (const uint8_t*)pairs;
        a.counters = m->counters;
        a.n = accepted;
        a.empty_key = m->empty_key;
        a.empty_val = m->empty_val;
        a.erased_key = m->empty_key;
        a.cap = (uint32_t)m->slots_pow2;
        a.bucket = 1;
        a.nbuckets = (uint32_t)m->slots_pow2;
        zero_counters_kernel<<<1, 128, 0, s>>>(m->counters, 2 * kCounterSlots);
        const unsigned ggrid = grid_for(accepted * 4, 256);
#define INS(K, V) oa_kernel<K, V, 4, OP_INSERT, 1, MmFns><<<ggrid, 256, 0, s>>>(a)
        MM_DISPATCH(m, INS);
#undef INS
        mm_counters_status_kernel<<<1, kCounterSlots, 0, s>>>(m->counters, m->status);
    } else if (accepted) {
#define INS(K, V)                                                                                                  \
    mm_insert_kernel<K, V><<<grid, 256, 0, s>>>(m->table, m->slots_pow2 - 1, (const uint8_t*)pairs, sd, esz, pred, \
                                                ranks, n, accepted, m->empty_key, m->empty_val, m->status)
        MM_DISPATCH(m, INS);
#undef INS
    }
    if (accepted) {
        MXCO_HIP(hipGetLastError(), "StaticMultimap: insert");
        MXCO_HIP(hipMemcpyAsync(m->host + 1, m->status, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "status");
        MXCO_HIP(hipStreamSynchronize(s), "StaticMultimap: insert");
        if (((uint32_t*)(m->host + 1))[0] & 2u)
            return fail(MXCO_RUNTIME_ERROR, "StaticMultimap: bounded insertion failed despite available capacity");
        m->size += accepted;
    }
    if (num_failed) *num_failed = selected - accepted;
    return ok();
}

MXCO_API mxco_status mxco_mm_contains(mxco_mm* m, const void* keys, const mxco_stencil* stencil, uint8_t* out,
                                      uint64_t n, hipStream_t s)
{
    if (mxco_status e = check_mm(m)) return e;
    if (n == 0) return ok();
    if (!keys || !out) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: null nonempty buffer");
    if (mxco_status e = query(m, Q_CONTAINS, keys, stencil, out, n, nullptr, nullptr, nullptr, nullptr, nullptr, s))
        return e;
    MXCO_HIP(hipStreamSynchronize(s), "StaticMultimap: contains");
    return ok();
}

MXCO_API mxco_status mxco_mm_find(mxco_mm* m, const void* keys, const mxco_stencil* stencil, void* out, uint64_t n,
                                  hipStream_t s)
{
    if (mxco_status e = check_mm(m)) return e;
    if (n == 0) return ok();
    if (!keys || !out) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: null nonempty buffer");
    if (mxco_status e = query(m, Q_FIND, keys, stencil, out, n, nullptr, nullptr, nullptr, nullptr, nullptr, s))
        return e;
    MXCO_HIP(hipStreamSynchronize(s), "StaticMultimap: find");
    return ok();
}

MXCO_API mxco_status mxco_mm_count(mxco_mm* m, const void* keys, uint64_t n, hipStream_t s, uint64_t* count)
{
    if (mxco_status e = check_mm(m)) return e;
    if (!count) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: null output");
    *count = 0;
    if (n == 0) return ok();
    if (!keys) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: null nonempty buffer");
    if (mxco_status e = reserve(m, 2, sizeof(unsigned long long))) return e;
    unsigned long long* total = sbuf<unsigned long long>(m, 2);
    MXCO_HIP(hipMemsetAsync(total, 0, 8, s), "StaticMultimap: count");
    if (mxco_status e = query(m, Q_COUNT, keys, nullptr, nullptr, n, nullptr, nullptr, nullptr, nullptr, total, s))
        return e;
    MXCO_HIP(hipMemcpyAsync(m->host, total, 8, hipMemcpyDeviceToHost, s), "StaticMultimap: count");
    MXCO_HIP(hipStreamSynchronize(s), "StaticMultimap: count");
    *count = m->host[0];
    return ok();
}

MXCO_API mxco_status mxco_mm_retrieve(mxco_mm* m, const void* keys, uint64_t n, void* probe_out, void* value_out,
                                      uint64_t output_capacity, hipStream_t s, uint64_t* num_out)
{
    if (mxco_status e = check_mm(m)) return e;
    if (!num_out) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: null output");
    *num_out = 0;
    if (n == 0) return ok();
    if (!keys) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: null nonempty buffer");
    // counts -> exclusive offsets (n + 1 entries); afterwards counts[] holds
    // the list of segments too long to sort in place (2n + 2 entries)
    if (mxco_status e = reserve(m, 0, (2 * n + 2) * sizeof(uint64_t))) return e;
    if (mxco_status e = reserve(m, 4, (n + 1) * sizeof(uint64_t))) return e;
    if (mxco_status e = reserve(m, 2, 2 * sizeof(unsigned long long))) return e;
    uint64_t* counts = sbuf<uint64_t>(m, 0);
    uint64_t* offsets = sbuf<uint64_t>(m, 4);
    // res[0]: long segments listed by the gather, res[1]: matches
    unsigned long long* res = sbuf<unsigned long long>(m, 2);
    if (!probe_out || !value_out) {
        // no output: only legal when nothing matches
        uint64_t total = 0;
        if (mxco_status e = query(m, Q_COUNT, keys, nullptr, nullptr, n, counts, nullptr, nullptr, nullptr, nullptr,
                                  s))
            return e;
        if (mxco_status e = exclusive_sum(m, counts, offsets, n, s, &total)) return e;
        if (total > output_capacity)
            return fail(MXCO_LENGTH_ERROR, "StaticMultimap: output capacity is insufficient");
        if (total) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: null nonempty buffer");
        return ok();
    }
    // count -> scan -> gather back to back, one read-back: the gather checks
    // the output capacity on the device and writes nothing when it is short.
    // Segments of up to kInlineSort values are sorted in place in value_out;
    // longer ones land there unsorted and are listed for a segmented sort.
    if (mxco_status e = query(m, Q_COUNT, keys, nullptr, nullptr, n, counts, nullptr, nullptr, nullptr, nullptr, s,
                              0, res))
        return e;
    size_t temp = 0;
    MXCO_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, counts, offsets, (int64_t)n + 1, s), "scan");
    if (mxco_status e = reserve(m, 3, temp)) return e;
    MXCO_HIP(hipcub::DeviceScan::ExclusiveSum(m->buf[3], temp, counts, offsets, (int64_t)n + 1, s), "scan");
    if (mxco_status e = query(m, Q_GATHER, keys, nullptr, nullptr, n, counts, offsets, probe_out, value_out, res, s,
                              output_capacity))
        return e;
    MXCO_HIP(hipMemcpyAsync(m->host, res, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s), "StaticMultimap: retrieve");
    MXCO_HIP(hipStreamSynchronize(s), "StaticMultimap: retrieve");
    const int64_t nlong = (int64_t)m->host[0];
    const uint64_t total = m->host[1];
    if (total > output_capacity) return fail(MXCO_LENGTH_ERROR, "StaticMultimap: output capacity is insufficient");
    if (nlong > 0) {
        // the sort reads a copy of the output and writes the listed segments back
        if (mxco_status e = reserve(m, 1, total * m->vs)) return e;
        MXCO_HIP(hipMemcpyAsync(m->buf[1], value_out, total * m->vs, hipMemcpyDeviceToDevice, s),
                 "StaticMultimap: retrieve");
        const uint64_t* begins = counts;
        const uint64_t* ends = counts + n + 1;
        hipError_t he;
        if (m->vs == 4) {
            he = hipcub::DeviceSegmentedSort::SortKeys(nullptr, temp, (const int32_t*)m->buf[1], (int32_t*)value_out,
                                                      (int64_t)total, nlong, begins, ends, s);
        } else {
            he = hipcub::DeviceSegmentedSort::SortKeys(nullptr, temp, (const int64_t*)m->buf[1], (int64_t*)value_out,
                                                      (int64_t)total, nlong, begins, ends, s);
        }
        MXCO_HIP(he, "StaticMultimap: sort");
        if (mxco_status e = reserve(m, 3, temp)) return e;
        if (m->vs == 4) {
            he = hipcub::DeviceSegmentedSort::SortKeys(m->buf[3], temp, (const int32_t*)m->buf[1], (int32_t*)value_out,
                                                      (int64_t)total, nlong, begins, ends, s);
        } else {
            he = hipcub::DeviceSegmentedSort::SortKeys(m->buf[3], temp, (const int64_t*)m->buf[1], (int64_t*)value_out,
                                                      (int64_t)total, nlong, begins, ends, s);
        }
        MXCO_HIP(he, "StaticMultimap: sort");
        MXCO_HIP(hipStreamSynchronize(s), "StaticMultimap: retrieve");
    }
    *num_out = total;
    return ok();
}

MXCO_API mxco_status mxco_mm_retrieve_all(mxco_mm* m, void* key_out, void* value_out, uint64_t output_capacity,
                                          hipStream_t s, uint64_t* num_out)
{
    if (mxco_status e = check_mm(m)) return e;
    if (!num_out) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: null output");
    *num_out = 0;
    if (m->size > output_capacity) return fail(MXCO_LENGTH_ERROR, "StaticMultimap: output capacity is insufficient");
    if (m->size == 0) return ok();
    if (!key_out || !value_out) return fail(MXCO_INVALID_ARGUMENT, "StaticMultimap: null nonempty buffer");
    const uint64_t P = m->slots_pow2;
    if (mxco_status e = reserve(m, 0, (P + 1) * sizeof(uint64_t))) return e;
    if (mxco_status e = reserve(m, 4, (P + 1) * sizeof(uint64_t))) return e;
    uint64_t* flags = sbuf<uint64_t>(m, 0);
    uint64_t* offsets = sbuf<uint64_t>(m, 4);
    const unsigned grid = grid_for(P, 8);
#define OCC(K, V) mm_occupied_kernel<K, V><<<grid, 256, 0, s>>>(m->table, P, m->empty_key, flags)
    MM_DISPATCH(m, OCC);
#undef OCC
    MXCO_HIP(hipGetLastError(), "StaticMultimap: retrieve all");
    uint64_t total = 0;
    if (mxco_status e = exclusive_sum(m, flags, offsets, P, s, &total)) return e;
#define EXP(K, V) \
    mm_export_kernel<K, V><<<grid, 256, 0, s>>>(m->table, P, m->empty_key, offsets, (uint8_t*)key_out, (uint8_t*)value_out)
    MM_DISPATCH(m, EXP);
#undef EXP
    MXCO_HIP(hipGetLastError(), "StaticMultimap: retrieve all");
    MXCO_HIP(hipStreamSynchronize(s), "StaticMultimap: retrieve all");
    *num_out = total;
    return ok();
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched reads over many tables (multimap_batch.h)
// ---------------------------------------------------------------------------
namespace mxco {

namespace {

struct MmBatchItem {
    const uint8_t* table;
    const uint8_t* keys;
    uint8_t* out;   // Find: values; Retrieve: probe keys
    uint8_t* out2;  // Retrieve: values
    uint64_t mask, ek, ev;
    uint64_t n;
    uint64_t base;      // Count / Retrieve: the item's first entry of the counts and offsets
    uint64_t capacity;  // Retrieve
    uint32_t mode;      // Q_FIND, Q_COUNT (Count) or Q_GATHER (Retrieve)
    uint32_t pad;
};

struct MmBatchChunk {
    uint32_t item, len;
    uint64_t start;
};

constexpr uint32_t kMmChunk = 1024;

// Walk pass: Find writes its values; Count and Retrieve write each query's
// match count at counts[base + i].
__global__ __launch_bounds__(256) void mm_batch_walk_kernel(const MmBatchItem* items, const MmBatchChunk* chunks,
                                                            uint32_t nchunks, uint64_t* counts)
{
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const MmBatchChunk ch = chunks[c];
        const MmBatchItem it = items[ch.item];
        for (uint32_t j = threadIdx.x; j < ch.len; j += blockDim.x) {
            const uint64_t i = ch.start + j;
            const uint64_t key = load_bits<8>(it.keys + i * 8);
            int64_t best = 0;
            bool any = false;
            if (it.mode == Q_FIND) {
                if (key != it.ek)
                    (void)mm_walk<8, 8, Q_FIND>(it.table, it.mask, key, it.ek, best, any, 0, 0, 0, nullptr, nullptr);
                store_bits<8>(it.out + i * 8, any ? (uint64_t)best : it.ev);
            } else {
                uint64_t m = 0;
                if (key != it.ek)
                    m = mm_walk<8, 8, Q_COUNT>(it.table, it.mask, key, it.ek, best, any, 0, 0, 0, nullptr, nullptr);
                counts[it.base + i] = m;
            }
        }
    }
}

// res[4 r] = matches of item r (Count / Retrieve); res[4 r + 1] = 0, the
// long-segment count the gather raises.
__global__ __launch_bounds__(256) void mm_batch_totals_kernel(const MmBatchItem* items, uint32_t nitems,
                                                              const uint64_t* offsets, uint64_t* res)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nitems) return;
    const MmBatchItem it = items[r];
    res[4 * r] = it.mode == Q_FIND ? 0 : offsets[it.base + it.n] - offsets[it.base];
    res[4 * r + 1] = 0;
}

// Gather pass of the Retrieves (chunks of Retrieve items only): matches at
// their offsets within the item's own output; an item whose matches exceed
// its capacity writes nothing.  Segments longer than kInlineSort land
// unsorted and are listed: begins at longs[2 base + k], ends at
// longs[2 base + n + k].
__global__ __launch_bounds__(256) void mm_batch_gather_kernel(const MmBatchItem* items, const MmBatchChunk* chunks,
                                                              uint32_t nchunks, const uint64_t* offsets,
                                                              unsigned long long* res, uint64_t* longs)
{
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const MmBatchChunk ch = chunks[c];
        const MmBatchItem it = items[ch.item];
        const uint64_t first = offsets[it.base];
        if (offsets[it.base + it.n] - first > it.capacity) continue;
        for (uint32_t j = threadIdx.x; j < ch.len; j += blockDim.x) {
            const uint64_t i = ch.start + j;
            const uint64_t key = load_bits<8>(it.keys + i * 8);
            const uint64_t seg_off = offsets[it.base + i] - first;
            const uint64_t seg_len = offsets[it.base + i + 1] - offsets[it.base + i];
            if (seg_len == 0) continue;
            if (seg_len > kInlineSort) {
                const unsigned long long k = atomicAdd(&res[4 * ch.item + 1], 1ull);
                longs[2 * it.base + k] = seg_off;
                longs[2 * it.base + it.n + k] = seg_off + seg_len;
            }
            int64_t best = 0;
            bool any = false;
            (void)mm_walk<8, 8, Q_GATHER>(it.table, it.mask, key, it.ek, best, any, seg_off, seg_off, seg_len,
                                          it.out, it.out2);
        }
    }
}

// A device or pinned buffer that grows by doubling and keeps what it outgrew.
struct Grow {
    void* p = nullptr;
    size_t cap = 0;
    bool pinned = false;
    std::vector<void*> retired;
    hipError_t need(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        const size_t want = std::max<size_t>({bytes, 2 * cap, 4096});
        void* q = nullptr;
        hipError_t e = pinned ? hipHostMalloc(&q, want, 0) : hipMalloc(&q, want);
        if (e != hipSuccess) return e;
        if (p) retired.push_back(p);
        p = q;
        cap = want;
        return hipSuccess;
    }
    void release()
    {
        for (void* q : retired) (void)(pinned ? hipHostFree(q) : hipFree(q));
        if (p) (void)(pinned ? hipHostFree(p) : hipFree(p));
        retired.clear();
        p = nullptr;
        cap = 0;
    }
};

}  // namespace

struct MmBatchWs {
    Grow host_up{nullptr, 0, true, {}};  // items + walk chunks + gather chunks
    Grow host_res{nullptr, 0, true, {}};
    Grow dev_up, counts, offsets, longs, res, scan_tmp, sort_tmp, sort_copy;
};

MmBatchWs* mm_batch_ws_create() { return new (std::nothrow) MmBatchWs(); }

mxco_status mm_batch_ws_reserve(MmBatchWs* ws, uint32_t reads, uint64_t keys)
{
    const uint64_t chunks = 2 * (keys / kMmChunk + reads);
    const size_t up_b = reads * sizeof(MmBatchItem) + chunks * sizeof(MmBatchChunk) + 32;
    MXCO_HIP(ws->host_up.need(up_b), "multimap batch: staging");
    MXCO_HIP(ws->dev_up.need(up_b), "multimap batch: upload");
    MXCO_HIP(ws->counts.need((keys + 1) * 8), "multimap batch: counts");
    MXCO_HIP(ws->offsets.need((keys + 1) * 8), "multimap batch: offsets");
    MXCO_HIP(ws->longs.need(2 * keys * 8), "multimap batch: segments");
    MXCO_HIP(ws->res.need(4ull * reads * 8), "multimap batch: results");
    MXCO_HIP(ws->host_res.need(4ull * reads * 8), "multimap batch: results");
    size_t temp = 0;
    MXCO_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                              (int64_t)keys + 1, (hipStream_t)0),
             "scan");
    MXCO_HIP(ws->scan_tmp.need(temp), "multimap batch: scan");
    return MXCO_OK;
}

void mm_batch_ws_destroy(MmBatchWs* ws)
{
    if (!ws) return;
    for (Grow* g : {&ws->host_up, &ws->host_res, &ws->dev_up, &ws->counts, &ws->offsets, &ws->longs, &ws->res,
                    &ws->scan_tmp, &ws->sort_tmp, &ws->sort_copy})
        g->release();
    delete ws;
}

mxco_status mm_batch_reads(MmRead* reads, uint32_t n, MmBatchWs* ws, hipStream_t s)
{
    if (n == 0) return MXCO_OK;
    std::vector<MmBatchItem> items(n);
    std::vector<MmBatchChunk> walk, gather;
    uint64_t L = 0;  // counts of Count and Retrieve items, concatenated
    for (uint32_t r = 0; r < n; ++r) {
        MmRead& q = reads[r];
        q.result = 0;
        q.status = MXCO_OK;
        const mxco_mm* m = q.table;
        if (!m || !m->table || m->ks != 8 || m->vs != 8) return fail(MXCO_INVALID_ARGUMENT, "multimap batch: table");
        MmBatchItem& it = items[r];
        memset(&it, 0, sizeof(it));
        it.table = m->table;
        it.keys = (const uint8_t*)q.keys;
        it.out = (uint8_t*)q.out;
        it.out2 = (uint8_t*)q.out2;
        it.mask = m->slots_pow2 - 1;
        it.ek = m->empty_key;
        it.ev = m->empty_val;
        it.n = q.keys ? q.n : 0;
        it.capacity = q.out_capacity;
        if (q.op == MXCO_REQ_MM_FIND) {
            it.mode = Q_FIND;
            if (it.n && !q.out) return fail(MXCO_INVALID_ARGUMENT, "multimap batch: null output");
        } else if (q.op == MXCO_REQ_MM_COUNT) {
            it.mode = Q_COUNT;
        } else if (q.op == MXCO_REQ_MM_RETRIEVE) {
            it.mode = Q_GATHER;
        } else {
            return fail(MXCO_INVALID_ARGUMENT, "multimap batch: not a read");
        }
        if (it.mode != Q_FIND) {
            it.base = L;
            L += it.n;
        }
        for (uint64_t st = 0; st < it.n; st += kMmChunk) {
            const MmBatchChunk ch{r, (uint32_t)std::min<uint64_t>(kMmChunk, it.n - st), st};
            walk.push_back(ch);
            if (it.mode == Q_GATHER) gather.push_back(ch);
        }
    }
    const size_t items_b = (n * sizeof(MmBatchItem) + 15) & ~size_t(15);
    const size_t walk_b = (walk.size() * sizeof(MmBatchChunk) + 15) & ~size_t(15);
    const size_t up_b = items_b + walk_b + gather.size() * sizeof(MmBatchChunk);
    MXCO_HIP(ws->host_up.need(up_b), "multimap batch: staging");
    MXCO_HIP(ws->dev_up.need(up_b), "multimap batch: upload");
    MXCO_HIP(ws->counts.need((L + 1) * 8), "multimap batch: counts");
    MXCO_HIP(ws->offsets.need((L + 1) * 8), "multimap batch: offsets");
    MXCO_HIP(ws->longs.need(std::max<uint64_t>(1, 2 * L) * 8), "multimap batch: segments");
    MXCO_HIP(ws->res.need(4ull * n * 8), "multimap batch: results");
    MXCO_HIP(ws->host_res.need(4ull * n * 8), "multimap batch: results");
    uint8_t* hu = (uint8_t*)ws->host_up.p;
    memcpy(hu, items.data(), n * sizeof(MmBatchItem));
    if (!walk.empty()) memcpy(hu + items_b, walk.data(), walk.size() * sizeof(MmBatchChunk));
    if (!gather.empty()) memcpy(hu + items_b + walk_b, gather.data(), gather.size() * sizeof(MmBatchChunk));
    uint8_t* du = (uint8_t*)ws->dev_up.p;
    const MmBatchItem* ditems = (const MmBatchItem*)du;
    const MmBatchChunk* dwalk = (const MmBatchChunk*)(du + items_b);
    const MmBatchChunk* dgather = (const MmBatchChunk*)(du + items_b + walk_b);
    uint64_t* counts = (uint64_t*)ws->counts.p;
    uint64_t* offsets = (uint64_t*)ws->offsets.p;
    uint64_t* res = (uint64_t*)ws->res.p;
    MXCO_HIP(hipMemcpyAsync(du, hu, up_b, hipMemcpyHostToDevice, s), "multimap batch: upload");
    const unsigned cap = (unsigned)cu_count() * 16;
    if (!walk.empty())
        mm_batch_walk_kernel<<<std::min<unsigned>((unsigned)walk.size(), cap), 256, 0, s>>>(
            ditems, dwalk, (uint32_t)walk.size(), counts);
    MXCO_HIP(hipGetLastError(), "multimap batch: walk");
    // offsets over the L counts: offsets[L] is the total (counts[L] is read
    // but does not change the result)
    size_t temp = 0;
    MXCO_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, counts, offsets, (int64_t)L + 1, s), "scan");
    MXCO_HIP(ws->scan_tmp.need(temp), "multimap batch: scan");
    MXCO_HIP(hipcub::DeviceScan::ExclusiveSum(ws->scan_tmp.p, temp, counts, offsets, (int64_t)L + 1, s), "scan");
    mm_batch_totals_kernel<<<(n + 255) / 256, 256, 0, s>>>(ditems, n, offsets, res);
    if (!gather.empty())
        mm_batch_gather_kernel<<<std::min<unsigned>((unsigned)gather.size(), cap), 256, 0, s>>>(
            ditems, dgather, (uint32_t)gather.size(), offsets, (unsigned long long*)res, (uint64_t*)ws->longs.p);
    MXCO_HIP(hipGetLastError(), "multimap batch: gather");
    uint64_t* hres = (uint64_t*)ws->host_res.p;
    MXCO_HIP(hipMemcpyAsync(hres, res, 4ull * n * 8, hipMemcpyDeviceToHost, s), "multimap batch: results");
    MXCO_HIP(hipStreamSynchronize(s), "multimap batch");
    bool sorted_any = false;
    for (uint32_t r = 0; r < n; ++r) {
        MmRead& q = reads[r];
        const MmBatchItem& it = items[r];
        if (it.mode == Q_FIND) continue;
        const uint64_t total = hres[4 * r];
        if (it.mode == Q_COUNT) {
            q.result = total;
            continue;
        }
        if (total > it.capacity) {
            q.status = MXCO_LENGTH_ERROR;
            continue;
        }
        if (total && (!q.out || !q.out2)) {
            q.status = MXCO_INVALID_ARGUMENT;
            continue;
        }
        q.result = total;
        const int64_t nlong = (int64_t)hres[4 * r + 1];
        if (nlong > 0) {
            // sort the listed segments from a copy of the values back into the output
            MXCO_HIP(ws->sort_copy.need(total * 8), "multimap batch: sort");
            MXCO_HIP(hipMemcpyAsync(ws->sort_copy.p, q.out2, total * 8, hipMemcpyDeviceToDevice, s), "sort");
            const uint64_t* begins = (const uint64_t*)ws->longs.p + 2 * it.base;
            const uint64_t* ends = begins + it.n;
            size_t stemp = 0;
            MXCO_HIP(hipcub::DeviceSegmentedSort::SortKeys(nullptr, stemp, (const int64_t*)ws->sort_copy.p,
                                                          (int64_t*)q.out2, (int64_t)total, nlong, begins, ends, s),
                     "multimap batch: sort");
            MXCO_HIP(ws->sort_tmp.need(stemp), "multimap batch: sort");
            MXCO_HIP(hipcub::DeviceSegmentedSort::SortKeys(ws->sort_tmp.p, stemp, (const int64_t*)ws->sort_copy.p,
                                                          (int64_t*)q.out2, (int64_t)total, nlong, begins, ends, s),
                     "multimap batch: sort");
            // the copy and the sort's scratch are reused by the next item's sort: in stream order
            sorted_any = true;
        }
    }
    if (sorted_any) MXCO_HIP(hipStreamSynchronize(s), "multimap batch: sort");
    return MXCO_OK;
}

}  // namespace mxco
