// DynamicMap (include/dynamic_map.h:26-625): a growing map made of StaticMap
// submaps, orchestrated on the host exactly as the reference does:
//   - usable room of a submap of capacity c: floor(0.6 c) - 10000 (0 below);
//   - growth appends a submap of the next capacity (doubling from twice the
//     first submap), doubled again until its room covers max(2 * size, n);
//   - Insert deduplicates against older non-empty submaps (Contains mask,
//     then InsertIf on the active submap for the keys absent elsewhere);
//   - Find / Contains query every non-empty submap and merge on the device
//     (a key lives in at most one submap); Erase visits every non-empty
//     submap; submaps use the erased key as tombstone.
// The reference prepares the next submap ahead of need (StartPrepareNext,
// dynamic_map.h:470-486) and clears it on a second stream; the same rule
// decides here when that submap is made and what capacity it gets, so
// Capacity() and NumSubmaps() follow the reference, but it is cleared at
// once on the caller's stream.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "mxco_common.h"

using namespace mxco;

namespace {

constexpr uint64_t kMinInsert = 10000;  // DEFAULT_MIN_INSERT (dynamic_map.h:39)
constexpr float kMaxLoad = 0.60f;       // DEFAULT_MAX_LOAD (dynamic_map.h:40)
constexpr uint64_t kGrowth = 2;         // GROWTH_FACTOR
constexpr uint64_t kSubBatch = 800000;  // INSERT_SUBBATCH

__global__ __launch_bounds__(256) void dm_extract_keys_kernel(uint8_t* keys, const uint8_t* pairs, uint64_t n,
                                                              uint32_t ks, uint32_t ss)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (ks == 4) ((uint32_t*)keys)[i] = *(const uint32_t*)(pairs + i * ss);
        else ((uint64_t*)keys)[i] = *(const uint64_t*)(pairs + i * ss);
    }
}

__global__ __launch_bounds__(256) void dm_count_zero_kernel(const uint8_t* mask, uint64_t n, unsigned long long* out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long z = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) z += mask[i] == 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) z += __shfl_xor(z, o);
    if (__lane_id() == 0 && z) atomicAdd(out, z);
}

__global__ void dm_zero_kernel(unsigned long long* p)
{
    *p = 0;
}

// out[i] = tmp[i] where out[i] is the empty value (MergeFind,
// detail/dynamic_map/kernels.h:74-97: == for floating payloads, plus a
// bitwise match for a NaN sentinel).
template <int VS>
__global__ __launch_bounds__(256) void dm_merge_find_kernel(uint8_t* out, const uint8_t* tmp, uint64_t empty,
                                                            uint32_t is_float, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        bool is_empty;
        if constexpr (VS == 4) {
            const uint32_t cur = ((const uint32_t*)out)[i];
            if (is_float) {
                const float c = __uint_as_float(cur), e = __uint_as_float((uint32_t)empty);
                is_empty = c == e || (e != e && cur == (uint32_t)empty);
            } else {
                is_empty = cur == (uint32_t)empty;
            }
            if (is_empty) ((uint32_t*)out)[i] = ((const uint32_t*)tmp)[i];
        } else {
            const uint64_t cur = ((const uint64_t*)out)[i];
            if (is_float) {
                const double c = __longlong_as_double((long long)cur), e = __longlong_as_double((long long)empty);
                is_empty = c == e || (e != e && cur == empty);
            } else {
                is_empty = cur == empty;
            }
            if (is_empty) ((uint64_t*)out)[i] = ((const uint64_t*)tmp)[i];
        }
    }
}

__global__ __launch_bounds__(256) void dm_merge_contains_kernel(uint8_t* out, const uint8_t* tmp, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        if (tmp[i]) out[i] = 1;
}

}  // namespace

struct mxco_dm {
    mxco_oa_desc d;
    uint32_t ss;  // slot (pair) size
    uint8_t empty_key[8], empty_val[8], erased_key[8];
    uint64_t empty_val_bits;
    std::vector<mxco_oa*> subs;
    std::vector<uint64_t> sub_size;
    uint64_t size = 0, total_cap = 0, next_cap = 0;
    mxco_oa* next_sub = nullptr;  // prepared next submap (not yet active)
    uint64_t next_sub_cap = 0;    // its requested capacity
    uint8_t* scratch = nullptr;  // Find / Contains merge input
    size_t scratch_bytes = 0;
    uint8_t* dedup = nullptr;  // keys | mask | tmp
    size_t dedup_bytes = 0;
    unsigned long long* counter = nullptr;
    unsigned long long* counter_host = nullptr;  // pinned
};

namespace {

uint64_t usable(uint64_t cap)
{
    const uint64_t u = (uint64_t)floor((double)kMaxLoad * (double)cap);
    return u > kMinInsert ? u - kMinInsert : 0;
}

uint64_t capacity_of(const mxco_oa* t)
{
    uint64_t c = 0;
    (void)mxco_oa_capacity(t, &c);
    return c;
}

mxco_status make_submap(mxco_dm* m, uint64_t cap, hipStream_t s, mxco_oa** out)
{
    mxco_oa* t = nullptr;
    if (mxco_status e = mxco_oa_create(&m->d, cap, m->empty_key, m->empty_val, s, &t)) return e;
    if (mxco_status e = mxco_oa_set_erased_key(t, m->erased_key, s)) {
        (void)mxco_oa_destroy(t);
        return e;
    }
    *out = t;
    return MXCO_OK;
}

void push_submap(mxco_dm* m, mxco_oa* t)
{
    m->subs.push_back(t);
    m->sub_size.push_back(0);
    m->total_cap += capacity_of(t);
}

mxco_status add_submap(mxco_dm* m, uint64_t cap, hipStream_t s)
{
    mxco_oa* t = nullptr;
    if (mxco_status e = make_submap(m, cap, s, &t)) return e;
    push_submap(m, t);
    return MXCO_OK;
}

void discard_prepared(mxco_dm* m)
{
    if (m->next_sub) (void)mxco_oa_destroy(m->next_sub);
    m->next_sub = nullptr;
    m->next_sub_cap = 0;
}

size_t active(const mxco_dm* m)
{
    return m->subs.size() - 1;
}

uint64_t active_room(const mxco_dm* m)
{
    const size_t a = active(m);
    const uint64_t u = usable(capacity_of(m->subs[a]));
    return u > m->sub_size[a] ? u - m->sub_size[a] : 0;
}

uint64_t grow_target(const mxco_dm* m, uint64_t min_usable)
{
    const uint64_t by_growth = m->size * kGrowth;
    return by_growth > min_usable ? by_growth : min_usable;
}

// StartPrepareNext (dynamic_map.h:470-486): the next submap, sized for
// GrowTarget(min_usable), made before it is needed.
mxco_status start_prepare_next(mxco_dm* m, uint64_t min_usable, hipStream_t s)
{
    if (m->next_sub) return MXCO_OK;
    uint64_t cap = m->next_cap;
    while (usable(cap) < grow_target(m, min_usable)) cap *= kGrowth;
    if (mxco_status e = make_submap(m, cap, s, &m->next_sub)) return e;
    m->next_sub_cap = cap;
    return MXCO_OK;
}

// ActivateNext (dynamic_map.h:488-510): the prepared submap when its room
// suffices, else a new one of the next capacity doubled until it does.
mxco_status activate_next(mxco_dm* m, uint64_t min_usable, hipStream_t s)
{
    if (m->next_sub && usable(m->next_sub_cap) >= min_usable) {
        push_submap(m, m->next_sub);
        m->next_cap = m->next_sub_cap * kGrowth;
        m->next_sub = nullptr;
        m->next_sub_cap = 0;
        return MXCO_OK;
    }
    discard_prepared(m);
    uint64_t cap = m->next_cap;
    while (usable(cap) < min_usable) cap *= kGrowth;
    if (mxco_status e = add_submap(m, cap, s)) return e;
    m->next_cap = cap * kGrowth;
    return MXCO_OK;
}

mxco_status ensure_active_room(mxco_dm* m, uint64_t n, hipStream_t s)
{
    if (active_room(m) >= n) return MXCO_OK;
    return activate_next(m, grow_target(m, n), s);
}

bool nonempty_before(const mxco_dm* m, size_t a)
{
    for (size_t i = 0; i < a; ++i)
        if (m->sub_size[i]) return true;
    return false;
}

size_t first_nonempty(const mxco_dm* m)
{
    for (size_t i = 0; i < m->subs.size(); ++i)
        if (m->sub_size[i]) return i;
    return m->subs.size();
}

mxco_status grow_buffer(uint8_t** p, size_t* have, size_t want, const char* what)
{
    if (*have >= want && *p) return MXCO_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *have = 0;
    MXCO_HIP(hipMalloc((void**)p, want), what);
    *have = want;
    return MXCO_OK;
}

size_t round64(size_t b)
{
    return (b + 63) / 64 * 64;
}

// mask[i] = key of pair i is in an older non-empty submap; returns the number
// of zero mask bytes (the keys the active submap may take).
mxco_status contains_mask(mxco_dm* m, const void* pairs, uint64_t n, hipStream_t s, uint8_t** mask_out,
                          uint64_t* zeros)
{
    const size_t kb = round64(n * m->d.key_size), mb = round64(n);
    if (mxco_status e = grow_buffer(&m->dedup, &m->dedup_bytes, kb + 2 * mb, "DynamicMap: dedup scratch")) return e;
    uint8_t* keys = m->dedup;
    uint8_t* mask = keys + kb;
    uint8_t* tmp = mask + mb;
    const unsigned grid = grid_for(n, 8);
    dm_extract_keys_kernel<<<grid, 256, 0, s>>>(keys, (const uint8_t*)pairs, n, m->d.key_size, m->ss);
    MXCO_HIP(hipGetLastError(), "DynamicMap: extract keys");
    bool first = true;
    for (size_t i = 0; i < active(m); ++i) {
        if (!m->sub_size[i]) continue;
        if (mxco_status e = mxco_oa_contains(m->subs[i], keys, nullptr, first ? mask : tmp, n, s, 0)) return e;
        if (!first) {
            dm_merge_contains_kernel<<<grid, 256, 0, s>>>(mask, tmp, n);
            MXCO_HIP(hipGetLastError(), "DynamicMap: merge mask");
        }
        first = false;
    }
    dm_zero_kernel<<<1, 1, 0, s>>>(m->counter);
    dm_count_zero_kernel<<<grid, 256, 0, s>>>(mask, n, m->counter);
    MXCO_HIP(hipGetLastError(), "DynamicMap: count");
    MXCO_HIP(hipMemcpyAsync(m->counter_host, m->counter, 8, hipMemcpyDeviceToHost, s), "DynamicMap: count");
    MXCO_HIP(hipStreamSynchronize(s), "DynamicMap: count");
    *zeros = *m->counter_host;
    *mask_out = mask;
    return MXCO_OK;
}

// InsertChunk (dynamic_map.h:531-553).
mxco_status insert_chunk(mxco_dm* m, const uint8_t* pairs, uint64_t n, hipStream_t s, uint64_t* inserted)
{
    const size_t a = active(m);
    uint64_t ins;
    if (!nonempty_before(m, a)) {
        uint64_t failed = 0;
        if (mxco_status e = mxco_oa_insert(m->subs[a], pairs, nullptr, n, s, 1, &failed)) return e;
        ins = n - failed;
    } else {
        uint8_t* mask;
        uint64_t cand = 0;
        if (mxco_status e = contains_mask(m, pairs, n, s, &mask, &cand)) return e;
        const mxco_stencil zero{mask, 1, MXCO_PRED_EVEN};  // mask byte 0 or 1: even = 0
        uint64_t failed = 0;
        if (mxco_status e = mxco_oa_insert(m->subs[a], pairs, &zero, n, s, 1, &failed)) return e;
        ins = cand - failed;
    }
    m->sub_size[a] += ins;
    m->size += ins;
    *inserted = ins;
    return MXCO_OK;
}

// InsertOrAssignChunk (dynamic_map.h:555-585).
mxco_status insert_or_assign_chunk(mxco_dm* m, const uint8_t* pairs, uint64_t n, hipStream_t s, uint64_t* inserted)
{
    const size_t a = active(m);
    uint64_t ins;
    if (!nonempty_before(m, a)) {
        uint64_t failed = 0, assigned = 0;
        if (mxco_status e = mxco_oa_insert_or_assign(m->subs[a], pairs, n, s, 1, &failed, &assigned)) return e;
        ins = n - failed - assigned;
    } else {
        uint8_t* mask;
        uint64_t cand = 0;
        if (mxco_status e = contains_mask(m, pairs, n, s, &mask, &cand)) return e;
        const mxco_stencil hit{mask, 1, MXCO_PRED_NONZERO};
        const mxco_stencil zero{mask, 1, MXCO_PRED_EVEN};
        for (size_t i = 0; i < a; ++i) {
            if (!m->sub_size[i]) continue;
            if (mxco_status e = mxco_oa_assign_if(m->subs[i], pairs, &hit, n, s, 0)) return e;
        }
        if (mxco_status e = mxco_oa_assign_if(m->subs[a], pairs, &zero, n, s, 0)) return e;
        uint64_t failed = 0;
        if (mxco_status e = mxco_oa_insert(m->subs[a], pairs, &zero, n, s, 1, &failed)) return e;
        ins = cand - failed;
    }
    m->sub_size[a] += ins;
    m->size += ins;
    *inserted = ins;
    return MXCO_OK;
}

mxco_status check_dm(const mxco_dm* m)
{
    if (!m || m->subs.empty()) return fail(MXCO_LOGIC_ERROR, "DynamicMap: moved-from object");
    return MXCO_OK;
}

}  // namespace

extern "C" {

MXCO_API mxco_status mxco_dm_create(const mxco_oa_desc* desc, uint64_t initial_capacity, const void* empty_key,
                                    const void* empty_value, const void* erased_key, hipStream_t stream, mxco_dm** out)
{
    if (!out || !desc || !empty_key || !empty_value || !erased_key)
        return fail(MXCO_INVALID_ARGUMENT, "DynamicMap: null argument");
    *out = nullptr;
    if (desc->value_size == 0) return fail(MXCO_UNSUPPORTED, "DynamicMap: maps only (value_size 4 or 8)");
    if ((desc->key_size != 4 && desc->key_size != 8) || (desc->value_size != 4 && desc->value_size != 8))
        return fail(MXCO_UNSUPPORTED, "DynamicMap: keys and payloads of 4 or 8 bytes");
    if (memcmp(empty_key, erased_key, desc->key_size) == 0)
        return fail(MXCO_INVALID_ARGUMENT, "DynamicMap: emptyKey and erasedKey must differ");
    mxco_dm* m = new (std::nothrow) mxco_dm();
    if (!m) return fail(MXCO_OUT_OF_MEMORY, "DynamicMap: host allocation");
    m->d = *desc;
    m->ss = (desc->key_size == 4 && desc->value_size == 4) ? 8 : 16;
    memcpy(m->empty_key, empty_key, desc->key_size);
    memcpy(m->empty_val, empty_value, desc->value_size);
    memcpy(m->erased_key, erased_key, desc->key_size);
    m->empty_val_bits = 0;
    memcpy(&m->empty_val_bits, empty_value, desc->value_size);
    hipError_t he = hipMalloc((void**)&m->counter, 8);
    if (he == hipSuccess) he = hipHostMalloc((void**)&m->counter_host, 8, 0);
    mxco_status st = he == hipSuccess ? MXCO_OK : hip_fail(he, "DynamicMap: allocation");
    if (st == MXCO_OK) st = add_submap(m, initial_capacity, stream);
    if (st != MXCO_OK) {
        (void)mxco_dm_destroy(m);
        return st;
    }
    m->next_cap = capacity_of(m->subs[0]) * kGrowth;
    *out = m;
    return ok();
}

MXCO_API mxco_status mxco_dm_destroy(mxco_dm* m)
{
    if (!m) return ok();
    for (mxco_oa* t : m->subs) (void)mxco_oa_destroy(t);
    discard_prepared(m);
    if (m->scratch) (void)hipFree(m->scratch);
    if (m->dedup) (void)hipFree(m->dedup);
    if (m->counter) (void)hipFree(m->counter);
    if (m->counter_host) (void)hipHostFree(m->counter_host);
    delete m;
    return ok();
}

MXCO_API mxco_status mxco_dm_info(const mxco_dm* m, uint64_t* size, uint64_t* capacity, uint64_t* num_submaps)
{
    if (mxco_status e = check_dm(m)) return e;
    if (size) *size = m->size;
    if (capacity) *capacity = m->total_cap;
    if (num_submaps) *num_submaps = m->subs.size();
    return ok();
}

MXCO_API mxco_status mxco_dm_reserve(mxco_dm* m, uint64_t n, hipStream_t stream)
{
    if (mxco_status e = check_dm(m)) return e;
    uint64_t remaining = n;
    size_t idx = 0;
    while (remaining > 0) {
        uint64_t cap;
        if (idx < m->subs.size()) {
            cap = capacity_of(m->subs[idx]);
        } else {
            cap = m->next_cap;
            if (mxco_status e = add_submap(m, cap, stream)) return e;
            m->next_cap = cap * kGrowth;
        }
        const uint64_t u = usable(cap);
        if (u >= remaining) break;
        remaining -= u;
        ++idx;
    }
    return ok();
}

MXCO_API mxco_status mxco_dm_insert(mxco_dm* m, const void* pairs, uint64_t n, int mode, int sync, hipStream_t stream,
                                    uint64_t* inserted)
{
    if (mxco_status e = check_dm(m)) return e;
    if (!inserted) return fail(MXCO_INVALID_ARGUMENT, "DynamicMap: null output");
    *inserted = 0;
    if (n == 0 || !pairs) return ok();
    if (mxco_status e = ensure_active_room(m, n, stream)) return e;
    const size_t a = active(m);
    const uint8_t* p = (const uint8_t*)pairs;
    if (!sync || mode == MXCO_DM_APPEND_UNIQUE) {
        // optimistic count: InsertAsync, or unique keys appended in sub-batches
        if (!sync && active_room(m) < n * 2)
            if (mxco_status e = start_prepare_next(m, n, stream)) return e;
        const uint64_t step = sync ? kSubBatch : n;
        for (uint64_t off = 0; off < n; off += step) {
            const uint64_t cur = std::min(step, n - off);
            if (mxco_status e = mxco_oa_insert(m->subs[a], p + off * m->ss, nullptr, cur, stream, 0, nullptr)) return e;
        }
        m->sub_size[a] += n;
        m->size += n;
        *inserted = n;
        return ok();
    }
    if (!nonempty_before(m, a)) {
        uint64_t failed = 0;
        if (mxco_status e = mxco_oa_insert(m->subs[a], p, nullptr, n, stream, 1, &failed)) return e;
        const uint64_t ins = n >= failed ? n - failed : 0;
        m->sub_size[a] += ins;
        m->size += ins;
        *inserted = ins;
        return ok();
    }
    uint64_t total = 0, off = 0;
    while (off < n) {
        const uint64_t room = active_room(m);
        if (room == 0) {
            if (mxco_status e = activate_next(m, grow_target(m, n - off), stream)) return e;
            continue;
        }
        if (room < (n - off) * 2)
            if (mxco_status e = start_prepare_next(m, n - off, stream)) return e;
        const uint64_t chunk = std::min(n - off, room);
        uint64_t ins = 0;
        if (mxco_status e = insert_chunk(m, p + off * m->ss, chunk, stream, &ins)) return e;
        total += ins;
        off += chunk;
    }
    *inserted = total;
    return ok();
}

MXCO_API mxco_status mxco_dm_insert_or_assign(mxco_dm* m, const void* pairs, uint64_t n, hipStream_t stream,
                                              uint64_t* inserted)
{
    if (mxco_status e = check_dm(m)) return e;
    if (!inserted) return fail(MXCO_INVALID_ARGUMENT, "DynamicMap: null output");
    *inserted = 0;
    if (n == 0 || !pairs) return ok();
    if (mxco_status e = ensure_active_room(m, n, stream)) return e;
    const uint8_t* p = (const uint8_t*)pairs;
    uint64_t total = 0, off = 0;
    while (off < n) {
        const uint64_t room = active_room(m);
        if (room == 0) {
            if (mxco_status e = activate_next(m, grow_target(m, n - off), stream)) return e;
            continue;
        }
        if (room < (n - off) * 2)
            if (mxco_status e = start_prepare_next(m, n - off, stream)) return e;
        const uint64_t chunk = std::min(n - off, room);
        uint64_t ins = 0;
        if (mxco_status e = insert_or_assign_chunk(m, p + off * m->ss, chunk, stream, &ins)) return e;
        total += ins;
        off += chunk;
    }
    *inserted = total;
    return ok();
}

MXCO_API mxco_status mxco_dm_find(mxco_dm* m, const void* keys, void* out, uint64_t n, hipStream_t stream)
{
    if (mxco_status e = check_dm(m)) return e;
    if (n == 0 || !keys || !out) return ok();
    const size_t base = first_nonempty(m);
    if (base == m->subs.size()) {
        if (mxco_status e = mxco_oa_find(m->subs[0], keys, nullptr, out, n, stream, 1)) return e;
        return ok();
    }
    if (mxco_status e = mxco_oa_find(m->subs[base], keys, nullptr, out, n, stream, 0)) return e;
    const unsigned grid = grid_for(n, 8);
    for (size_t i = base + 1; i < m->subs.size(); ++i) {
        if (!m->sub_size[i]) continue;
        if (mxco_status e = grow_buffer(&m->scratch, &m->scratch_bytes, n * m->d.value_size, "DynamicMap: scratch"))
            return e;
        if (mxco_status e = mxco_oa_find(m->subs[i], keys, nullptr, m->scratch, n, stream, 0)) return e;
        if (m->d.value_size == 4)
            dm_merge_find_kernel<4><<<grid, 256, 0, stream>>>((uint8_t*)out, m->scratch, m->empty_val_bits,
                                                              m->d.value_kind == MXCO_KEY_FLOAT, n);
        else
            dm_merge_find_kernel<8><<<grid, 256, 0, stream>>>((uint8_t*)out, m->scratch, m->empty_val_bits,
                                                              m->d.value_kind == MXCO_KEY_FLOAT, n);
        MXCO_HIP(hipGetLastError(), "DynamicMap: merge find");
    }
    MXCO_HIP(hipStreamSynchronize(stream), "DynamicMap: find");
    return ok();
}

MXCO_API mxco_status mxco_dm_contains(mxco_dm* m, const void* keys, uint8_t* out, uint64_t n, hipStream_t stream)
{
    if (mxco_status e = check_dm(m)) return e;
    if (n == 0 || !keys || !out) return ok();
    const size_t base = first_nonempty(m);
    if (base == m->subs.size()) {
        if (mxco_status e = mxco_oa_contains(m->subs[0], keys, nullptr, out, n, stream, 1)) return e;
        return ok();
    }
    if (mxco_status e = mxco_oa_contains(m->subs[base], keys, nullptr, out, n, stream, 0)) return e;
    const unsigned grid = grid_for(n, 8);
    for (size_t i = base + 1; i < m->subs.size(); ++i) {
        if (!m->sub_size[i]) continue;
        if (mxco_status e = grow_buffer(&m->scratch, &m->scratch_bytes, n, "DynamicMap: scratch")) return e;
        if (mxco_status e = mxco_oa_contains(m->subs[i], keys, nullptr, m->scratch, n, stream, 0)) return e;
        dm_merge_contains_kernel<<<grid, 256, 0, stream>>>(out, m->scratch, n);
        MXCO_HIP(hipGetLastError(), "DynamicMap: merge contains");
    }
    MXCO_HIP(hipStreamSynchronize(stream), "DynamicMap: contains");
    return ok();
}

MXCO_API mxco_status mxco_dm_erase(mxco_dm* m, const void* keys, uint64_t n, int sync, hipStream_t stream,
                                   uint64_t* erased)
{
    if (mxco_status e = check_dm(m)) return e;
    if (erased) *erased = 0;
    if (n == 0 || !keys) return ok();
    uint64_t total = 0;
    for (size_t i = 0; i < m->subs.size(); ++i) {
        if (!m->sub_size[i]) continue;
        if (!sync) {
            if (mxco_status e = mxco_oa_erase(m->subs[i], keys, n, stream, 0, nullptr)) return e;
            continue;
        }
        uint64_t failed = 0;
        if (mxco_status e = mxco_oa_erase(m->subs[i], keys, n, stream, 1, &failed)) return e;
        const uint64_t del = n - failed;
        m->sub_size[i] -= std::min(del, m->sub_size[i]);
        total += del;
    }
    if (sync) m->size -= std::min(total, m->size);
    if (erased) *erased = total;
    return ok();
}

MXCO_API mxco_status mxco_dm_clear(mxco_dm* m, hipStream_t stream)
{
    if (mxco_status e = check_dm(m)) return e;
    discard_prepared(m);
    for (mxco_oa* t : m->subs)
        if (mxco_status e = mxco_oa_clear(t, stream, 1)) return e;
    for (uint64_t& s : m->sub_size) s = 0;
    m->size = 0;
    return ok();
}

MXCO_API mxco_status mxco_dm_submap(const mxco_dm* m, uint32_t index, mxco_oa** submap)
{
    if (mxco_status e = check_dm(m)) return e;
    if (!submap || index >= m->subs.size()) return fail(MXCO_INVALID_ARGUMENT, "DynamicMap: bad submap index");
    *submap = m->subs[index];
    return ok();
}

}  // extern "C"
