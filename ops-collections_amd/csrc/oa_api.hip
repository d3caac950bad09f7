// C ABI of StaticMap / StaticSet (include/mxco.h, mxco_oa_*).
//
// Host semantics follow OpenAddressingImpl
// (include/detail/open_addressing/open_addressing_impl.h:58-559): zero-sized
// batches are no-ops, null inputs count as failures on the synchronous
// counting calls and are ignored by the asynchronous ones, oversized batches
// raise a length error, and a synchronous call is the asynchronous call plus
// a stream synchronisation and a read-back of the per-object counters.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>

#include "batch_build.h"
#include "mxco_common.h"
#include "oa_launch.h"

using namespace mxco;

// Workspace of the region-built Insert, allocated on first use and grown
// with the batch (batch_build.h): the batch's pairs grouped by first-level
// bin (part1: nbins1 bins of G slices of build_cap_slice pairs), then by
// region (part2: nreg regions of build_cap_region pairs), and the pairs that
// find their slice or region full (spill, up to the whole batch).
struct RegionWs {
    uint64_t part1_pairs, part2_pairs, spill_pairs;  // what the buffers hold
    uint8_t* part1;
    uint8_t* part2;
    uint8_t* spill;
    uint32_t* fill;      // nbins1 * kBuildMaxSlices slice counts, then the spill count
    uint32_t* offsets2;  // nreg: region starts in part2
    uint32_t* ovc;       // nreg
    uint32_t* apply_end; // nreg
    BatchTable* table;   // the table's descriptor
    uint32_t* reg0;      // {0, nbins1}
    BatchTable table_host;
    uint32_t reg0_host[2];
    uint32_t nreg, nbins1, group;
    // A captured region build keeps part1 / part2 in its kernel arguments,
    // so once one is captured the buffers are never reallocated: a larger
    // batch takes the CAS walk instead.
    bool captured;
    // Eager region builds on different streams share the workspace: each
    // one records `done` on its stream and the next one, on another
    // stream, waits for it.
    hipEvent_t done;
    hipStream_t last_stream;
    bool has_last;
};

struct mxco_oa {
    mxco_oa_desc d;
    uint64_t cap;
    uint32_t ss;
    uint32_t voff;
    uint8_t* slots;
    unsigned long long* counters;       // device: kCounterSlots pairs [failed, assigned]
    unsigned long long* host_counters;  // pinned mirror
    uint64_t empty_key;
    uint64_t empty_val;
    uint64_t erased_key;
    bool has_tomb;
    int device;
    int group;
    unsigned blocks_per_cu;
    uint32_t store_mode;
    int inflight;
    uint32_t probe_flags;
    // Probe kernels compiled in the caller's translation unit (its own
    // KeyEqual / hasher types); null: the library's kernels.
    mxco_oa_probe_launcher caller_launch;
    RegionWs* rws;
    // No key can be in the slots: set by creation and by an eager Clear,
    // cleared by every call that can add keys (Insert, InsertOrAssign,
    // InsertAndFind).  The region build then starts its regions empty
    // instead of reading them.  A captured Clear leaves it unset: the fill
    // has not run when the capture returns.
    mutable bool empty;
    // Data() has handed out the slot pointer, through which device code may
    // write at any later time (also after a Clear): from then on the table
    // is never taken to be empty.
    mutable bool exposed;
};

namespace {

uint64_t load_le(const void* p, uint32_t n)
{
    uint64_t v = 0;
    memcpy(&v, p, n);
    return v;
}

// Pair<K, V> (include/pair.h:41): the payload at its natural alignment after
// the key, the slot aligned to min(16, bit_ceil(sK + sV)).  Sets: the key.
void pair_layout(uint32_t ks, uint32_t vs, uint32_t* ss, uint32_t* voff)
{
    if (vs == 0) {
        *ss = ks;
        *voff = 0;
        return;
    }
    uint32_t align = 1;
    while (align < ks + vs) align <<= 1;
    *ss = std::min<uint32_t>(align, 16);
    *voff = std::max(ks, vs);
}

bool wide_layout(const mxco_oa* t)
{
    return t->d.key_size >= 4 && (t->d.value_size == 0 || t->d.value_size >= 4);
}

// Lanes per key when the descriptor leaves it to the library: 4, the best
// width for both Find and Insert in the group-width sweep on MI355X
// (bench.py --sweep).
int default_group(uint32_t ss)
{
    int g = env_int("MXCO_OA_GROUP", 0);
    if (g == 1 || g == 2 || g == 4 || g == 8) return g;
    return 4;
}

OaArgs base_args(const mxco_oa* t)
{
    OaArgs a;
    memset(&a, 0, sizeof(a));
    a.table = t->slots;
    a.cap = (uint32_t)t->cap;
    a.bucket = t->d.bucket_size;
    a.nbuckets = (uint32_t)(t->cap / t->d.bucket_size);
    a.probing = t->d.probing;
    a.hash = t->d.hash;
    a.key_float = t->d.key_kind == MXCO_KEY_FLOAT;
    a.empty_key = t->empty_key;
    a.empty_val = t->empty_val;
    a.erased_key = t->has_tomb ? t->erased_key : t->empty_key;
    a.has_tombstone = t->has_tomb;
    a.store_mode = t->store_mode;
    a.probe_flags = t->probe_flags;
    return a;
}

mxco_status check_table(const mxco_oa* t, const char* fn)
{
    if (!t || !t->slots) return fail(MXCO_LOGIC_ERROR, std::string(fn) + ": moved-from or null container");
    return MXCO_OK;
}

mxco_status check_batch(uint64_t n, const char* fn)
{
    if (n > max_batch())
        return fail(MXCO_LENGTH_ERROR, std::string(fn) + ": batch size exceeds the kernel index range");
    return MXCO_OK;
}

mxco_status set_stencil(OaArgs& a, const mxco_stencil* st, const char* fn)
{
    if (!st || st->pred == MXCO_PRED_ALWAYS) return MXCO_OK;
    if (st->elem_size != 1 && st->elem_size != 4 && st->elem_size != 8)
        return fail(MXCO_INVALID_ARGUMENT, std::string(fn) + ": stencil element size must be 1, 4 or 8");
    if (st->pred > MXCO_PRED_EVEN) return fail(MXCO_INVALID_ARGUMENT, std::string(fn) + ": unknown predicate");
    a.stencil = (const uint8_t*)st->data;
    a.stencil_size = st->elem_size;
    a.stencil_pred = st->pred;
    return MXCO_OK;
}

mxco_status launch(const mxco_oa* t, int op, OaArgs& a, hipStream_t stream, const char* fn)
{
    if (t->caller_launch) {
        const unsigned grid = grid_for(a.n * (uint64_t)kCallerGroup, t->blocks_per_cu);
        const hipError_t e = (hipError_t)t->caller_launch(op, &a, grid, 256, stream);
        if (e != hipSuccess) return hip_fail(e, fn);
        return MXCO_OK;
    }
    const unsigned grid = grid_for(a.n * (uint64_t)t->group, t->blocks_per_cu);
    hipError_t e = launch_oa(t->d.key_size, t->d.value_size, t->group, op, a, dim3(grid), stream, t->inflight);
    if (e != hipSuccess) return hip_fail(e, fn);
    return MXCO_OK;
}

mxco_status zero_counters(mxco_oa* t, hipStream_t stream)
{
    zero_counters_kernel<<<1, 128, 0, stream>>>(t->counters, 2 * kCounterSlots);
    MXCO_HIP(hipGetLastError(), "zero counters");
    return MXCO_OK;
}

mxco_status read_counters(mxco_oa* t, hipStream_t stream, uint64_t* c0, uint64_t* c1)
{
    MXCO_HIP(hipMemcpyAsync(t->host_counters, t->counters, 2 * kCounterSlots * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost,
                            stream),
             "read counters");
    MXCO_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
    uint64_t s0 = 0, s1 = 0;
    for (uint32_t i = 0; i < kCounterSlots; ++i) {
        s0 += t->host_counters[2 * i];
        s1 += t->host_counters[2 * i + 1];
    }
    if (c0) *c0 = s0;
    if (c1) *c1 = s1;
    return MXCO_OK;
}

mxco_status fill(mxco_oa* t, hipStream_t stream)
{
    uint8_t img[16];
    memset(img, 0, sizeof(img));
    for (uint32_t off = 0; off < 16; off += t->ss) {
        memcpy(img + off, &t->empty_key, t->d.key_size);
        if (t->d.value_size) memcpy(img + off + t->voff, &t->empty_val, t->d.value_size);
    }
    u64x2 pat;
    memcpy(&pat, img, 16);
    const uint64_t bytes = t->cap * t->ss;
    const unsigned grid = std::max(1u, grid_for(bytes / 16, kFillBlocksPerCu));
    oa_fill_kernel<<<grid, 256, 0, stream>>>(t->slots, bytes, pat);
    MXCO_HIP(hipGetLastError(), "Clear");
    return MXCO_OK;
}

__global__ void zero_u32_pair_kernel(uint32_t* a, uint32_t na, uint32_t* b)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < na; i += gridDim.x * blockDim.x) a[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) b[0] = 0;
}

bool capturing(hipStream_t s)
{
    // the null stream is never captured
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (!s) return false;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) return true;  // unknown: assume captured
    return cs != hipStreamCaptureStatusNone;
}

void free_region_ws(mxco_oa* t)
{
    RegionWs* w = t->rws;
    if (!w) return;
    if (w->done) (void)hipEventDestroy(w->done);
    for (void* p : {(void*)w->part1, (void*)w->part2, (void*)w->spill, (void*)w->fill, (void*)w->offsets2,
                    (void*)w->ovc, (void*)w->apply_end, (void*)w->table, (void*)w->reg0})
        if (p) (void)hipFree(p);
    delete w;
    t->rws = nullptr;
}

// Insert of n pairs takes the region build when the batch is large next to
// the table (n >= 2^20 and n >= capacity / 2) and the table is larger than
// the 256 MiB last-level cache (MXCO_OA_REGION_INSERT_MIN_BYTES), under
// LinearProbing without a stencil; MXCO_OA_REGION_INSERT=0 keeps the CAS
// walk (both read at each call).  The build moves 80 bytes per 16-byte pair
// (read and written by the scatter and by the regroup, read by the apply)
// plus 16 per slot written (32 when the regions are first read from a table
// that may hold keys); the CAS walk about 300 per distinct key at load factor
// 0.8 on a table that misses the cache, less on one that fits it or for a
// stream of repeated keys, so those keep the walk.  Under stream capture the
// workspace must already be large enough, and once a build has been captured
// it never grows.  A captured build shares the table's workspace with every
// other Insert on the table: replays must be ordered with them by the caller,
// as for any two calls on one container.
// Regions, regions per first-level bin and first-level bins of the region
// build: about 512 bins, at most 256 regions per bin.
struct RegionShape {
    uint32_t nreg, group, nbins1;
};

RegionShape region_shape(const mxco_oa* t)
{
    const uint32_t R = build_region_slots(t->ss);
    RegionShape r;
    r.nreg = (uint32_t)((t->cap + R - 1) / R);
    r.group = std::min<uint32_t>(256u, (r.nreg + 511) / 512);
    r.nbins1 = (r.nreg + r.group - 1) / r.group;
    return r;
}

// The buffers a region build of n pairs needs, and its slice count.
struct RegionNeed {
    uint32_t slices, cap_slice, cap_region;
    uint64_t part1, part2, spill;
};

RegionNeed region_need(const mxco_oa* t, uint64_t n)
{
    const RegionShape rs = region_shape(t);
    RegionNeed q;
    q.slices = kBuildMaxSlices;
    q.cap_slice = build_cap_slice(n, rs.nbins1, q.slices);
    q.cap_region = build_cap_region(n, rs.nreg);
    q.part1 = (uint64_t)rs.nbins1 * q.slices * q.cap_slice;
    q.part2 = (uint64_t)rs.nreg * q.cap_region;
    q.spill = n;
    return q;
}

bool region_ws_holds(const RegionWs* w, const RegionNeed& q)
{
    return w && w->part1_pairs >= q.part1 && w->part2_pairs >= q.part2 && w->spill_pairs >= q.spill;
}

bool region_insert_fits(const mxco_oa* t, uint64_t n, const OaArgs& a, hipStream_t stream)
{
    const char* env = getenv("MXCO_OA_REGION_INSERT");
    if ((env && atoi(env) == 0) || t->d.probing != 0 || a.stencil || !wide_layout(t) || t->caller_launch)
        return false;
    if (n < (1ull << 20) || n * 2 < t->cap || n >= 0xffffffffull) return false;
    const char* mb = getenv("MXCO_OA_REGION_INSERT_MIN_BYTES");
    if (t->cap * t->ss < (mb ? strtoull(mb, nullptr, 10) : (256ull << 20))) return false;
    const RegionShape rs = region_shape(t);
    if (rs.nbins1 > build_bounded_max_bins(t->ss)) return false;  // the scatter's LDS
    // part1 / part2 positions are 32-bit
    const RegionNeed q = region_need(t, n);
    if (q.part1 > 0xffffffffull || q.part2 > 0xffffffffull) return false;
    // under capture the workspace must exist and be large enough; after a
    // capture it never grows
    if (capturing(stream) || (t->rws && t->rws->captured)) return region_ws_holds(t->rws, q);
    return true;
}

hipError_t region_ws(mxco_oa* t, uint64_t n, hipStream_t stream)
{
    if (!t->rws) {
        RegionWs* w = new (std::nothrow) RegionWs();
        if (!w) return hipErrorOutOfMemory;
        t->rws = w;
        // about 512 bins: runs of a line or more per bin and 4096-pair tile
        // in the scatter; at most 256 regions per bin (the regroup's LDS)
        const RegionShape rs = region_shape(t);
        w->nreg = rs.nreg;
        w->group = rs.group;
        w->nbins1 = rs.nbins1;
        hipError_t e = hipMalloc((void**)&w->fill, ((uint64_t)w->nbins1 * kBuildMaxSlices + 1) * 4);
        if (e == hipSuccess) e = hipMalloc((void**)&w->offsets2, w->nreg * 4ull);
        if (e == hipSuccess) e = hipMalloc((void**)&w->ovc, w->nreg * 4ull);
        if (e == hipSuccess) e = hipMalloc((void**)&w->apply_end, w->nreg * 4ull);
        if (e == hipSuccess) e = hipMalloc((void**)&w->table, sizeof(BatchTable));
        if (e == hipSuccess) e = hipMalloc((void**)&w->reg0, 2 * sizeof(uint32_t));
        if (e == hipSuccess) e = hipEventCreateWithFlags(&w->done, hipEventDisableTiming);
        if (e == hipSuccess) {
            // stream-ordered from host memory the workspace keeps
            w->table_host = BatchTable{t->slots, (uint32_t)t->cap, (uint32_t)(t->cap / t->d.bucket_size)};
            w->reg0_host[0] = 0u;
            w->reg0_host[1] = w->nbins1;
            e = hipMemcpyAsync(w->table, &w->table_host, sizeof(BatchTable), hipMemcpyHostToDevice, stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(w->reg0, w->reg0_host, sizeof(w->reg0_host), hipMemcpyHostToDevice, stream);
        }
        if (e != hipSuccess) {
            free_region_ws(t);
            return e;
        }
    }
    RegionWs* w = t->rws;
    const RegionNeed q = region_need(t, n);
    if (!region_ws_holds(w, q)) {
        if (w->captured) return hipErrorInvalidValue;  // a captured build holds these buffers
        // hipFree waits for the device, so no earlier build still reads them
        for (uint8_t** b : {&w->part1, &w->part2, &w->spill}) {
            if (*b) (void)hipFree(*b);
            *b = nullptr;
        }
        w->part1_pairs = w->part2_pairs = w->spill_pairs = 0;
        hipError_t e = hipMalloc((void**)&w->part1, q.part1 * t->ss);
        if (e == hipSuccess) e = hipMalloc((void**)&w->part2, q.part2 * t->ss);
        if (e == hipSuccess) e = hipMalloc((void**)&w->spill, q.spill * t->ss);
        if (e != hipSuccess) return e;
        w->part1_pairs = q.part1;
        w->part2_pairs = q.part2;
        w->spill_pairs = q.spill;
    }
    return hipSuccess;
}

template <int KS, int VS>
hipError_t launch_region_insert(const mxco_oa* t, const OaArgs& oa, hipStream_t s, bool table_empty)
{
    constexpr int SS = Lay<KS, VS>::SS;
    constexpr uint32_t R = BuildShape<SS>::kRegion;
    static const hipError_t attr = [] {
        hipError_t e = hipFuncSetAttribute((const void*)build_bounded_scatter_kernel<KS, VS>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)build_bounded_scatter_lds<SS>(build_bounded_max_bins(SS)));
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)build_bounded_regroup_kernel<KS, VS>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)build_bounded_regroup_lds<SS>(256));
        return e;
    }();
    if (attr != hipSuccess) return attr;
    const RegionWs& w = *t->rws;
    if (w.group > 256 || w.nbins1 > build_bounded_max_bins(SS)) return hipErrorInvalidValue;
    const uint64_t n = oa.n;
    const RegionNeed q = region_need(t, n);
    if (!region_ws_holds(&w, q)) return hipErrorInvalidValue;
    uint32_t* nspill = w.fill + (uint64_t)w.nbins1 * kBuildMaxSlices;
    zero_u32_pair_kernel<<<grid_for((uint64_t)w.nbins1 * q.slices, 1), 256, 0, s>>>(w.fill, w.nbins1 * q.slices,
                                                                                  nspill);
    BuildArgs a1;
    memset(&a1, 0, sizeof(a1));
    a1.in = oa.in;
    a1.part = w.part1;
    a1.tables = w.table;
    a1.reg0 = w.reg0;
    a1.counters = oa.counters;
    a1.empty_key = oa.empty_key;
    a1.empty_val = oa.empty_val;
    a1.single_n = n;
    a1.nreg = w.nbins1;
    a1.maxbins = w.nbins1;
    a1.bin_slots = w.group * R;
    a1.bucket = oa.bucket;
    a1.hash = oa.hash;
    a1.key_float = oa.key_float;
    SpillArgs sp;
    sp.fill = w.fill;
    sp.spill = w.spill;
    sp.nspill = nspill;
    sp.cap_slice = q.cap_slice;
    sp.cap_region = q.cap_region;
    sp.slices = q.slices;
    const unsigned ntiles = (unsigned)((n + BuildShape<SS>::kTile - 1) / BuildShape<SS>::kTile);
    const unsigned sgrid = std::min<unsigned>(ntiles, (unsigned)cu_count() * 2);
    build_bounded_scatter_kernel<KS, VS><<<sgrid, kBuildThreads, build_bounded_scatter_lds<SS>(w.nbins1), s>>>(a1, sp);
    BuildArgs a2 = a1;
    a2.part = w.part2;
    a2.offsets = w.offsets2;
    a2.ovc = w.ovc;
    a2.apply_end = w.apply_end;
    a2.nreg = w.nreg;
    a2.bin_slots = R;
    build_bounded_regroup_kernel<KS, VS><<<w.nbins1, kBuildThreads, build_bounded_regroup_lds<SS>(w.group), s>>>(
        a2, sp, w.part1, w.group);
    // an empty table's regions start from the empty slot
    if (table_empty)
        build_apply_kernel<KS, VS, true, false><<<w.nreg, kBuildThreads, 0, s>>>(a2);
    else
        build_apply_kernel<KS, VS, true, true><<<w.nreg, kBuildThreads, 0, s>>>(a2);
    const unsigned ogrid = std::max(1u, std::min<unsigned>((w.nreg + 3) / 4, (unsigned)cu_count() * 8));
    build_overflow_kernel<KS, VS, true><<<ogrid, 256, 0, s>>>(a2);
    build_spill_kernel<KS, VS><<<(unsigned)cu_count() * 4, 256, 0, s>>>(a2, sp);
    return hipGetLastError();
}

hipError_t region_insert(mxco_oa* t, const OaArgs& a, hipStream_t s, bool table_empty)
{
    RegionWs& w = *t->rws;
    const bool cap = capturing(s);
    // a captured Insert may be replayed on a table that holds keys by then
    if (cap) table_empty = false;
    if (!cap && w.has_last && w.last_stream != s) {
        hipError_t e = hipStreamWaitEvent(s, w.done, 0);
        if (e != hipSuccess) return e;
    }
    const uint32_t ks = t->d.key_size, vs = t->d.value_size;
    hipError_t e = hipErrorInvalidValue;
    if (ks == 4 && vs == 4) e = launch_region_insert<4, 4>(t, a, s, table_empty);
    else if (ks == 8 && vs == 8) e = launch_region_insert<8, 8>(t, a, s, table_empty);
    else if (ks == 4 && vs == 8) e = launch_region_insert<4, 8>(t, a, s, table_empty);
    else if (ks == 8 && vs == 4) e = launch_region_insert<8, 4>(t, a, s, table_empty);
    else if (ks == 4 && vs == 0) e = launch_region_insert<4, 0>(t, a, s, table_empty);
    else if (ks == 8 && vs == 0) e = launch_region_insert<8, 0>(t, a, s, table_empty);
    if (e != hipSuccess) return e;
    if (cap) {
        w.captured = true;
    } else {
        e = hipEventRecord(w.done, s);
        w.last_stream = s;
        w.has_last = true;
    }
    return e;
}

}  // namespace

extern "C" {

MXCO_API mxco_status mxco_oa_create(const mxco_oa_desc* desc, uint64_t capacity, const void* empty_key,
                                    const void* empty_value, hipStream_t stream, mxco_oa** out)
{
    if (!desc || !out || !empty_key) return fail(MXCO_INVALID_ARGUMENT, "mxco_oa_create: null argument");
    *out = nullptr;
    const mxco_oa_desc& d = *desc;
    if (d.key_size != 1 && d.key_size != 2 && d.key_size != 4 && d.key_size != 8)
        return fail(MXCO_UNSUPPORTED, "mxco_oa_create: key size must be 1, 2, 4 or 8 bytes");
    if (d.value_size != 0 && d.value_size != 1 && d.value_size != 2 && d.value_size != 4 && d.value_size != 8)
        return fail(MXCO_UNSUPPORTED, "mxco_oa_create: payload size must be 0, 1, 2, 4 or 8 bytes");
    if (d.value_size && !empty_value) return fail(MXCO_INVALID_ARGUMENT, "mxco_oa_create: null empty value");
    if (d.bucket_size < 1 || d.bucket_size > 1024)
        return fail(MXCO_UNSUPPORTED, "mxco_oa_create: bucket size must be in [1, 1024]");
    if (d.probing > MXCO_PROBING_DOUBLE) return fail(MXCO_INVALID_ARGUMENT, "mxco_oa_create: unknown probing scheme");
    if (d.hash > MXCO_HASH_FMIX_64) return fail(MXCO_INVALID_ARGUMENT, "mxco_oa_create: unknown hash");
    if ((d.hash == MXCO_HASH_FMIX_32 && d.key_size != 4) || (d.hash == MXCO_HASH_FMIX_64 && d.key_size != 8))
        return fail(MXCO_INVALID_ARGUMENT, "mxco_oa_create: fmix hash width differs from the key width");
    if (d.key_kind > MXCO_KEY_FLOAT) return fail(MXCO_INVALID_ARGUMENT, "mxco_oa_create: unknown key kind");
    if (d.key_kind == MXCO_KEY_FLOAT && d.key_size == 1)
        return fail(MXCO_UNSUPPORTED, "mxco_oa_create: floating-point keys are 2 (binary16), 4 or 8 bytes");
    if (d.group_size != 0 && d.group_size != 1 && d.group_size != 2 && d.group_size != 4 && d.group_size != 8)
        return fail(MXCO_INVALID_ARGUMENT, "mxco_oa_create: group size must be 0, 1, 2, 4 or 8");

    const uint64_t cap = valid_extent(capacity, d.bucket_size, d.probing);
    if (cap == 0) return fail(MXCO_INVALID_ARGUMENT, "Invalid input Extent");
    if (cap > 0xffffffffull)
        return fail(MXCO_INVALID_ARGUMENT,
                    "mxco_oa_create: capacity exceeds the 32-bit slot index of the device probe sequence");

    mxco_oa* t = new (std::nothrow) mxco_oa();
    if (!t) return fail(MXCO_OUT_OF_MEMORY, "mxco_oa_create: host allocation");
    t->d = d;
    t->cap = cap;
    pair_layout(d.key_size, d.value_size, &t->ss, &t->voff);
    t->empty_key = load_le(empty_key, d.key_size);
    t->empty_val = d.value_size ? load_le(empty_value, d.value_size) : 0;
    t->erased_key = t->empty_key;
    t->has_tomb = false;
    t->group = d.group_size ? (int)d.group_size : default_group(t->ss);
    t->blocks_per_cu = (unsigned)env_int("MXCO_OA_BLOCKS_PER_CU", 256);
    t->store_mode = (uint32_t)env_int("MXCO_OA_STORE_MODE", 0);
    {
        // Probe kernel variant: 4 (default) stages Find / Contains outputs per
        // wavefront; 1 plain; 2 two keys in flight per group; 3 block-staged
        // chunks.  Operations a variant does not cover use the plain kernel.
        const int m = env_int("MXCO_OA_INFLIGHT", 4);
        t->inflight = (m == 2 || m == 3 || m == 4) ? m : 1;
        if (!wide_layout(t)) t->inflight = 1;
    }
    t->probe_flags = (uint32_t)env_int("MXCO_OA_PROBE_FLAGS", 0);
    (void)hipGetDevice(&t->device);

    // rounded up to 16 bytes: the CAS of a 1- or 2-byte slot works on the
    // aligned 32-bit word that holds it
    hipError_t e = hipMalloc((void**)&t->slots, (cap * t->ss + 15) & ~15ull);
    if (e != hipSuccess) {
        delete t;
        return hip_fail(e, "mxco_oa_create: slot array");
    }
    e = hipMalloc((void**)&t->counters, 2 * kCounterSlots * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipHostMalloc((void**)&t->host_counters, 2 * kCounterSlots * sizeof(unsigned long long), 0);
    if (e != hipSuccess) {
        (void)hipFree(t->slots);
        if (t->counters) (void)hipFree(t->counters);
        delete t;
        return hip_fail(e, "mxco_oa_create: counters");
    }
    mxco_status st = fill(t, stream);
    t->empty = !capturing(stream);
    if (st == MXCO_OK) {
        e = hipStreamSynchronize(stream);
        if (e != hipSuccess) st = hip_fail(e, "mxco_oa_create: clear");
    }
    if (st != MXCO_OK) {
        (void)hipFree(t->slots);
        (void)hipFree(t->counters);
        (void)hipHostFree(t->host_counters);
        delete t;
        return st;
    }
    *out = t;
    return ok();
}

MXCO_API mxco_status mxco_oa_destroy(mxco_oa* t)
{
    if (!t) return ok();
    if (t->slots) (void)hipFree(t->slots);
    if (t->counters) (void)hipFree(t->counters);
    if (t->host_counters) (void)hipHostFree(t->host_counters);
    free_region_ws(t);
    delete t;
    return ok();
}

MXCO_API mxco_status mxco_oa_capacity(const mxco_oa* t, uint64_t* capacity)
{
    if (mxco_status s = check_table(t, "Capacity")) return s;
    if (!capacity) return fail(MXCO_INVALID_ARGUMENT, "Capacity: null");
    *capacity = t->cap;
    return ok();
}

MXCO_API mxco_status mxco_oa_data(const mxco_oa* t, void** slots)
{
    if (mxco_status s = check_table(t, "Data")) return s;
    if (!slots) return fail(MXCO_INVALID_ARGUMENT, "Data: null");
    *slots = t->slots;
    t->empty = false;
    t->exposed = true;
    return ok();
}

MXCO_API mxco_status mxco_oa_slot_size(const mxco_oa* t, uint32_t* ss)
{
    if (mxco_status s = check_table(t, "slot_size")) return s;
    if (!ss) return fail(MXCO_INVALID_ARGUMENT, "slot_size: null");
    *ss = t->ss;
    return ok();
}

MXCO_API mxco_status mxco_oa_set_erased_key(mxco_oa* t, const void* erased_key, hipStream_t stream)
{
    if (mxco_status s = check_table(t, "SetErasedKey")) return s;
    if (!erased_key) return fail(MXCO_INVALID_ARGUMENT, "SetErasedKey: null key");
    t->erased_key = load_le(erased_key, t->d.key_size);
    t->has_tomb = t->erased_key != t->empty_key;
    return ok();
}

MXCO_API mxco_status mxco_oa_clear(mxco_oa* t, hipStream_t stream, int sync)
{
    if (mxco_status s = check_table(t, "Clear")) return s;
    if (mxco_status s = fill(t, stream)) return s;
    t->empty = !t->exposed && !capturing(stream);
    if (sync) MXCO_HIP(hipStreamSynchronize(stream), "Clear: hipStreamSynchronize");
    return ok();
}

MXCO_API mxco_status mxco_oa_insert(mxco_oa* t, const void* pairs, const mxco_stencil* stencil, uint64_t n,
                                    hipStream_t stream, int sync, uint64_t* num_failed)
{
    if (mxco_status s = check_table(t, "Insert")) return s;
    if (num_failed) *num_failed = 0;
    if (n == 0) return ok();
    const bool stencil_null = stencil && stencil->pred != MXCO_PRED_ALWAYS && !stencil->data;
    if (!pairs || stencil_null) {
        if (sync && num_failed) *num_failed = n;
        return ok();
    }
    if (mxco_status s = check_batch(n, "Insert")) return s;
    OaArgs a = base_args(t);
    if (mxco_status s = set_stencil(a, stencil, "InsertIf")) return s;
    a.in = (const uint8_t*)pairs;
    a.n = n;
    if (sync) {
        if (mxco_status s = zero_counters(t, stream)) return s;
        a.counters = t->counters;
    }
    bool region = region_insert_fits(t, n, a, stream);
    if (region && region_ws(t, n, stream) != hipSuccess) {
        (void)hipGetLastError();  // no room for the workspace: the CAS walk needs none
        region = false;
    }
    const bool was_empty = t->empty;
    t->empty = false;
    if (region) {
        MXCO_HIP(region_insert(t, a, stream, was_empty), "Insert");
    } else if (mxco_status s = launch(t, OP_INSERT, a, stream, "Insert")) {
        return s;
    }
    if (sync) {
        uint64_t f = 0;
        if (mxco_status s = read_counters(t, stream, &f, nullptr)) return s;
        if (num_failed) *num_failed = f;
    }
    return ok();
}

MXCO_API mxco_status mxco_oa_insert_path(const mxco_oa* t, uint64_t n, hipStream_t stream, int* region_build)
{
    if (mxco_status s = check_table(t, "insert_path")) return s;
    if (!region_build) return fail(MXCO_INVALID_ARGUMENT, "insert_path: null output");
    OaArgs a = base_args(t);
    *region_build = region_insert_fits(t, n, a, stream) ? 1 : 0;
    return ok();
}

MXCO_API mxco_status mxco_oa_insert_or_assign(mxco_oa* t, const void* pairs, uint64_t n, hipStream_t stream, int sync,
                                              uint64_t* num_failed, uint64_t* num_assigned)
{
    if (mxco_status s = check_table(t, "InsertOrAssign")) return s;
    if (t->d.value_size == 0) return fail(MXCO_UNSUPPORTED, "InsertOrAssign: containers with a payload only");
    if (num_failed) *num_failed = 0;
    if (num_assigned) *num_assigned = 0;
    if (n == 0) return ok();
    if (!pairs) {
        if (sync && num_failed) *num_failed = n;
        return ok();
    }
    if (mxco_status s = check_batch(n, "InsertOrAssign")) return s;
    t->empty = false;
    OaArgs a = base_args(t);
    a.in = (const uint8_t*)pairs;
    a.n = n;
    if (sync) {
        if (mxco_status s = zero_counters(t, stream)) return s;
        a.counters = t->counters;
    }
    if (mxco_status s = launch(t, OP_INSERT_OR_ASSIGN, a, stream, "InsertOrAssign")) return s;
    if (sync) {
        uint64_t f = 0, g = 0;
        if (mxco_status s = read_counters(t, stream, &f, &g)) return s;
        if (num_failed) *num_failed = f;
        if (num_assigned) *num_assigned = g;
    }
    return ok();
}

MXCO_API mxco_status mxco_oa_assign_if(mxco_oa* t, const void* pairs, const mxco_stencil* stencil, uint64_t n,
                                       hipStream_t stream, int sync)
{
    if (mxco_status s = check_table(t, "AssignIf")) return s;
    if (t->d.value_size == 0) return fail(MXCO_UNSUPPORTED, "AssignIf: containers with a payload only");
    if (n == 0 || !pairs) return ok();
    if (stencil && stencil->pred != MXCO_PRED_ALWAYS && !stencil->data) return ok();
    if (mxco_status s = check_batch(n, "AssignIf")) return s;
    OaArgs a = base_args(t);
    if (mxco_status s = set_stencil(a, stencil, "AssignIf")) return s;
    a.in = (const uint8_t*)pairs;
    a.n = n;
    if (mxco_status s = launch(t, OP_ASSIGN, a, stream, "AssignIf")) return s;
    if (sync) MXCO_HIP(hipStreamSynchronize(stream), "AssignIf: hipStreamSynchronize");
    return ok();
}

MXCO_API mxco_status mxco_oa_insert_and_find(mxco_oa* t, const void* pairs, void* out_found, uint8_t* out_inserted,
                                             uint64_t n, hipStream_t stream, int sync)
{
    if (mxco_status s = check_table(t, "InsertAndFind")) return s;
    if (n == 0 || !pairs) return ok();
    if (!out_found || !out_inserted) return fail(MXCO_INVALID_ARGUMENT, "InsertAndFind: null output");
    if (mxco_status s = check_batch(n, "InsertAndFind")) return s;
    t->empty = false;
    OaArgs a = base_args(t);
    a.in = (const uint8_t*)pairs;
    a.out = (uint8_t*)out_found;
    a.out2 = out_inserted;
    a.n = n;
    if (mxco_status s = launch(t, OP_INSERT_AND_FIND, a, stream, "InsertAndFind")) return s;
    if (sync) MXCO_HIP(hipStreamSynchronize(stream), "InsertAndFind: hipStreamSynchronize");
    return ok();
}

static mxco_status find_like(mxco_oa* t, int op, const void* keys, const mxco_stencil* stencil, void* out, uint64_t n,
                             hipStream_t stream, int sync, const char* fn)
{
    if (mxco_status s = check_table(t, fn)) return s;
    if (n == 0 || !keys || !out) return ok();
    if (stencil && stencil->pred != MXCO_PRED_ALWAYS && !stencil->data) return ok();
    if (mxco_status s = check_batch(n, fn)) return s;
    OaArgs a = base_args(t);
    if (mxco_status s = set_stencil(a, stencil, fn)) return s;
    a.in = (const uint8_t*)keys;
    a.out = (uint8_t*)out;
    a.n = n;
    if (mxco_status s = launch(t, op, a, stream, fn)) return s;
    if (sync) MXCO_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
    return ok();
}

MXCO_API mxco_status mxco_oa_find(mxco_oa* t, const void* keys, const mxco_stencil* stencil, void* out, uint64_t n,
                                  hipStream_t stream, int sync)
{
    return find_like(t, OP_FIND, keys, stencil, out, n, stream, sync, "Find");
}

MXCO_API mxco_status mxco_oa_contains(mxco_oa* t, const void* keys, const mxco_stencil* stencil, uint8_t* out,
                                      uint64_t n, hipStream_t stream, int sync)
{
    return find_like(t, OP_CONTAINS, keys, stencil, out, n, stream, sync, "Contains");
}

MXCO_API mxco_status mxco_oa_count(mxco_oa* t, const void* keys, uint64_t n, hipStream_t stream, uint64_t* count)
{
    if (mxco_status s = check_table(t, "Count")) return s;
    if (!count) return fail(MXCO_INVALID_ARGUMENT, "Count: null output");
    *count = 0;
    if (n == 0 || !keys) return ok();
    if (mxco_status s = check_batch(n, "Count")) return s;
    OaArgs a = base_args(t);
    a.in = (const uint8_t*)keys;
    a.n = n;
    if (mxco_status s = zero_counters(t, stream)) return s;
    a.counters = t->counters;
    if (mxco_status s = launch(t, OP_COUNT, a, stream, "Count")) return s;
    if (mxco_status s = read_counters(t, stream, count, nullptr)) return s;
    return ok();
}

MXCO_API mxco_status mxco_oa_erase(mxco_oa* t, const void* keys, uint64_t n, hipStream_t stream, int sync,
                                   uint64_t* num_failed)
{
    if (mxco_status s = check_table(t, "Erase")) return s;
    if (num_failed) *num_failed = 0;
    if (n == 0) return ok();
    if (!keys) {
        if (sync && num_failed) *num_failed = n;
        return ok();
    }
    if (mxco_status s = check_batch(n, "Erase")) return s;
    OaArgs a = base_args(t);
    a.in = (const uint8_t*)keys;
    a.n = n;
    if (sync) {
        if (mxco_status s = zero_counters(t, stream)) return s;
        a.counters = t->counters;
    }
    if (mxco_status s = launch(t, OP_ERASE, a, stream, "Erase")) return s;
    if (sync) {
        uint64_t f = 0;
        if (mxco_status s = read_counters(t, stream, &f, nullptr)) return s;
        if (num_failed) *num_failed = f;
    }
    return ok();
}

MXCO_API mxco_status mxco_oa_probe_stats(mxco_oa* t, const void* keys, uint64_t n, hipStream_t stream,
                                         uint64_t* lines, uint64_t* slots)
{
    if (mxco_status s = check_table(t, "probe_stats")) return s;
    if (!lines || !slots) return fail(MXCO_INVALID_ARGUMENT, "probe_stats: null output");
    *lines = *slots = 0;
    if (n == 0 || !keys) return ok();
    if (mxco_status s = check_batch(n, "probe_stats")) return s;
    OaArgs a = base_args(t);
    a.in = (const uint8_t*)keys;
    a.n = n;
    if (mxco_status s = zero_counters(t, stream)) return s;
    a.counters = t->counters;
    if (mxco_status s = launch(t, OP_PROBE_STATS, a, stream, "probe_stats")) return s;
    if (mxco_status s = read_counters(t, stream, lines, slots)) return s;
    return ok();
}

MXCO_API mxco_status mxco_oa_set_probe_launcher(mxco_oa* t, mxco_oa_probe_launcher launch)
{
    if (mxco_status s = check_table(t, "set_probe_launcher")) return s;
    t->caller_launch = launch;
    return ok();
}

MXCO_API mxco_status mxco_oa_insert_stats(mxco_oa* t, const void* pairs, uint64_t n, hipStream_t stream,
                                          uint64_t* cas_lost, uint64_t* cas_made)
{
    if (mxco_status s = check_table(t, "insert_stats")) return s;
    if (!cas_lost || !cas_made) return fail(MXCO_INVALID_ARGUMENT, "insert_stats: null output");
    *cas_lost = *cas_made = 0;
    if (n == 0 || !pairs) return ok();
    if (mxco_status s = check_batch(n, "insert_stats")) return s;
    t->empty = false;
    OaArgs a = base_args(t);
    a.in = (const uint8_t*)pairs;
    a.n = n;
    if (mxco_status s = zero_counters(t, stream)) return s;
    a.counters = t->counters;
    if (mxco_status s = launch(t, OP_INSERT_STATS, a, stream, "insert_stats")) return s;
    if (mxco_status s = read_counters(t, stream, cas_lost, cas_made)) return s;
    return ok();
}

}  // extern "C"
