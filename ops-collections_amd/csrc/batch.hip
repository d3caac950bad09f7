// Batched form: thousands of independent StaticMap / StaticSet tables of
// uneven sizes in one slot arena, built and probed by one persistent kernel
// per phase from a device work list.
//
// A plan turns host work items (table, input offset, count) into a device
// chunk list; the persistent kernel's workgroups take chunks in grid-stride
// order, load the chunk's table descriptor and run the same probe state
// machine as the single-table kernels (oa_process), so every table gets the
// reference's results (open_addressing_ref_impl.h:490-605).  Every launch is
// capture-safe (no host copies, no memset nodes): a caller can record Clear
// + Insert + Find into one hipGraph and replay it.  A built-in tuner times
// the plan under each lane-group width and grid size and keeps the fastest;
// results do not depend on that choice.
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <new>
#include <string>
#include <vector>

#include "mxco_common.h"
#include "batch_build.h"
#include "batch_kernels.h"
#include "mxco/detail/oa_kernels.h"

using namespace mxco;

namespace mxco {

template <int KS, int VS, int G, typename Item>
static hipError_t launch_batch_g(int op, const OaArgs& a, const BatchTable* tables, const Item* items,
                                 const BatchChunk* chunks, uint32_t nchunks, unsigned long long* counters,
                                 unsigned grid, hipStream_t s, unsigned block_threads, bool wave_staged)
{
    const dim3 block(block_threads);
    if constexpr (std::is_same<Item, mxco_batch_item>::value) {
        // plans: Find / Contains with wave-staged outputs, 256 threads (MXCO_BATCH_WAVE_STAGED=0 or a
        // plan shape without it selects the plain kernel)
        static const int staged_env = env_int("MXCO_BATCH_WAVE_STAGED", 1);
        if (staged_env && wave_staged && (op == OP_FIND || op == OP_CONTAINS)) {
            // a chunk is 16 wave sub-chunks, 4 wavefronts per workgroup: a
            // grid capped below the chunk count stays as tuned, a smaller one
            // grows to one workgroup per 4 sub-chunks
            const unsigned sg = grid >= nchunks ? nchunks * (kChunk / kWaveChunk / 4) : grid;
            if (op == OP_FIND)
                oa_batch_wave_staged_kernel<KS, VS, G, OP_FIND><<<sg, 256, 0, s>>>(a, tables, items, chunks,
                                                                                   nchunks);
            else
                oa_batch_wave_staged_kernel<KS, VS, G, OP_CONTAINS><<<sg, 256, 0, s>>>(a, tables, items, chunks,
                                                                                       nchunks);
            return hipGetLastError();
        }
    }
    switch (op) {
    case OP_INSERT:
        oa_batch_kernel<KS, VS, G, OP_INSERT, Item><<<grid, block, 0, s>>>(a, tables, items, chunks, nchunks,
                                                                          counters);
        break;
    case OP_FIND:
        oa_batch_kernel<KS, VS, G, OP_FIND, Item><<<grid, block, 0, s>>>(a, tables, items, chunks, nchunks, counters);
        break;
    case OP_CONTAINS:
        oa_batch_kernel<KS, VS, G, OP_CONTAINS, Item><<<grid, block, 0, s>>>(a, tables, items, chunks, nchunks,
                                                                            counters);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int KS, int VS, typename Item>
static hipError_t launch_batch_l(int g, int op, const OaArgs& a, const BatchTable* tables, const Item* items,
                                 const BatchChunk* chunks, uint32_t nchunks, unsigned long long* counters,
                                 unsigned grid, hipStream_t s, unsigned block, bool staged)
{
#define LG(G) launch_batch_g<KS, VS, G, Item>(op, a, tables, items, chunks, nchunks, counters, grid, s, block, staged)
    switch (g) {
    case 1: return LG(1);
    case 2: return LG(2);
    case 4: return LG(4);
    case 8: return LG(8);
    default: return hipErrorInvalidValue;
    }
#undef LG
}

hipError_t launch_batch_any(int ks, int vs, int g, int op, bool serve, const OaArgs& a, const BatchTable* tables,
                            const void* items, const BatchChunk* chunks, uint32_t nchunks,
                            unsigned long long* counters, unsigned grid, hipStream_t s, unsigned block, bool staged)
{
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
#define LB(K, V)                                                                                              \
    return serve ? launch_batch_l<K, V, ServeItem>(g, op, a, tables, (const ServeItem*)items, chunks, nchunks, \
                                                   counters, grid, s, block, staged)                          \
                 : launch_batch_l<K, V, mxco_batch_item>(g, op, a, tables, (const mxco_batch_item*)items,     \
                                                         chunks, nchunks, counters, grid, s, block, staged)
    if (ks == 4 && vs == 4) LB(4, 4);
    if (ks == 8 && vs == 8) LB(8, 8);
    if (ks == 4 && vs == 8) LB(4, 8);
    if (ks == 8 && vs == 4) LB(8, 4);
    if (ks == 4 && vs == 0) LB(4, 0);
    if (ks == 8 && vs == 0) LB(8, 0);
#undef LB
    return hipErrorInvalidValue;
}

}  // namespace mxco

namespace {

__global__ void zero_u64_kernel(unsigned long long* p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ __launch_bounds__(256) void arena_fill_kernel(uint8_t* p, uint64_t bytes, u64x2 pattern)
{
    const uint64_t n16 = bytes / 16;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        ((u64x2*)p)[i] = pattern;
}

}  // namespace

struct mxco_batch {
    mxco_oa_desc d;
    uint32_t ss;
    uint32_t ntables;
    uint8_t* arena;
    uint64_t arena_bytes;
    BatchTable* tables_dev;
    std::vector<BatchTable> tables;
    uint64_t empty_key, empty_val;
};

struct mxco_batch_plan {
    mxco_batch* batch;
    int op;
    uint32_t nitems;
    uint32_t nchunks;
    mxco_batch_item* items_dev;
    BatchChunk* chunks_dev;
    unsigned long long* counters;  // 2 per item
    unsigned long long* host_counters;
    int group;
    unsigned blocks_per_cu;
    unsigned block = 256;     // threads per workgroup of the probe kernel
    bool wave_staged = true;  // Find / Contains: the wave-staged kernel (256 threads)
    // MXCO_BATCH_BUILD
    bool build;   // Clear of every table, then this insert plan
    bool region;  // ... by the region build (otherwise Clear + the insert kernel)
    uint32_t nreg, ntiles, maxbins;
    uint64_t total;
    BatchChunk* tiles_dev;
    uint32_t* reg0_dev;
    uint32_t* reg_table_dev;
    int32_t* table_item_dev;
    uint32_t* counts;   // nreg + 1
    uint32_t* offsets;  // nreg + 1
    uint32_t* ovc;      // nreg
    uint8_t* part;      // total pairs
};

namespace {

OaArgs batch_args(const mxco_batch* b)
{
    OaArgs a;
    memset(&a, 0, sizeof(a));
    a.bucket = b->d.bucket_size;
    a.probing = b->d.probing;
    a.hash = b->d.hash;
    a.key_float = b->d.key_kind == MXCO_KEY_FLOAT;
    a.empty_key = b->empty_key;
    a.empty_val = b->empty_val;
    a.erased_key = b->empty_key;
    return a;
}

// A probe launch shape of a plan: lanes per key, workgroups per CU,
// threads per workgroup, and (Find / Contains) the wave-staged kernel.
struct Shape {
    int group;
    unsigned bpc, block;
    bool staged;
};

Shape plan_shape(const mxco_batch_plan* p) { return {p->group, p->blocks_per_cu, p->block, p->wave_staged}; }

hipError_t dispatch(const mxco_batch* b, int op, const OaArgs& a, const mxco_batch_plan* p, unsigned grid,
                    const Shape& sh, hipStream_t s)
{
    return launch_batch_any(b->d.key_size, b->d.value_size, sh.group, op, false, a, b->tables_dev, p->items_dev,
                            p->chunks_dev, p->nchunks, op == OP_INSERT ? p->counters : nullptr, grid, s, sh.block,
                            sh.staged);
}

__global__ void zero_u32_kernel(uint32_t* p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

template <int KS, int VS>
hipError_t launch_build(const BuildArgs& a, uint32_t ntiles, hipStream_t s)
{
    constexpr int SS = Lay<KS, VS>::SS;
    static const hipError_t attr = hipFuncSetAttribute((const void*)build_scatter_kernel<KS, VS, false>,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                                       (int)build_scatter_lds<SS>(kBuildMaxBins));
    if (attr != hipSuccess) return attr;
    if (ntiles) build_count_kernel<KS, VS, false><<<ntiles, kBuildThreads, 0, s>>>(a);
    build_scan_kernel<<<1, 1024, 0, s>>>(a.counts, a.offsets, a.nreg);
    if (ntiles) build_scatter_kernel<KS, VS, false><<<ntiles, kBuildThreads, build_scatter_lds<SS>(a.maxbins), s>>>(a);
    build_apply_kernel<KS, VS, false, false><<<a.nreg, kBuildThreads, 0, s>>>(a);
    const unsigned waves_grid = std::max(1u, std::min<unsigned>((a.nreg + 3) / 4, (unsigned)cu_count() * 8));
    build_overflow_kernel<KS, VS, false><<<waves_grid, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_build_any(int ks, int vs, const BuildArgs& a, uint32_t ntiles, hipStream_t s)
{
    if (ks == 4 && vs == 4) return launch_build<4, 4>(a, ntiles, s);
    if (ks == 8 && vs == 8) return launch_build<8, 8>(a, ntiles, s);
    if (ks == 4 && vs == 8) return launch_build<4, 8>(a, ntiles, s);
    if (ks == 8 && vs == 4) return launch_build<8, 4>(a, ntiles, s);
    if (ks == 4 && vs == 0) return launch_build<4, 0>(a, ntiles, s);
    if (ks == 8 && vs == 0) return launch_build<8, 0>(a, ntiles, s);
    return hipErrorInvalidValue;
}

mxco_status run_plan(mxco_batch_plan* p, const void* in, void* out, uint8_t* out2, hipStream_t s, const Shape& sh);

// MXCO_BATCH_BUILD: Clear + Insert (capture-safe: kernels only).
mxco_status run_build(mxco_batch_plan* p, const void* in, hipStream_t s)
{
    mxco_batch* b = p->batch;
    if (!p->region) {
        if (mxco_status st = mxco_batch_clear(b, s)) return st;
        return run_plan(p, in, nullptr, nullptr, s, plan_shape(p));
    }
    zero_u64_kernel<<<grid_for(2ull * p->nitems, 1), 256, 0, s>>>(p->counters, 2ull * p->nitems);
    zero_u32_kernel<<<grid_for(p->nreg + 1ull, 1), 256, 0, s>>>(p->counts, p->nreg + 1ull);
    MXCO_HIP(hipGetLastError(), "batch build: zero");
    BuildArgs a;
    memset(&a, 0, sizeof(a));
    a.in = (const uint8_t*)in;
    a.part = p->part;
    a.tables = b->tables_dev;
    a.items = p->items_dev;
    a.tiles = p->tiles_dev;
    a.reg0 = p->reg0_dev;
    a.reg_table = p->reg_table_dev;
    a.table_item = p->table_item_dev;
    a.counts = p->counts;
    a.offsets = p->offsets;
    a.ovc = p->ovc;
    a.counters = p->counters;
    a.empty_key = b->empty_key;
    a.empty_val = b->empty_val;
    a.nreg = p->nreg;
    a.maxbins = p->maxbins;
    a.bin_slots = build_region_slots(b->ss);  // a bin is one region
    a.bucket = b->d.bucket_size;
    a.hash = b->d.hash;
    a.key_float = b->d.key_kind == MXCO_KEY_FLOAT;
    MXCO_HIP(launch_build_any(b->d.key_size, b->d.value_size, a, p->ntiles, s), "batch build: launch");
    return MXCO_OK;
}

// Region build data of a build plan; leaves p->region false when the plan
// does not qualify (double hashing, a table in two items, a table of more
// than kBuildMaxBins regions, 2^32 pairs or more).
hipError_t prepare_build(mxco_batch_plan* p, const mxco_batch_item* items)
{
    mxco_batch* b = p->batch;
    static const int enabled = env_int("MXCO_BATCH_REGION_BUILD", 1);
    if (!enabled || b->d.probing != 0) return hipSuccess;
    const uint32_t R = build_region_slots(b->ss);
    const uint32_t tile = build_tile_pairs(b->ss);
    std::vector<int32_t> table_item(b->ntables, -1);
    std::vector<uint32_t> reg0(b->ntables + 1);
    uint64_t nreg = 0, total = 0;
    uint32_t maxbins = 1;
    for (uint32_t t = 0; t < b->ntables; ++t) {
        reg0[t] = (uint32_t)nreg;
        nreg += (b->tables[t].cap + R - 1) / R;
        if (nreg >= 0x7fffffffull) return hipSuccess;
    }
    reg0[b->ntables] = (uint32_t)nreg;
    std::vector<BatchChunk> tiles;
    for (uint32_t i = 0; i < p->nitems; ++i) {
        const uint32_t t = items[i].table;
        if (table_item[t] >= 0) return hipSuccess;
        table_item[t] = (int32_t)i;
        const uint32_t bins = reg0[t + 1] - reg0[t];
        if (bins > kBuildMaxBins) return hipSuccess;
        maxbins = std::max(maxbins, bins);
        total += items[i].count;
        for (uint64_t s = 0; s < items[i].count; s += tile)
            tiles.push_back({i, (uint32_t)std::min<uint64_t>(tile, items[i].count - s), s});
    }
    if (total >= 0xffffffffull) return hipSuccess;
    std::vector<uint32_t> reg_table(nreg);
    for (uint32_t t = 0; t < b->ntables; ++t)
        for (uint32_t r = reg0[t]; r < reg0[t + 1]; ++r) reg_table[r] = t;
    p->nreg = (uint32_t)nreg;
    p->ntiles = (uint32_t)tiles.size();
    p->maxbins = maxbins;
    p->total = total;
    hipError_t e = hipMalloc((void**)&p->tiles_dev, std::max<size_t>(1, tiles.size()) * sizeof(BatchChunk));
    if (e == hipSuccess) e = hipMalloc((void**)&p->reg0_dev, reg0.size() * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&p->reg_table_dev, nreg * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&p->table_item_dev, b->ntables * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&p->counts, (nreg + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&p->offsets, (nreg + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&p->ovc, nreg * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&p->part, std::max<uint64_t>(16, total * b->ss));
    if (e == hipSuccess && !tiles.empty())
        e = hipMemcpy(p->tiles_dev, tiles.data(), tiles.size() * sizeof(BatchChunk), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->reg0_dev, reg0.data(), reg0.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p->reg_table_dev, reg_table.data(), nreg * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p->table_item_dev, table_item.data(), b->ntables * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) p->region = true;
    return e;
}

mxco_status run_plan(mxco_batch_plan* p, const void* in, void* out, uint8_t* out2, hipStream_t s, const Shape& sh)
{
    if (p->op == OP_INSERT) {
        zero_u64_kernel<<<grid_for(2ull * p->nitems, 1), 256, 0, s>>>(p->counters, 2ull * p->nitems);
        MXCO_HIP(hipGetLastError(), "batch: zero counters");
    }
    OaArgs a = batch_args(p->batch);
    a.in = (const uint8_t*)in;
    a.out = (uint8_t*)out;
    a.out2 = out2;
    const unsigned grid = std::max(1u, std::min<unsigned>(p->nchunks, (unsigned)cu_count() * sh.bpc));
    MXCO_HIP(dispatch(p->batch, p->op, a, p, grid, sh, s), "batch: launch");
    return MXCO_OK;
}

}  // namespace

extern "C" {

MXCO_API mxco_status mxco_batch_create(const mxco_oa_desc* desc, uint32_t num_tables, const uint64_t* capacities,
                                       const void* empty_key, const void* empty_value, hipStream_t stream,
                                       mxco_batch** out)
{
    if (!desc || !capacities || !empty_key || !out || num_tables == 0)
        return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_create: bad arguments");
    *out = nullptr;
    const mxco_oa_desc& d = *desc;
    if ((d.key_size != 4 && d.key_size != 8) || (d.value_size != 0 && d.value_size != 4 && d.value_size != 8))
        return fail(MXCO_UNSUPPORTED, "mxco_batch_create: keys of 4/8 bytes, payloads of 0/4/8 bytes");
    if (d.value_size && !empty_value) return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_create: null empty value");
    if (d.bucket_size < 1 || d.bucket_size > 1024 || d.probing > 1 || d.hash > MXCO_HASH_FMIX_64)
        return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_create: bad descriptor");
    mxco_batch* b = new (std::nothrow) mxco_batch();
    if (!b) return fail(MXCO_OUT_OF_MEMORY, "mxco_batch_create: host allocation");
    b->d = d;
    b->ss = d.value_size == 0 ? d.key_size : ((d.key_size == 4 && d.value_size == 4) ? 8 : 16);
    b->ntables = num_tables;
    memcpy(&b->empty_key, empty_key, d.key_size);
    if (d.value_size) memcpy(&b->empty_val, empty_value, d.value_size);
    b->tables.resize(num_tables);
    uint64_t off = 0;
    std::vector<uint64_t> offs(num_tables);
    for (uint32_t t = 0; t < num_tables; ++t) {
        const uint64_t cap = valid_extent(capacities[t], d.bucket_size, d.probing);
        if (cap == 0 || cap > 0xffffffffull) {
            delete b;
            return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_create: invalid capacity for table " + std::to_string(t));
        }
        offs[t] = off;
        b->tables[t].cap = (uint32_t)cap;
        b->tables[t].nbuckets = (uint32_t)(cap / d.bucket_size);
        off += (cap * b->ss + 255) & ~255ull;  // each table starts on a 256-byte boundary
    }
    b->arena_bytes = off;
    hipError_t e = hipMalloc((void**)&b->arena, off);
    if (e == hipSuccess) {
        for (uint32_t t = 0; t < num_tables; ++t) b->tables[t].slots = b->arena + offs[t];
        e = hipMalloc((void**)&b->tables_dev, num_tables * sizeof(BatchTable));
    }
    if (e == hipSuccess)
        e = hipMemcpy(b->tables_dev, b->tables.data(), num_tables * sizeof(BatchTable), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (b->arena) (void)hipFree(b->arena);
        if (b->tables_dev) (void)hipFree(b->tables_dev);
        delete b;
        return hip_fail(e, "mxco_batch_create: allocation");
    }
    *out = b;
    mxco_status st = mxco_batch_clear(b, stream);
    if (st == MXCO_OK) st = hipStreamSynchronize(stream) == hipSuccess ? MXCO_OK : MXCO_RUNTIME_ERROR;
    return st == MXCO_OK ? ok() : st;
}

MXCO_API mxco_status mxco_batch_destroy(mxco_batch* b)
{
    if (!b) return ok();
    if (b->arena) (void)hipFree(b->arena);
    if (b->tables_dev) (void)hipFree(b->tables_dev);
    delete b;
    return ok();
}

MXCO_API mxco_status mxco_batch_table(const mxco_batch* b, uint32_t table, uint64_t* capacity, void** slots)
{
    if (!b || table >= b->ntables) return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_table: bad table");
    if (capacity) *capacity = b->tables[table].cap;
    if (slots) *slots = b->tables[table].slots;
    return ok();
}

MXCO_API mxco_status mxco_batch_clear(mxco_batch* b, hipStream_t stream)
{
    if (!b) return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_clear: null");
    uint8_t img[16];
    memset(img, 0, sizeof(img));
    const uint32_t voff = (b->ss == 8 && b->d.value_size) ? 4 : 8;
    for (uint32_t o = 0; o < 16; o += b->ss) {
        memcpy(img + o, &b->empty_key, b->d.key_size);
        if (b->d.value_size) memcpy(img + o + voff, &b->empty_val, b->d.value_size);
    }
    u64x2 pat;
    memcpy(&pat, img, 16);
    arena_fill_kernel<<<grid_for(b->arena_bytes / 16, kFillBlocksPerCu), 256, 0, stream>>>(b->arena, b->arena_bytes,
                                                                                           pat);
    MXCO_HIP(hipGetLastError(), "mxco_batch_clear");
    return ok();
}

MXCO_API mxco_status mxco_batch_plan_create(mxco_batch* b, int op, const mxco_batch_item* items, uint32_t num_items,
                                            uint32_t group_size, mxco_batch_plan** out)
{
    if (!b || !items || !out || num_items == 0) return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_plan_create: bad args");
    if (op != MXCO_BATCH_INSERT && op != MXCO_BATCH_FIND && op != MXCO_BATCH_CONTAINS && op != MXCO_BATCH_BUILD)
        return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_plan_create: unknown operation");
    if (group_size != 0 && group_size != 1 && group_size != 2 && group_size != 4 && group_size != 8)
        return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_plan_create: group size must be 0, 1, 2, 4 or 8");
    *out = nullptr;
    std::vector<BatchChunk> chunks;
    for (uint32_t i = 0; i < num_items; ++i) {
        if (items[i].table >= b->ntables)
            return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_plan_create: item names an unknown table");
        for (uint64_t s = 0; s < items[i].count; s += kChunk)
            chunks.push_back({i, (uint32_t)std::min<uint64_t>(kChunk, items[i].count - s), s});
    }
    // Chunks of one item stay adjacent, so neighbouring workgroups share a
    // table's lines in L2.
    mxco_batch_plan* p = new (std::nothrow) mxco_batch_plan();
    if (!p) return fail(MXCO_OUT_OF_MEMORY, "mxco_batch_plan_create: host allocation");
    p->batch = b;
    p->op = (op == MXCO_BATCH_INSERT || op == MXCO_BATCH_BUILD) ? OP_INSERT
                                                                 : (op == MXCO_BATCH_FIND ? OP_FIND : OP_CONTAINS);
    p->build = op == MXCO_BATCH_BUILD;
    p->nitems = num_items;
    p->nchunks = (uint32_t)chunks.size();
    p->group = group_size ? (int)group_size : (b->ss == 16 ? 4 : 4);
    p->blocks_per_cu = 16;
    hipError_t e = hipMalloc((void**)&p->items_dev, num_items * sizeof(mxco_batch_item));
    if (e == hipSuccess) e = hipMalloc((void**)&p->chunks_dev, std::max<size_t>(1, chunks.size()) * sizeof(BatchChunk));
    if (e == hipSuccess) e = hipMalloc((void**)&p->counters, 2ull * num_items * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipHostMalloc((void**)&p->host_counters, 2ull * num_items * sizeof(unsigned long long), 0);
    if (e == hipSuccess)
        e = hipMemcpy(p->items_dev, items, num_items * sizeof(mxco_batch_item), hipMemcpyHostToDevice);
    if (e == hipSuccess && !chunks.empty())
        e = hipMemcpy(p->chunks_dev, chunks.data(), chunks.size() * sizeof(BatchChunk), hipMemcpyHostToDevice);
    if (e == hipSuccess && p->build) e = prepare_build(p, items);
    if (e != hipSuccess) {
        (void)mxco_batch_plan_destroy(p);
        return hip_fail(e, "mxco_batch_plan_create: allocation");
    }
    *out = p;
    return ok();
}

MXCO_API mxco_status mxco_batch_plan_destroy(mxco_batch_plan* p)
{
    if (!p) return ok();
    if (p->items_dev) (void)hipFree(p->items_dev);
    if (p->chunks_dev) (void)hipFree(p->chunks_dev);
    if (p->counters) (void)hipFree(p->counters);
    if (p->host_counters) (void)hipHostFree(p->host_counters);
    for (void* d : {(void*)p->tiles_dev, (void*)p->reg0_dev, (void*)p->reg_table_dev, (void*)p->table_item_dev,
                    (void*)p->counts, (void*)p->offsets, (void*)p->ovc, (void*)p->part})
        if (d) (void)hipFree(d);
    delete p;
    return ok();
}

MXCO_API mxco_status mxco_batch_run(mxco_batch_plan* p, const void* in, void* out, uint8_t* out2, hipStream_t stream)
{
    if (!p || !in) return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_run: null argument");
    if (p->op != OP_INSERT && !out) return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_run: null output");
    if (p->build) {
        if (mxco_status s = run_build(p, in, stream)) return s;
        return ok();
    }
    if (mxco_status s = run_plan(p, in, out, out2, stream, plan_shape(p))) return s;
    return ok();
}

MXCO_API mxco_status mxco_batch_failures(mxco_batch_plan* p, hipStream_t stream, uint64_t* failed_per_item)
{
    if (!p || !failed_per_item || p->op != OP_INSERT)
        return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_failures: an insert plan and an output are required");
    MXCO_HIP(hipMemcpyAsync(p->host_counters, p->counters, 2ull * p->nitems * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, stream),
             "mxco_batch_failures");
    MXCO_HIP(hipStreamSynchronize(stream), "mxco_batch_failures");
    for (uint32_t i = 0; i < p->nitems; ++i) failed_per_item[i] = p->host_counters[2ull * i];
    return ok();
}

MXCO_API mxco_status mxco_batch_plan_config(mxco_batch_plan* p, uint32_t* group_size, uint32_t* blocks_per_cu)
{
    if (!p) return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_plan_config: null plan");
    if (group_size) *group_size = (uint32_t)p->group;
    if (blocks_per_cu) *blocks_per_cu = p->blocks_per_cu;
    return ok();
}

// Times the plan under every launch shape (lane-group width, grid size,
// workgroup size, Find / Contains kernel) and keeps the fastest.  The run must be repeatable on the given buffers: a Find or
// Contains plan always is; an insert plan is timed on a cleared table
// arena each time and left cleared.
MXCO_API mxco_status mxco_batch_plan_tune(mxco_batch_plan* p, const void* in, void* out, uint8_t* out2,
                                          hipStream_t stream, float* best_ms)
{
    if (!p || !in) return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_plan_tune: null argument");
    if (p->build && p->region) {
        // the region build has no launch shape to choose: time it (3 runs, the first discarded)
        hipEvent_t e0, e1;
        MXCO_HIP(hipEventCreate(&e0), "tune");
        MXCO_HIP(hipEventCreate(&e1), "tune");
        mxco_status st = MXCO_OK;
        float total = 0;
        for (int rep = 0; rep < 3 && st == MXCO_OK; ++rep) {
            if (hipEventRecord(e0, stream) != hipSuccess) st = fail(MXCO_RUNTIME_ERROR, "tune: event");
            if (st == MXCO_OK) st = run_build(p, in, stream);
            if (st == MXCO_OK && hipEventRecord(e1, stream) != hipSuccess) st = fail(MXCO_RUNTIME_ERROR, "tune");
            if (st == MXCO_OK && hipEventSynchronize(e1) != hipSuccess) st = fail(MXCO_RUNTIME_ERROR, "tune");
            float ms = 0;
            if (st == MXCO_OK) (void)hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0) total += ms;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (st != MXCO_OK) return st;
        if (best_ms) *best_ms = total / 2;
        return ok();
    }
    // the grid: lanes per key x workgroups per CU x threads per workgroup
    // (64 / 128 / 256) for the plain probe kernel, and for a Find / Contains
    // plan also lanes x workgroups for the wave-staged kernel (256 threads)
    static const int groups[] = {1, 2, 4, 8};
    static const unsigned bpcs[] = {8, 32, 64, 128, 256, 512};
    static const unsigned blocks[] = {64, 128, 256};
    std::vector<Shape> grid;
    for (int g : groups)
        for (unsigned bpc : bpcs) {
            if (p->op != OP_INSERT) grid.push_back({g, bpc, 256, true});
            for (unsigned blk : blocks) grid.push_back({g, bpc, blk, false});
        }
    hipEvent_t e0, e1;
    MXCO_HIP(hipEventCreate(&e0), "tune");
    MXCO_HIP(hipEventCreate(&e1), "tune");
    float best = 1e30f;
    Shape bs = plan_shape(p);
    mxco_status st = MXCO_OK;
    for (const Shape& sh : grid) {
        float total = 0;
        for (int rep = 0; rep < 3 && st == MXCO_OK; ++rep) {
            if (p->op == OP_INSERT) st = mxco_batch_clear(p->batch, stream);
            if (st != MXCO_OK) break;
            if (hipEventRecord(e0, stream) != hipSuccess) st = fail(MXCO_RUNTIME_ERROR, "tune: event");
            if (st == MXCO_OK) st = run_plan(p, in, out, out2, stream, sh);
            if (st == MXCO_OK && hipEventRecord(e1, stream) != hipSuccess) st = fail(MXCO_RUNTIME_ERROR, "tune");
            if (st == MXCO_OK && hipEventSynchronize(e1) != hipSuccess) st = fail(MXCO_RUNTIME_ERROR, "tune");
            float ms = 0;
            if (st == MXCO_OK) (void)hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0) total += ms;  // first repetition warms up
        }
        if (st != MXCO_OK) break;
        if (total < best) {
            best = total;
            bs = sh;
        }
    }
    // Leave the tables as one run() of the plan leaves them (an insert plan
    // tuned above ran several times), so a find plan tuned next sees the
    // filled tables.
    if (p->op == OP_INSERT && st == MXCO_OK) st = mxco_batch_clear(p->batch, stream);
    if (p->op == OP_INSERT && st == MXCO_OK) st = run_plan(p, in, out, out2, stream, bs);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != MXCO_OK) return st;
    MXCO_HIP(hipStreamSynchronize(stream), "tune");
    p->group = bs.group;
    p->blocks_per_cu = bs.bpc;
    p->block = bs.block;
    p->wave_staged = bs.staged;
    if (best_ms) *best_ms = best / 2;
    return ok();
}

MXCO_API mxco_status mxco_batch_plan_set_config(mxco_batch_plan* p, uint32_t group_size, uint32_t blocks_per_cu)
{
    if (!p || (group_size != 1 && group_size != 2 && group_size != 4 && group_size != 8) || blocks_per_cu == 0)
        return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_plan_set_config: bad arguments");
    p->group = (int)group_size;
    p->blocks_per_cu = blocks_per_cu;
    return ok();
}

MXCO_API mxco_status mxco_batch_plan_launch_shape(mxco_batch_plan* p, uint32_t* block_size, uint32_t* wave_staged)
{
    if (!p) return fail(MXCO_INVALID_ARGUMENT, "mxco_batch_plan_launch_shape: null plan");
    if (block_size) *block_size = p->block;
    if (wave_staged) *wave_staged = p->wave_staged ? 1u : 0u;
    return ok();
}

MXCO_API mxco_status mxco_batch_plan_set_launch_shape(mxco_batch_plan* p, uint32_t block_size, uint32_t wave_staged)
{
    if (!p || (block_size != 64 && block_size != 128 && block_size != 256) || wave_staged > 1 ||
        (wave_staged && block_size != 256))
        return fail(MXCO_INVALID_ARGUMENT,
                    "mxco_batch_plan_set_launch_shape: 64, 128 or 256 threads; the wave-staged kernel takes 256");
    p->block = block_size;
    p->wave_staged = wave_staged != 0;
    return ok();
}

}  // extern "C"
