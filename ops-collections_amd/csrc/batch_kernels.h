// Persistent multi-table probe kernel shared by the batched tables and the
// serving engine: workgroups take chunks of a device work list, load the
// chunk's table descriptor and run the single-table probe state machine.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/mxco.h"
#include "mxco/detail/oa_kernels.h"

namespace mxco {

struct BatchTable {
    uint8_t* slots;
    uint32_t cap;
    uint32_t nbuckets;
};

struct BatchChunk {
    uint32_t item;
    uint32_t len;
    uint64_t start;  // first key of the chunk, relative to the item
};

constexpr uint32_t kChunk = 2048;  // keys per chunk

// Work item of the serving engine: explicit device pointers per request.
struct ServeItem {
    uint32_t table;
    uint32_t op;  // read launches: MXCO_REQ_FIND or MXCO_REQ_CONTAINS
    const uint8_t* in;
    uint8_t* out;
    uint64_t count;
};

template <typename Item>
__device__ __forceinline__ void item_io(const Item& it, const OaArgs& base, int in_stride, int out_stride, OaArgs& a);

template <>
__device__ __forceinline__ void item_io<mxco_batch_item>(const mxco_batch_item& it, const OaArgs& base, int in_stride,
                                                         int out_stride, OaArgs& a)
{
    a.in = base.in + it.offset * in_stride;
    if (base.out) a.out = base.out + it.offset * out_stride;
    if (base.out2) a.out2 = base.out2 + it.offset;
}

template <>
__device__ __forceinline__ void item_io<ServeItem>(const ServeItem& it, const OaArgs&, int, int, OaArgs& a)
{
    a.in = it.in;
    a.out = it.out;
}

template <int KS, int VS, int G, int OP, typename Item>
__global__ __launch_bounds__(256) void oa_batch_kernel(OaArgs base, const BatchTable* tables, const Item* items,
                                                       const BatchChunk* chunks, uint32_t nchunks,
                                                       unsigned long long* item_counters)
{
    constexpr int SS = Lay<KS, VS>::SS;
    constexpr bool kInsertFamily =
        OP == OP_INSERT || OP == OP_INSERT_OR_ASSIGN || OP == OP_ASSIGN || OP == OP_INSERT_AND_FIND;
    constexpr int IN_STRIDE = kInsertFamily ? SS : KS;
    constexpr int OUT_STRIDE = OP == OP_CONTAINS ? 1 : (VS ? VS : KS);
    const uint32_t gpb = blockDim.x / G;
    const uint32_t gj = threadIdx.x / G;
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const BatchChunk ch = chunks[c];
        const Item it = items[ch.item];
        const BatchTable t = tables[it.table];
        OaArgs a = base;
        a.table = t.slots;
        a.cap = t.cap;
        a.nbuckets = t.nbuckets;
        item_io<Item>(it, base, IN_STRIDE, OUT_STRIDE, a);
        a.n = it.count;
        unsigned long long c0 = 0, c1 = 0;
        if constexpr (std::is_same<Item, ServeItem>::value && OP == OP_FIND) {
            // one read launch serves Find and Contains items (uniform per chunk)
            if (it.op == MXCO_REQ_CONTAINS)
                oa_process<KS, VS, G, OP_CONTAINS>(a, ch.start + gj, ch.start + ch.len, gpb, c0, c1);
            else
                oa_process<KS, VS, G, OP_FIND>(a, ch.start + gj, ch.start + ch.len, gpb, c0, c1);
        } else {
            oa_process<KS, VS, G, OP>(a, ch.start + gj, ch.start + ch.len, gpb, c0, c1);
        }
        if (item_counters) flush_counters<OP>(item_counters + 2ull * ch.item, c0, c1);
    }
}

// Wave-staged Find / Contains over a plan's chunk list: every 2048-key chunk
// is cut into kWaveChunk-key sub-chunks, and wavefront w takes sub-chunks
// w, w + waves, ... (skipping the empty tails of short chunks).
template <int KS, int OUTB>
struct BatchChunks {
    static constexpr bool kOneTable = false;
    const OaArgs* base;
    const BatchTable* tables;
    const mxco_batch_item* items;
    const BatchChunk* chunks;
    uint32_t nchunks;
    uint64_t next, step;
    __device__ __forceinline__ WsChunk open()
    {
        constexpr uint32_t kSub = kChunk / kWaveChunk;
        WsChunk c;
        c.len = 0;
        for (;;) {
            const uint64_t sc = next;
            next += step;
            const uint64_t ci = sc / kSub;
            if (ci >= nchunks) return c;
            const BatchChunk ch = chunks[ci];
            const uint32_t off = (uint32_t)(sc % kSub) * kWaveChunk;
            if (off >= ch.len) continue;
            const mxco_batch_item it = items[ch.item];
            const BatchTable t = tables[it.table];
            const uint64_t first = it.offset + ch.start + off;
            c.table = t.slots;
            c.cap = t.cap;
            c.nbuckets = t.nbuckets;
            c.in = base->in + first * KS;
            c.out = base->out + first * OUTB;
            c.stencil = nullptr;
            c.len = min(kWaveChunk, ch.len - off);
            return c;
        }
    }
};

template <int KS, int VS, int G, int OP>
__global__ __launch_bounds__(256) void oa_batch_wave_staged_kernel(OaArgs base, const BatchTable* tables,
                                                                   const mxco_batch_item* items,
                                                                   const BatchChunk* chunks, uint32_t nchunks)
{
    static_assert(kChunk % kWaveChunk == 0, "chunks split into whole sub-chunks");
    constexpr int OUTB = OP == OP_CONTAINS ? 1 : (VS ? VS : KS);
    __shared__ __align__(16) uint8_t stage[4][2][kWaveChunk * OUTB];
    __shared__ __align__(16) uint8_t keys[4][2][kWaveChunk * KS];
    const uint32_t wib = threadIdx.x >> 6;
    BatchChunks<KS, OUTB> src{&base, tables, items, chunks, nchunks,
                              (uint64_t)blockIdx.x * (blockDim.x >> 6) + wib, (uint64_t)gridDim.x * (blockDim.x >> 6)};
    wave_staged_find<KS, VS, G, OP>(base, src, stage[wib][0], stage[wib][1], keys[wib][0], keys[wib][1]);
}

// Launches the persistent multi-table kernel (defined in batch.hip).
// `serve` selects ServeItem items instead of mxco_batch_item.
hipError_t launch_batch_any(int ks, int vs, int g, int op, bool serve, const OaArgs& a, const BatchTable* tables,
                            const void* items, const BatchChunk* chunks, uint32_t nchunks,
                            unsigned long long* counters, unsigned grid, hipStream_t s, unsigned block = 256,
                            bool staged = true);

}  // namespace mxco
