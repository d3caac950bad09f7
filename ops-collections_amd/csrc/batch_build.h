// Region build of batched tables: Clear followed by Insert of one plan, in
// one pass over the slot arena.
//
// Every table is cut into regions of kBuildBytes of slots.  A key belongs to
// the region holding the first slot of its walk, (h % nbuckets) * bucket
// under LinearProbing (probing_scheme_impl.inl:33-47 of the reference).
//   1. count:   per tile of one item, an LDS histogram of the keys' regions,
//               added into per-region counts;
//   2. scan:    exclusive sum of the counts (one workgroup);
//   3. scatter: each tile sorted by region in LDS and written as one run per
//               region, so the plan's pairs end up grouped by region;
//   4. apply:   one workgroup per region fills the region with the empty slot
//               in LDS and walks each of its keys slot by slot from the first
//               slot, as the CAS walk does (open_addressing_ref_impl.h:490-532):
//               the first slot that is empty or holds an equal key ends the
//               walk -- EQUAL is a duplicate, EMPTY an LDS compare-and-swap,
//               re-examined on the same slot when lost.  The region is then
//               written once with 16-byte stores;
//   5. overflow: a key whose walk passes the region's last slot is kept and
//               inserted after every region is written, by the CAS walk on
//               the table resumed at the next slot.  The slots it skipped are
//               full and hold other keys, so the final table is one that the
//               CAS walk produces for some order of the same inserts.
// The walk order and the EMPTY / EQUAL rules are those of the CAS insert, so
// Find, Contains and the per-item failure counts equal Clear + Insert's.
#pragma once
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "batch_kernels.h"
#include "oa_kernels.h"

namespace mxco {

constexpr uint32_t kBuildBytes = 65536;    // slot bytes of one region (held in LDS)
constexpr uint32_t kBuildThreads = 1024;   // count / scatter / apply workgroups
constexpr uint32_t kBuildMaxBins = 4096;   // regions of one table
constexpr uint32_t kBuildTileBytes = 65536;  // pair bytes of one scatter tile

template <int SS>
struct BuildShape {
    static constexpr uint32_t kRegion = kBuildBytes / SS;     // slots per region
    static constexpr uint32_t kTile = kBuildTileBytes / SS;   // pairs per tile
    static constexpr uint32_t kPer = kTile / kBuildThreads;   // pairs per thread of a tile
};

struct BuildArgs {
    const uint8_t* in;  // the plan's input pairs (slot layout)
    uint8_t* part;      // the pairs grouped by region (total * SS bytes)
    const BatchTable* tables;
    const mxco_batch_item* items;
    const BatchChunk* tiles;    // tiles of one item each, at most kTile pairs
    const uint32_t* reg0;       // first region of each table (num_tables + 1)
    const uint32_t* reg_table;  // table of each region
    const int32_t* table_item;  // item of each table, or -1
    uint32_t* counts;           // per region; zero again after the scatter
    uint32_t* offsets;          // exclusive sum of the counts (nreg + 1)
    uint32_t* ovc;              // overflow pairs of each region, kept at its offset
    unsigned long long* counters;  // 2 per item: [0] failed inserts
    uint64_t empty_key, empty_val;
    uint32_t nreg;
    uint32_t maxbins;  // most regions of a table with an item (stride of the scatter's LDS arrays)
    uint32_t bucket;
    uint32_t hash;
    uint32_t key_float;
};

template <int KS>
__device__ __forceinline__ uint32_t first_slot(const BuildArgs& a, uint32_t nbuckets, uint64_t key)
{
    return (probe_hash<KS>(a.hash, key, 0u) % nbuckets) * a.bucket;
}

// The empty slot of layout L (unused bytes zero, as Clear writes them).
template <int KS, int VS>
__device__ __forceinline__ typename Lay<KS, VS>::Raw empty_slot(uint64_t ek, uint64_t ev)
{
    using Raw = typename Lay<KS, VS>::Raw;
    if constexpr (Lay<KS, VS>::SS == 16) {
        Raw r;
        r.x = KS == 4 ? (ek & 0xffffffffull) : ek;
        r.y = VS == 4 ? (ev & 0xffffffffull) : ev;
        return r;
    } else if constexpr (Lay<KS, VS>::SS == 8 && VS == 4) {
        return (Raw)((ek & 0xffffffffull) | (ev << 32));
    } else {
        return (Raw)ek;
    }
}

// 1. Region histogram of one tile per workgroup.
template <int KS, int VS>
__global__ __launch_bounds__(kBuildThreads) void build_count_kernel(BuildArgs a)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    using S = BuildShape<L::SS>;
    __shared__ uint32_t hist[kBuildMaxBins];
    const BatchChunk tl = a.tiles[blockIdx.x];
    const mxco_batch_item it = a.items[tl.item];
    const BatchTable t = a.tables[it.table];
    const uint32_t nbins = (t.cap + S::kRegion - 1) / S::kRegion;
    for (uint32_t j = threadIdx.x; j < nbins; j += blockDim.x) hist[j] = 0;
    __syncthreads();
    const Raw* src = (const Raw*)(a.in + (it.offset + tl.start) * L::SS);
    Raw p[S::kPer];
#pragma unroll
    for (uint32_t u = 0; u < S::kPer; ++u) {
        const uint32_t j = threadIdx.x + u * kBuildThreads;
        if (j < tl.len) p[u] = src[j];
    }
#pragma unroll
    for (uint32_t u = 0; u < S::kPer; ++u) {
        const uint32_t j = threadIdx.x + u * kBuildThreads;
        if (j < tl.len) atomicAdd(&hist[first_slot<KS>(a, t.nbuckets, L::key(p[u])) / S::kRegion], 1u);
    }
    __syncthreads();
    const uint32_t g0 = a.reg0[it.table];
    for (uint32_t j = threadIdx.x; j < nbins; j += blockDim.x)
        if (hist[j]) atomicAdd(&a.counts[g0 + j], hist[j]);
}

// 2. offsets[0..n] = exclusive sum of counts[0..n) (one workgroup of 1024).
__global__ __launch_bounds__(1024) void build_scan_kernel(const uint32_t* counts, uint32_t* offsets, uint32_t n)
{
    typedef hipcub::BlockScan<uint32_t, 1024> Scan;
    __shared__ typename Scan::TempStorage tmp;
    const uint32_t per = (n + 4095) / 4096 * 4;  // a multiple of 4: 16-byte loads
    const uint32_t b = min(n, threadIdx.x * per), e = min(n, b + per);
    uint32_t local = 0;
    uint32_t i = b;
#pragma unroll 4
    for (; i + 4 <= e; i += 4) {
        const uint4 v = *(const uint4*)(counts + i);
        local += v.x + v.y + v.z + v.w;
    }
    for (; i < e; ++i) local += counts[i];
    uint32_t excl, total;
    Scan(tmp).ExclusiveSum(local, excl, total);
    for (uint32_t i = b; i < e; ++i) {
        offsets[i] = excl;
        excl += counts[i];
    }
    if (threadIdx.x == 0) offsets[n] = total;
}

// Dynamic LDS of the scatter kernel for tables of up to `maxbins` regions.
template <int SS>
constexpr size_t build_scatter_lds(uint32_t maxbins)
{
    return (size_t)BuildShape<SS>::kTile * SS + (size_t)BuildShape<SS>::kTile * 2 + (size_t)maxbins * 12;
}

// 3. One tile per workgroup: sort its pairs by region in LDS, claim a run in
// every region it touches (counts count down to zero) and write the runs.
template <int KS, int VS>
__global__ __launch_bounds__(kBuildThreads) void build_scatter_kernel(BuildArgs a)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    using S = BuildShape<L::SS>;
    extern __shared__ __align__(16) uint8_t smem[];
    Raw* staged = (Raw*)smem;                             // kTile
    uint32_t* tcnt = (uint32_t*)(staged + S::kTile);      // nbins
    uint32_t* toff = tcnt + a.maxbins;                    // nbins (first of each region in the tile)
    uint32_t* gpos = toff + a.maxbins;                    // nbins (run start in `part`)
    uint16_t* rid = (uint16_t*)(gpos + a.maxbins);        // kTile
    typedef hipcub::BlockScan<uint32_t, kBuildThreads> Scan;
    __shared__ typename Scan::TempStorage scan_tmp;

    const BatchChunk tl = a.tiles[blockIdx.x];
    const mxco_batch_item it = a.items[tl.item];
    const BatchTable t = a.tables[it.table];
    const uint32_t nbins = (t.cap + S::kRegion - 1) / S::kRegion;
    const uint32_t g0 = a.reg0[it.table];
    for (uint32_t j = threadIdx.x; j < nbins; j += blockDim.x) tcnt[j] = 0;
    __syncthreads();
    const Raw* src = (const Raw*)(a.in + (it.offset + tl.start) * L::SS);
    Raw p[S::kPer];
    uint32_t bin[S::kPer], rank[S::kPer];
#pragma unroll
    for (uint32_t u = 0; u < S::kPer; ++u) {
        const uint32_t j = threadIdx.x + u * kBuildThreads;
        if (j < tl.len) p[u] = src[j];
    }
#pragma unroll
    for (uint32_t u = 0; u < S::kPer; ++u) {
        const uint32_t j = threadIdx.x + u * kBuildThreads;
        if (j < tl.len) {
            bin[u] = first_slot<KS>(a, t.nbuckets, L::key(p[u])) / S::kRegion;
            rank[u] = atomicAdd(&tcnt[bin[u]], 1u);
        }
    }
    __syncthreads();
    const uint32_t span = (nbins + kBuildThreads - 1) / kBuildThreads;
    uint32_t local = 0;
    for (uint32_t q = 0; q < span; ++q) {
        const uint32_t r = threadIdx.x * span + q;
        if (r < nbins) local += tcnt[r];
    }
    uint32_t excl;
    Scan(scan_tmp).ExclusiveSum(local, excl);
    for (uint32_t q = 0; q < span; ++q) {
        const uint32_t r = threadIdx.x * span + q;
        if (r < nbins) {
            const uint32_t c = tcnt[r];
            toff[r] = excl;
            excl += c;
            if (c) gpos[r] = a.offsets[g0 + r] + atomicSub(&a.counts[g0 + r], c) - c;
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t u = 0; u < S::kPer; ++u) {
        const uint32_t j = threadIdx.x + u * kBuildThreads;
        if (j < tl.len) {
            const uint32_t pos = toff[bin[u]] + rank[u];
            staged[pos] = p[u];
            rid[pos] = (uint16_t)bin[u];
        }
    }
    __syncthreads();
    Raw* part = (Raw*)a.part;
    for (uint32_t pos = threadIdx.x; pos < tl.len; pos += blockDim.x) {
        const uint32_t r = rid[pos];
        part[gpos[r] + (pos - toff[r])] = staged[pos];
    }
}

// Compare-and-swap of an empty slot to pair p; returns the slot's key as it
// was (the empty key on success).  SS <= 8: the whole slot; 16-byte slots:
// the key, then the payload store (the reference's two-step insert).
template <int KS, int VS>
__device__ __forceinline__ uint64_t claim_slot(typename Lay<KS, VS>::Raw* s, typename Lay<KS, VS>::Raw expect,
                                               typename Lay<KS, VS>::Raw p, uint64_t ekey)
{
    using L = Lay<KS, VS>;
    if constexpr (L::SS == 16) {
        uint64_t old;
        if constexpr (KS == 4)
            old = atomicCAS((unsigned int*)s, (unsigned int)ekey, (unsigned int)L::key(p));
        else
            old = atomicCAS((unsigned long long*)s, (unsigned long long)ekey, (unsigned long long)L::key(p));
        if (old == (KS == 4 ? (ekey & 0xffffffffull) : ekey)) {
            if constexpr (VS == 4) *(uint32_t*)((uint8_t*)s + 8) = (uint32_t)p.y;
            else *(uint64_t*)((uint8_t*)s + 8) = p.y;
        }
        return old;
    } else if constexpr (L::SS == 8) {
        return L::key((typename L::Raw)atomicCAS((unsigned long long*)s, (unsigned long long)expect,
                                                 (unsigned long long)p));
    } else {
        return atomicCAS((unsigned int*)s, (unsigned int)expect, (unsigned int)p);
    }
}

// 4. One region per workgroup: build it in LDS and write it once.
template <int KS, int VS>
__global__ __launch_bounds__(kBuildThreads) void build_apply_kernel(BuildArgs a)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    constexpr uint32_t SS = L::SS;
    constexpr uint32_t R = BuildShape<SS>::kRegion;
    __shared__ __align__(16) Raw ls[R];
    __shared__ uint32_t s_nov, s_fail;
    const uint32_t r = blockIdx.x;
    const uint32_t ti = a.reg_table[r];
    const BatchTable t = a.tables[ti];
    const uint32_t r0 = (r - a.reg0[ti]) * R;
    const uint32_t rlen = min(R, t.cap - r0);
    const bool whole = rlen == t.cap;  // a one-region table: the walk wraps inside the region
    const uint64_t ekey = a.empty_key;
    const Raw empty = empty_slot<KS, VS>(ekey, a.empty_val);
    for (uint32_t j = threadIdx.x; j < rlen; j += blockDim.x) ls[j] = empty;
    if (threadIdx.x == 0) {
        s_nov = 0;
        s_fail = 0;
    }
    __syncthreads();
    const uint32_t s = a.offsets[r], e = a.offsets[r + 1];
    Raw* part = (Raw*)a.part;
    uint32_t fails = 0;
    for (uint32_t b = s; b < e; b += blockDim.x) {
        const uint32_t j = b + threadIdx.x;
        Raw p;
        if (j < e) p = part[j];
        // every pair of this round is read before any overflow pair is
        // written back over positions [s, b + blockDim.x)
        __syncthreads();
        if (j >= e) continue;
        const uint64_t key = L::key(p);
        if (key == ekey) {  // the reserved empty key is never inserted
            ++fails;
            continue;
        }
        uint32_t pos = first_slot<KS>(a, t.nbuckets, key) - r0;
        uint32_t left = t.cap;
        for (;;) {
            const Raw v = ls[pos];
            const uint64_t k = L::key(v);
            if (k == ekey) {
                if (claim_slot<KS, VS>(&ls[pos], v, p, ekey) == ekey) break;  // inserted
                continue;  // lost the slot: examine it again
            }
            if (key_eq<KS>(k, key, a.key_float)) {
                ++fails;  // duplicate
                break;
            }
            if (--left == 0) {
                ++fails;  // a one-region table with no empty slot
                break;
            }
            if (++pos == rlen) {
                if (!whole) {
                    part[s + atomicAdd(&s_nov, 1u)] = p;  // finished by the overflow kernel
                    break;
                }
                pos = 0;
            }
        }
    }
    if (fails) atomicAdd(&s_fail, fails);
    __syncthreads();
    uint8_t* dst = t.slots + (uint64_t)r0 * SS;
    const uint32_t bytes = rlen * SS, n16 = bytes / 16;
    for (uint32_t j = threadIdx.x; j < n16; j += blockDim.x) ((u64x2*)dst)[j] = ((const u64x2*)ls)[j];
    for (uint32_t j = n16 * 16 / SS + threadIdx.x; j < rlen; j += blockDim.x) ((Raw*)dst)[j] = ls[j];
    if (threadIdx.x == 0) {
        a.ovc[r] = s_nov;
        const int32_t item = a.table_item[ti];
        if (s_fail && item >= 0) atomicAdd(&a.counters[2ull * item], (unsigned long long)s_fail);
    }
}

// 5. Keys whose walk left their region: the CAS walk on the table, resumed
// at the slot after the region.  One wavefront per region.
template <int KS, int VS>
__global__ __launch_bounds__(256) void build_overflow_kernel(BuildArgs a)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    constexpr uint32_t SS = L::SS;
    constexpr uint32_t R = BuildShape<SS>::kRegion;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    const uint64_t ekey = a.empty_key;
    for (uint32_t r = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; r < a.nreg; r += waves) {
        const uint32_t n = a.ovc[r];
        if (n == 0) continue;
        const uint32_t ti = a.reg_table[r];
        const BatchTable t = a.tables[ti];
        const uint32_t r0 = (r - a.reg0[ti]) * R;
        const uint32_t rend = r0 + min(R, t.cap - r0);
        const Raw* part = (const Raw*)a.part + a.offsets[r];
        uint32_t fails = 0;
        for (uint32_t k = lane; k < n; k += 64) {
            const Raw p = part[k];
            const uint64_t key = L::key(p);
            const uint32_t first = first_slot<KS>(a, t.nbuckets, key);
            uint32_t left = t.cap - (rend - first);  // slots of the walk not yet examined
            uint32_t pos = rend == t.cap ? 0 : rend;
            bool done = false;
            while (left > 0) {
                Raw* slot = (Raw*)(t.slots + (uint64_t)pos * SS);
                const Raw v = *slot;
                uint64_t kk = L::key(v);
                if (kk == ekey) {
                    kk = claim_slot<KS, VS>(slot, v, p, ekey);
                    if (kk == ekey) {
                        done = true;
                        break;
                    }
                    // lost: the slot now holds kk
                }
                if (key_eq<KS>(kk, key, a.key_float)) {
                    ++fails;
                    done = true;
                    break;
                }
                --left;
                if (++pos == t.cap) pos = 0;
            }
            if (!done) ++fails;  // no empty slot on the walk
        }
        // one add per wavefront
        for (int o = 32; o > 0; o >>= 1) fails += __shfl_xor(fails, o);
        const int32_t item = a.table_item[ti];
        if (lane == 0 && fails && item >= 0) atomicAdd(&a.counters[2ull * item], (unsigned long long)fails);
    }
}

}  // namespace mxco
