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
//
// The apply and overflow kernels also serve Insert into one large table
// (oa_api.hip), after a two-level partition of its own (the bounded scatter
// and regroup at the end of this file: the table's regions outnumber what one
// LDS histogram holds).  The apply then loads each region from the table (the
// table may hold keys and tombstones already, unless it is known to be
// empty).  An empty region of at most kPlaceMax pairs with integer keys is
// placed without walks (place_sorted); others walk as above.
#pragma once
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <type_traits>

#include "batch_kernels.h"
#include "mxco/detail/oa_kernels.h"

namespace mxco {

constexpr uint32_t kBuildBytes = 65536;    // most slot bytes of one region (held in LDS)
constexpr uint32_t kRegionSlots = 4096;    // most slots of one region
constexpr uint32_t kBuildThreads = 1024;   // count / scatter / apply workgroups
constexpr uint32_t kBuildMaxBins = 4096;   // bins of one table in count / scatter

// Pairs of one scatter / regroup tile: 64 KiB of pairs.  (Halving it for the
// narrower slots lets two of their scatter workgroups share a CU, but doubles
// the count pass's tiles and per-bin adds: a net loss.)
constexpr uint32_t build_tile_pairs(uint32_t ss)
{
    return 65536u / ss;
}

// Slots of one region for slot size ss (BuildShape<ss>::kRegion on the host).
constexpr uint32_t build_region_slots(uint32_t ss)
{
    return kBuildBytes / ss < kRegionSlots ? kBuildBytes / ss : kRegionSlots;
}

template <int SS>
struct BuildShape {
    static constexpr uint32_t kRegion = kBuildBytes / SS < kRegionSlots ? kBuildBytes / SS : kRegionSlots;  // slots per region
    static constexpr uint32_t kTile = build_tile_pairs(SS);   // pairs per tile
    static constexpr uint32_t kPer = kTile / kBuildThreads;   // pairs per thread of a tile
};

struct BuildArgs {
    const uint8_t* in;  // the plan's input pairs (slot layout)
    uint8_t* part;      // the pairs grouped by bin (total * SS bytes)
    const BatchTable* tables;
    const mxco_batch_item* items;
    const BatchChunk* tiles;    // batched plans: tiles of one item each, at most kTile pairs (the
                                // one-table Insert cuts pairs [0, single_n) of `in` into tiles)
    const uint32_t* reg0;       // first bin of each table (num_tables + 1)
    const uint32_t* reg_table;  // batched plans: table of each region (the one-table Insert: table 0)
    const int32_t* table_item;  // item of each table, or -1 (null: item 0)
    uint32_t* counts;           // per bin; zero again after the scatter
    uint32_t* offsets;          // exclusive sum of the counts (nbins + 1)
    uint32_t* ovc;              // overflow pairs of each region, kept at its offset
    uint32_t* apply_end;        // one-table: end of each region's pairs for the apply
    unsigned long long* counters;  // 2 per item: [0] failed inserts (null: not counted)
    uint64_t empty_key, empty_val;
    uint64_t single_n;
    uint32_t nreg;
    uint32_t maxbins;     // most bins of a table with an item (stride of the scatter's LDS arrays)
    uint32_t bin_slots;   // one-table Insert: slots per first-level bin (a multiple of the region;
                          // batched plans bin by region)
    uint32_t bucket;
    uint32_t hash;
    uint32_t key_float;
};

// The tile of workgroup b: its table, its first pair and its length.
struct TileRef {
    uint32_t table;
    const uint8_t* src;
    uint32_t len;
};

template <int SS, bool kOne>
__device__ __forceinline__ TileRef tile_ref(const BuildArgs& a, uint32_t b)
{
    constexpr uint32_t kTile = BuildShape<SS>::kTile;
    if constexpr (!kOne) {
        const BatchChunk tl = a.tiles[b];
        const mxco_batch_item it = a.items[tl.item];
        return TileRef{it.table, a.in + (it.offset + tl.start) * SS, tl.len};
    }
    const uint64_t st = (uint64_t)b * kTile;
    return TileRef{0u, a.in + st * SS, (uint32_t)min<uint64_t>(kTile, a.single_n - st)};
}

// Failures of a table: into its item's counter (batched plans), or into one
// of the table's counter slots (the one-table Insert: table_item is null).
__device__ __forceinline__ void count_failures(const BuildArgs& a, uint32_t table, uint32_t n)
{
    if (!n || !a.counters) return;
    if (!a.table_item) {
        atomicAdd(counter_slot(a.counters), (unsigned long long)n);
        return;
    }
    const int32_t item = a.table_item[table];
    if (item >= 0) atomicAdd(&a.counters[2ull * item], (unsigned long long)n);
}

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

// kOne: the one-table Insert (implicit tiles, bins of a.bin_slots slots, a
// two-level partition); otherwise a batched plan (bins are regions).

// 1. Region histogram of one tile per workgroup.
template <int KS, int VS, bool kOne>
__global__ __launch_bounds__(kBuildThreads) void build_count_kernel(BuildArgs a)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    using S = BuildShape<L::SS>;
    __shared__ uint32_t hist[kBuildMaxBins];
    // a batched plan: one tile per workgroup (tiles of different tables);
    // one table: tiles blockIdx.x, blockIdx.x + gridDim.x, ... into one
    // histogram, added once (every tile shares the bins, so per-tile adds
    // would contend on the same counters)
    const uint32_t ntiles = kOne ? (uint32_t)((a.single_n + S::kTile - 1) / S::kTile) : blockIdx.x + 1;
    const BatchTable t = a.tables[kOne ? 0u : tile_ref<L::SS, kOne>(a, blockIdx.x).table];
    const uint32_t bs = kOne ? a.bin_slots : S::kRegion;
    const uint32_t nbins = (t.cap + bs - 1) / bs;
    for (uint32_t j = threadIdx.x; j < nbins; j += blockDim.x) hist[j] = 0;
    __syncthreads();
    uint32_t table = 0;
    for (uint32_t b = blockIdx.x; b < ntiles; b += kOne ? gridDim.x : 1u) {
        const TileRef tl = tile_ref<L::SS, kOne>(a, b);
        table = tl.table;
        const Raw* src = (const Raw*)tl.src;
        Raw p[S::kPer];
#pragma unroll
        for (uint32_t u = 0; u < S::kPer; ++u) {
            const uint32_t j = threadIdx.x + u * kBuildThreads;
            if (j < tl.len) p[u] = src[j];
        }
#pragma unroll
        for (uint32_t u = 0; u < S::kPer; ++u) {
            const uint32_t j = threadIdx.x + u * kBuildThreads;
            if (j < tl.len) atomicAdd(&hist[first_slot<KS>(a, t.nbuckets, L::key(p[u])) / bs], 1u);
        }
    }
    __syncthreads();
    const uint32_t g0 = a.reg0[table];
    for (uint32_t j = threadIdx.x; j < nbins; j += blockDim.x)
        if (hist[j]) atomicAdd(&a.counts[g0 + j], hist[j]);
}

// 2. offsets[0..n] = exclusive sum of counts[0..n) (one workgroup of 1024).
static __global__ __launch_bounds__(1024) void build_scan_kernel(const uint32_t* counts, uint32_t* offsets, uint32_t n)
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
template <int KS, int VS, bool kOne>
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
    // The scan's storage shares `staged`, which is first written after the
    // scan and the barrier behind it: with no static LDS a workgroup of
    // 16-byte pairs over 512 bins needs 78 KiB, and two fit on a CU.
    static_assert(sizeof(typename Scan::TempStorage) <= S::kTile * sizeof(Raw), "scan storage exceeds the tile");
    typename Scan::TempStorage& scan_tmp = *reinterpret_cast<typename Scan::TempStorage*>(smem);

    const TileRef tl = tile_ref<L::SS, kOne>(a, blockIdx.x);
    const BatchTable t = a.tables[tl.table];
    const uint32_t bs = kOne ? a.bin_slots : S::kRegion;
    const uint32_t nbins = (t.cap + bs - 1) / bs;
    const uint32_t g0 = a.reg0[tl.table];
    for (uint32_t j = threadIdx.x; j < nbins; j += blockDim.x) tcnt[j] = 0;
    __syncthreads();
    const Raw* src = (const Raw*)tl.src;
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
            bin[u] = first_slot<KS>(a, t.nbuckets, L::key(p[u])) / bs;
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

// Placement without walks (an empty region of at most kPlaceMax slots, at
// most kPlaceMax pairs, integer keys; the sort scratch, keys of KS bytes and
// two 16-bit words per pair, must fit the region's LDS).  Inserting keys one after another in
// the order of their first slots, each walk ends at the first slot after the
// previous key's that is not below its own first slot: the i-th kept key
// lands at slot i + max_{j <= i}(first_j - j), a prefix maximum.  The pairs
// are counting-sorted by first slot in LDS, and the second and later copies
// of a key within its first-slot group are duplicates (equal integer keys
// share a first slot).  A key placed past the region's end goes to the
// overflow kernel.  Returns false, with ls refilled with the empty slot, when
// a first-slot group exceeds kPlaceGroup pairs; the caller then walks.
constexpr uint32_t kPlaceMax = 4096;

template <int KS, int SS>
constexpr bool place_fits()
{
    return BuildShape<SS>::kRegion <= kPlaceMax && kPlaceMax * (KS + 2 + 2) <= BuildShape<SS>::kRegion * SS;
}
constexpr uint32_t kPlaceGroup = 64;

// The region's pairs part[s + tid + u * kBuildThreads], u < kPer, below e,
// into registers (place_sorted's layout).
template <int KS, int VS>
__device__ __forceinline__ void load_region_pairs(const typename Lay<KS, VS>::Raw* part, uint32_t s, uint32_t e,
                                                  typename Lay<KS, VS>::Raw (&p)[kPlaceMax / kBuildThreads])
{
#pragma unroll
    for (uint32_t u = 0; u < kPlaceMax / kBuildThreads; ++u) {
        const uint32_t i = s + threadIdx.x + u * kBuildThreads;
        if (i < e) p[u] = part[i];
    }
}

template <int KS, int VS>
__device__ __forceinline__ bool place_sorted(const BuildArgs& a, const BatchTable& t, uint32_t r0, uint32_t rlen,
                                             uint32_t s, uint32_t e, typename Lay<KS, VS>::Raw* ls,
                                             typename Lay<KS, VS>::Raw* part,
                                             const typename Lay<KS, VS>::Raw (&pin)[kPlaceMax / kBuildThreads],
                                             uint32_t& s_nov, uint32_t& fails)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    constexpr uint32_t SS = L::SS;
    constexpr uint32_t kPer = kPlaceMax / kBuildThreads;  // pairs (and sorted positions) per thread
    using SKey = typename std::conditional<KS == 4, uint32_t, uint64_t>::type;
    static_assert(place_fits<KS, SS>(), "sort scratch fits the region's LDS");
    typedef hipcub::BlockScan<int32_t, kBuildThreads> Scan;
    __shared__ typename Scan::TempStorage scan_tmp;
    __shared__ uint32_t hcnt[kPlaceMax / 2];  // two 16-bit counts per word: pairs per first slot, then group starts
    __shared__ uint32_t s_big;
    SKey* skey = (SKey*)ls;                       // sorted keys
    uint16_t* shome = (uint16_t*)(skey + kPlaceMax);  // sorted first slots | 0x8000 for a kept pair
    uint16_t* sq = shome + kPlaceMax;             // slot of each kept sorted pair, 0xffff past the region
    const uint64_t ekey = a.empty_key;
    const uint32_t tid = threadIdx.x;

    Raw p[kPer];
    uint32_t home[kPer], rank[kPer], si[kPer];
    bool valid[kPer];
    for (uint32_t j = tid; j < kPlaceMax / 2; j += kBuildThreads) hcnt[j] = 0;
    if (tid == 0) s_big = 0;
    __syncthreads();
#pragma unroll
    for (uint32_t u = 0; u < kPer; ++u) {
        const uint32_t i = s + tid + u * kBuildThreads;
        valid[u] = false;
        if (i < e) {
            p[u] = pin[u];  // loaded by the caller (load_region_pairs)
            const uint64_t key = L::key(p[u]);
            if (key == ekey) {
                ++fails;  // the reserved empty key is never inserted
            } else {
                valid[u] = true;
                home[u] = first_slot<KS>(a, t.nbuckets, key) - r0;
                const uint32_t sh = (home[u] & 1u) * 16u;
                rank[u] = (atomicAdd(&hcnt[home[u] >> 1], 1u << sh) >> sh) & 0xffffu;
            }
        }
    }
    __syncthreads();
    // group starts: exclusive sum of the counts, kPer first slots per thread;
    // `total` = pairs sorted (all but the reserved key)
    uint32_t total;
    {
        uint32_t c[kPer], sum = 0, big = 0;
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t h = tid * kPer + k;
            c[k] = (hcnt[h >> 1] >> ((h & 1u) * 16u)) & 0xffffu;
            sum += c[k];
            big |= c[k] > kPlaceGroup;
        }
        int32_t excl, agg;
        Scan(scan_tmp).ExclusiveSum((int32_t)sum, excl, agg);
        total = (uint32_t)agg;
        if (big) s_big = 1;
        __syncthreads();  // every count read before the starts overwrite them
        uint32_t o = (uint32_t)excl;
#pragma unroll
        for (uint32_t k = 0; k < kPer; k += 2) {
            hcnt[(tid * kPer + k) >> 1] = o | (o + c[k]) << 16;
            o += c[k] + c[k + 1];
        }
    }
    __syncthreads();
    if (s_big) {
        const Raw empty = empty_slot<KS, VS>(ekey, a.empty_val);
        for (uint32_t j = tid; j < rlen; j += kBuildThreads) ls[j] = empty;
        __syncthreads();
        fails = 0;
        return false;
    }
#pragma unroll
    for (uint32_t u = 0; u < kPer; ++u) {
        if (!valid[u]) continue;
        const uint32_t h = home[u];
        si[u] = ((hcnt[h >> 1] >> ((h & 1u) * 16u)) & 0xffffu) + rank[u];
        skey[si[u]] = (SKey)L::key(p[u]);
    }
    __syncthreads();
    // duplicates: an equal key earlier in the same first-slot group
#pragma unroll
    for (uint32_t u = 0; u < kPer; ++u) {
        if (!valid[u]) continue;
        const uint32_t h = home[u];
        const uint32_t g0 = (hcnt[h >> 1] >> ((h & 1u) * 16u)) & 0xffffu;
        const SKey key = skey[si[u]];
        bool dup = false;
        for (uint32_t j = g0; j < si[u]; ++j) dup |= skey[j] == key;
        shome[si[u]] = (uint16_t)(h | (dup ? 0u : 0x8000u));
        if (dup) {
            ++fails;
            valid[u] = false;
        }
    }
    __syncthreads();
    // kept pairs in sorted order: rank K (exclusive count) and the running
    // maximum of first - K, kPer sorted positions per thread
    {
        uint32_t kept[kPer], hh[kPer], nk = 0;
        int32_t lmax = INT32_MIN;
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t j = tid * kPer + k;
            const uint16_t v = j < total ? shome[j] : 0;
            kept[k] = v >> 15;
            hh[k] = v & 0x7fffu;
            if (kept[k]) lmax = max(lmax, (int32_t)hh[k] - (int32_t)nk);  // relative to the thread's first K
            nk += kept[k];
        }
        int32_t kx;
        Scan(scan_tmp).ExclusiveSum((int32_t)nk, kx);
        __syncthreads();
        // the thread's own maximum, shifted by its first K
        const int32_t tmax = lmax == INT32_MIN ? INT32_MIN : lmax - kx;
        int32_t mx;
        Scan(scan_tmp).ExclusiveScan(tmax, mx, INT32_MIN, hipcub::Max());
        int32_t K = kx;
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t j = tid * kPer + k;
            if (j < total && kept[k]) {
                mx = max(mx, (int32_t)hh[k] - K);
                const int32_t q = K + mx;
                sq[j] = q < (int32_t)rlen ? (uint16_t)q : (uint16_t)0xffffu;
                ++K;
            }
        }
    }
    __syncthreads();
    uint32_t q[kPer];
#pragma unroll
    for (uint32_t u = 0; u < kPer; ++u) q[u] = valid[u] ? sq[si[u]] : 0xffffu;
    __syncthreads();
    const Raw empty = empty_slot<KS, VS>(ekey, a.empty_val);
    for (uint32_t j = tid; j < rlen; j += kBuildThreads) ls[j] = empty;
    __syncthreads();
#pragma unroll
    for (uint32_t u = 0; u < kPer; ++u) {
        if (!valid[u]) continue;
        if (q[u] != 0xffffu) ls[q[u]] = p[u];
        else part[s + atomicAdd(&s_nov, 1u)] = p[u];  // every pair was read before the first barrier
    }
    return true;
}

// The CAS walk of the region's pairs part[s, e) in LDS (the rules of
// open_addressing_ref_impl.h:490-532): each pair walks from its first slot to
// the first empty slot or equal key; a walk past the region's end sends the
// pair to the overflow kernel (back into part[s, ...)), and a one-region
// table wraps.  Entered after a barrier.
template <int KS, int VS>
__device__ __forceinline__ void apply_walk(const BuildArgs& a, const BatchTable& t, uint32_t r0, uint32_t rlen,
                                           bool whole, uint32_t s, uint32_t e, typename Lay<KS, VS>::Raw* ls,
                                           typename Lay<KS, VS>::Raw* part, uint32_t& s_nov, uint32_t& fails)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    const uint64_t ekey = a.empty_key;
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
}

// 4. One region per workgroup: build it in LDS and write it once.
// kLoad: start from the table's slots (otherwise from the empty slot: a
// batched plan, or a table known to be empty).
template <int KS, int VS, bool kOne, bool kLoad>
__global__ __launch_bounds__(kBuildThreads) void build_apply_kernel(BuildArgs a)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    constexpr uint32_t SS = L::SS;
    constexpr uint32_t R = BuildShape<SS>::kRegion;
    __shared__ __align__(16) Raw ls[R];
    __shared__ uint32_t s_nov, s_fail;
    const uint32_t r = blockIdx.x;
    const uint32_t ti = kOne ? 0u : a.reg_table[r];
    const BatchTable t = a.tables[ti];
    const uint32_t r0 = (r - a.reg0[ti]) * R;
    const uint32_t rlen = min(R, t.cap - r0);
    const bool whole = rlen == t.cap;  // a one-region table: the walk wraps inside the region
    const uint32_t s = a.offsets[r], e = kOne ? a.apply_end[r] : a.offsets[r + 1];
    const uint64_t ekey = a.empty_key;
    uint8_t* dst = t.slots + (uint64_t)r0 * SS;
    const uint32_t bytes = rlen * SS, n16 = bytes / 16;
    // regions of an empty table placed without walks are filled by place_sorted
    constexpr bool kPlace = place_fits<KS, SS>();
    const bool placing = !kLoad && kPlace && !a.key_float && e - s <= kPlaceMax;
    if constexpr (kLoad) {
        for (uint32_t j = threadIdx.x; j < n16; j += blockDim.x) ((u64x2*)ls)[j] = ((const u64x2*)dst)[j];
        for (uint32_t j = n16 * 16 / SS + threadIdx.x; j < rlen; j += blockDim.x) ls[j] = ((const Raw*)dst)[j];
    } else if (!placing) {
        const Raw empty = empty_slot<KS, VS>(ekey, a.empty_val);
        for (uint32_t j = threadIdx.x; j < rlen; j += blockDim.x) ls[j] = empty;
    }
    if (threadIdx.x == 0) {
        s_nov = 0;
        s_fail = 0;
    }
    Raw* part = (Raw*)a.part;
    uint32_t fails = 0;
    if constexpr (place_fits<KS, SS>()) {
        // an empty region, integer keys: placed without walks
        bool empty_region = true;
        if constexpr (kLoad)
            for (uint32_t j = threadIdx.x; j < rlen; j += blockDim.x) empty_region &= L::key(ls[j]) == ekey;
        Raw pin[kPlaceMax / kBuildThreads];
        const bool try_place = empty_region && !a.key_float && e - s <= kPlaceMax;
        if (try_place) load_region_pairs<KS, VS>(part, s, e, pin);
        if (__syncthreads_and(try_place) && place_sorted<KS, VS>(a, t, r0, rlen, s, e, ls, part, pin, s_nov, fails)) {
            if (fails) atomicAdd(&s_fail, fails);
            __syncthreads();
            for (uint32_t j = threadIdx.x; j < n16; j += blockDim.x) ((u64x2*)dst)[j] = ((const u64x2*)ls)[j];
            for (uint32_t j = n16 * 16 / SS + threadIdx.x; j < rlen; j += blockDim.x) ((Raw*)dst)[j] = ls[j];
            if (threadIdx.x == 0) {
                a.ovc[r] = s_nov;
                count_failures(a, ti, s_fail);
            }
            return;
        }
    }
    __syncthreads();
    apply_walk<KS, VS>(a, t, r0, rlen, whole, s, e, ls, part, s_nov, fails);
    if (fails) atomicAdd(&s_fail, fails);
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < n16; j += blockDim.x) ((u64x2*)dst)[j] = ((const u64x2*)ls)[j];
    for (uint32_t j = n16 * 16 / SS + threadIdx.x; j < rlen; j += blockDim.x) ((Raw*)dst)[j] = ls[j];
    if (threadIdx.x == 0) {
        a.ovc[r] = s_nov;
        count_failures(a, ti, s_fail);
    }
}

// The CAS walk of pair p (key `key`) on table t from slot `pos` over `left`
// slots (open_addressing_ref_impl.h:490-532): the first slot that is empty
// or holds the key ends it.  Returns true when the pair was not inserted
// (a duplicate, or no empty slot on the walk).
template <int KS, int VS>
__device__ __forceinline__ bool cas_walk(const BuildArgs& a, const BatchTable& t, typename Lay<KS, VS>::Raw p,
                                         uint64_t key, uint32_t pos, uint32_t left)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    constexpr uint32_t SS = L::SS;
    const uint64_t ekey = a.empty_key;
    while (left > 0) {
        Raw* slot = (Raw*)(t.slots + (uint64_t)pos * SS);
        const Raw v = *slot;
        uint64_t kk = L::key(v);
        if (kk == ekey) {
            kk = claim_slot<KS, VS>(slot, v, p, ekey);
            if (kk == ekey) return false;
            // lost: the slot now holds kk
        }
        if (key_eq<KS>(kk, key, a.key_float)) return true;
        --left;
        if (++pos == t.cap) pos = 0;
    }
    return true;
}

// A slot value of lane `src` of the wavefront.
template <typename Raw>
__device__ __forceinline__ Raw shfl_raw(Raw v, uint32_t src)
{
    if constexpr (sizeof(Raw) == 16) {
        Raw r;
        r.x = __shfl(v.x, (int)src);
        r.y = __shfl(v.y, (int)src);
        return r;
    } else if constexpr (sizeof(Raw) == 8) {
        return (Raw)__shfl((unsigned long long)v, (int)src);
    } else {
        return (Raw)__shfl((unsigned int)v, (int)src);
    }
}

// 5. Keys whose walk left their region: the CAS walk on the table, resumed
// at the slot after the region.  One wavefront per region, whose keys all
// resume at the same slot, walks them together line by line: lane j loads
// slot j of the line, and the wavefront inserts its keys as the sequential
// walk inserts them in lane order.  A key found in the line is a duplicate
// (no empty slot precedes it there: the key was inserted by a walk that
// found every earlier slot full, and slots are never emptied during an
// Insert); the other keys take the line's empty slots in lane order.  A
// compare-and-swap lost to another wavefront's key makes that slot full, and
// the keys still without a slot take the line's remaining empty slots: the
// result is the sequential walk's for the order that puts the other key
// first.  Keys of one region that are equal are inserted once, by the lowest
// lane.  Each key keeps its walk's slot budget (a full cycle of the table
// from its first slot).
template <int KS, int VS, bool kOne>
__global__ __launch_bounds__(256) void build_overflow_kernel(BuildArgs a)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    constexpr uint32_t SS = L::SS;
    constexpr uint32_t R = BuildShape<SS>::kRegion;
    constexpr uint32_t SPL = 128 / SS;  // slots of a 128-byte line (at most 32)
    static_assert(SPL <= 64, "a line's slots fit a wavefront");
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t ekey = a.empty_key;
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    for (uint32_t r = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; r < a.nreg; r += waves) {
        const uint32_t n = a.ovc[r];
        if (n == 0) continue;
        const uint32_t ti = kOne ? 0u : a.reg_table[r];
        const BatchTable t = a.tables[ti];
        const uint32_t r0 = (r - a.reg0[ti]) * R;
        const uint32_t rend = r0 + min(R, t.cap - r0);
        const uint32_t start = rend == t.cap ? 0u : rend;
        const Raw* part = (const Raw*)a.part + a.offsets[r];
        uint32_t fails = 0;
        for (uint32_t b0 = 0; b0 < n; b0 += 64) {
            const bool in = b0 + lane < n;
            Raw p{};
            uint64_t key = 0;
            uint32_t left = 0;  // slots this key's walk may still examine
            if (in) {
                p = part[b0 + lane];
                key = L::key(p);
                const uint32_t first = first_slot<KS>(a, t.nbuckets, key);
                // the region's slots from `first` on are full and hold other keys
                left = t.cap - (rend - first);
            }
            // a key equal to a lower lane's is a duplicate
            const uint64_t inmask = __ballot(in);
            bool dup = false;
            for (uint32_t j = 0; j < 64 && (inmask >> j); ++j) {
                const uint64_t kj = __shfl(key, (int)j);
                if (j < lane && ((inmask >> j) & 1ull) && key_eq<KS>(kj, key, a.key_float)) dup = true;
            }
            bool active = in && !dup;
            fails += in && dup;
            uint32_t pos = start;
            while (__ballot(active)) {
                const uint32_t wlen = min(SPL - pos % SPL, t.cap - pos);
                Raw* line = (Raw*)(t.slots + (uint64_t)pos * SS);
                Raw sv{};
                if (lane < wlen) sv = line[lane];
                uint64_t empt = __ballot(lane < wlen && L::key(sv) == ekey);
                // slots of the line this key may examine, and the first holding it
                const uint32_t lim = active ? min(wlen, left) : 0u;
                uint32_t eq = 64;
                for (uint32_t j = 0; j < wlen; ++j) {
                    const uint64_t kj = L::key(shfl_raw(sv, j));
                    if (j < lim && eq == 64 && kj != ekey && key_eq<KS>(kj, key, a.key_float)) eq = j;
                }
                if (active && eq < 64) {
                    ++fails;  // already in the table
                    active = false;
                }
                // the other keys take the empty slots in lane order, until the line
                // has none left or every key has a slot
                for (;;) {
                    uint64_t rem = __ballot(active);
                    if (!rem || !empt) break;
                    uint32_t mine = 64;
                    uint64_t em = empt;
                    while (rem && em) {
                        const uint32_t l = (uint32_t)__builtin_ctzll(rem);
                        rem &= rem - 1;
                        const uint32_t e = (uint32_t)__builtin_ctzll(em);
                        if (e < (uint32_t)__builtin_amdgcn_readlane((int)lim, (int)l)) {
                            em &= em - 1;  // lane l takes slot e
                            if (lane == l) mine = e;
                        }
                    }
                    if (!__ballot(mine < 64)) break;  // the empty slots lie past every budget
                    uint64_t kk = 0;
                    const Raw ev = shfl_raw(sv, mine < 64 ? mine : 0u);
                    if (mine < 64) kk = claim_slot<KS, VS>(&line[mine], ev, p, ekey);
                    // the claimed slots are full now, won or lost
                    const uint64_t took = __ballot(mine < 64);
                    uint64_t taken = 0;
                    for (uint64_t tk = took; tk; tk &= tk - 1) {
                        const uint32_t l = (uint32_t)__builtin_ctzll(tk);
                        taken |= 1ull << (uint32_t)__builtin_amdgcn_readlane((int)mine, (int)l);
                    }
                    empt &= ~taken;
                    if (mine < 64) {
                        if (kk == ekey) {
                            active = false;  // inserted
                        } else if (key_eq<KS>(kk, key, a.key_float)) {
                            ++fails;  // another wavefront inserted it first
                            active = false;
                        }
                    }
                }
                // keys whose budget ends in this line: no empty slot on their walk
                if (active && left <= wlen) {
                    ++fails;
                    active = false;
                }
                left -= active ? wlen : 0u;
                pos += wlen;
                if (pos == t.cap) pos = 0;
            }
        }
        // one add per wavefront
        for (int o = 32; o > 0; o >>= 1) fails += __shfl_xor(fails, o);
        if (lane == 0) count_failures(a, ti, fails);
    }
}

// Exclusive prefix sums of cnt[0 .. n) into off[0 .. n), n <= 256, by the 64
// lanes of wavefront 0 (four entries a lane, then a shuffle scan of the lane
// sums); the other wavefronts do not call it.
__device__ __forceinline__ void wave0_offsets(const uint32_t* cnt, uint32_t* off, uint32_t n)
{
    const uint32_t lane = threadIdx.x;
    uint32_t v[4], sum = 0;
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
        const uint32_t j = lane * 4 + q;
        v[q] = j < n ? cnt[j] : 0u;
        sum += v[q];
    }
    uint32_t incl = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o);
        if (lane >= (uint32_t)o) incl += y;
    }
    uint32_t run = incl - sum;
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
        const uint32_t j = lane * 4 + q;
        if (j < n) off[j] = run;
        run += v[q];
    }
}

// ---------------------------------------------------------------------------
// One-table Insert with bounded bins: no count pass.
//
// The pairs go to first-level bins of `group` regions, then to their
// regions.  Every bin is cut into one slice per XCD, appended to by the
// workgroups of that XCD (so the partial lines of their runs meet in that
// XCD's L2), and every slice and region holds at most a fixed number of pairs
// (capS, capR), set from the batch size with room to spare for a uniform
// hash, so no pass needs a count of the pairs first.  Pairs
// that find their slice or region full go to one spill list and are
// inserted by the CAS walk after every region is written
// (build_spill_kernel), so the final table is still one that the CAS walk
// produces for some order of the inserts: the regions' keys first, then
// their walks past the region's end, then the spilled pairs.  Each pass reads
// and writes every pair once:
//   scatter: pairs -> part1 (bin b, slice x at (b * G + x) * capS), tiles
//            sorted by bin in LDS, a run per bin reserved with one atomic on
//            the slice's fill count;
//   regroup: bin b's G slices -> part2 (region r at r * capR), tiles sorted
//            by region in LDS, runs placed on the bin's LDS cursors;
//   apply:   region by region, as above (offsets[r] = r * capR).
// ---------------------------------------------------------------------------

constexpr uint32_t kBuildMaxSlices = 8;  // slices per bin: one per XCD (blockIdx.x % 8)

struct SpillArgs {
    uint32_t* fill;       // pairs reserved in each slice: fill[b * G + x] (nbins1 * G)
    uint8_t* spill;       // pairs that found no room (at most the batch)
    uint32_t* nspill;     // their count
    uint32_t cap_slice;   // capS
    uint32_t cap_region;  // capR
    uint32_t slices;      // G
};

// Slice and region capacities for n pairs over nbins1 bins of `slices`
// slices and nreg regions.
__host__ __device__ inline uint32_t build_cap_slice(uint64_t n, uint32_t nbins1, uint32_t slices)
{
    const uint64_t q = (n + (uint64_t)nbins1 * slices - 1) / ((uint64_t)nbins1 * slices);
    return (uint32_t)(q + q / 16 + 1024);
}
__host__ __device__ inline uint32_t build_cap_region(uint64_t n, uint32_t nreg)
{
    const uint64_t q = (n + nreg - 1) / nreg;
    return (uint32_t)(q + q / 4 + 256);
}

// Dynamic LDS of build_bounded_scatter_kernel: the staged tile and a bin id
// per pair; per bin the tile's count and the run's start in part1 (32-bit),
// and the bin's first position in the sorted tile, the pairs that fit the
// slice and the spilled pairs' start in the tile's spill run (16-bit: at most
// a tile).
template <int SS>
constexpr size_t build_bounded_scatter_lds(uint32_t nbins)
{
    return (size_t)BuildShape<SS>::kTile * SS + (size_t)nbins * 8 + (size_t)nbins * 6 + (size_t)BuildShape<SS>::kTile * 2;
}

// Most bins of the bounded scatter with slot size ss: two workgroups per CU.
constexpr uint32_t build_bounded_max_bins(uint32_t ss)
{
    const uint32_t tile = build_tile_pairs(ss);
    return (81920u - 64u - tile * ss - tile * 2) / 14;
}

// Scatter of the one-table Insert: persistent workgroups take tiles
// blockIdx.x, blockIdx.x + gridDim.x, ...; the next tile's pairs are loaded
// while the current one is sorted and written.
template <int KS, int VS>
__global__ __launch_bounds__(kBuildThreads) void build_bounded_scatter_kernel(BuildArgs a, SpillArgs sp)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    using S = BuildShape<L::SS>;
    extern __shared__ __align__(16) uint8_t smem[];
    Raw* staged = (Raw*)smem;                           // kTile
    uint32_t* tcnt = (uint32_t*)(staged + S::kTile);    // pairs of the tile in each bin
    uint32_t* base = tcnt + a.maxbins;                  // run start in part1
    uint16_t* toff = (uint16_t*)(base + a.maxbins);     // first of each bin in the sorted tile
    uint16_t* keep = toff + a.maxbins;                  // pairs of the run that fit the bin
    uint16_t* sbase = keep + a.maxbins;                 // start of the bin's spilled pairs in the tile's spill run
    uint16_t* rid = sbase + a.maxbins;                  // kTile
    typedef hipcub::BlockScan<uint32_t, kBuildThreads> Scan;
    static_assert(sizeof(typename Scan::TempStorage) <= S::kTile * sizeof(Raw), "scan storage exceeds the tile");
    static_assert(S::kTile <= 65535, "16-bit tile positions");
    typename Scan::TempStorage& scan_tmp = *reinterpret_cast<typename Scan::TempStorage*>(smem);
    __shared__ uint32_t s_spill;  // spilled pairs of the tile, then their start in the spill list

    const BatchTable t = a.tables[0];
    const uint32_t bs = a.bin_slots;
    const uint32_t nbins = (t.cap + bs - 1) / bs;
    const uint32_t ntiles = (uint32_t)((a.single_n + S::kTile - 1) / S::kTile);
    const uint32_t span = (nbins + kBuildThreads - 1) / kBuildThreads;
    Raw* part = (Raw*)a.part;
    Raw* spill = (Raw*)sp.spill;
    Raw p[S::kPer], pn[S::kPer];
    const uint32_t x = blockIdx.x % sp.slices;
    auto load_tile = [&](uint32_t tile, Raw (&q)[S::kPer]) {
        if (tile >= ntiles) return;
        const uint64_t st = (uint64_t)tile * S::kTile;
        const uint32_t len = (uint32_t)min<uint64_t>(S::kTile, a.single_n - st);
        const Raw* src = (const Raw*)a.in + st;
#pragma unroll
        for (uint32_t u = 0; u < S::kPer; ++u) {
            const uint32_t j = threadIdx.x + u * kBuildThreads;
            if (j < len) q[u] = src[j];
        }
    };
    load_tile(blockIdx.x, p);
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint32_t len = (uint32_t)min<uint64_t>(S::kTile, a.single_n - (uint64_t)tile * S::kTile);
        for (uint32_t j = threadIdx.x; j < nbins; j += blockDim.x) tcnt[j] = 0;
        if (threadIdx.x == 0) s_spill = 0;
        __syncthreads();
        uint32_t bin[S::kPer], rank[S::kPer];
#pragma unroll
        for (uint32_t u = 0; u < S::kPer; ++u) {
            const uint32_t j = threadIdx.x + u * kBuildThreads;
            if (j < len) {
                bin[u] = first_slot<KS>(a, t.nbuckets, L::key(p[u])) / bs;
                rank[u] = atomicAdd(&tcnt[bin[u]], 1u);
            }
        }
        load_tile(tile + gridDim.x, pn);  // in flight during the sort and the writes
        __syncthreads();
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
                toff[r] = (uint16_t)excl;
                excl += c;
                uint32_t k = 0;
                if (c) {
                    const uint64_t f = (uint64_t)r * sp.slices + x;
                    const uint32_t at = atomicAdd(&sp.fill[f], c);
                    k = at < sp.cap_slice ? min(c, sp.cap_slice - at) : 0u;
                    base[r] = (uint32_t)(f * sp.cap_slice + at);
                    if (k < c) sbase[r] = (uint16_t)atomicAdd(&s_spill, c - k);
                }
                keep[r] = (uint16_t)k;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0 && s_spill) s_spill = atomicAdd(sp.nspill, s_spill);
#pragma unroll
        for (uint32_t u = 0; u < S::kPer; ++u) {
            const uint32_t j = threadIdx.x + u * kBuildThreads;
            if (j < len) {
                const uint32_t pos = toff[bin[u]] + rank[u];
                staged[pos] = p[u];
                rid[pos] = (uint16_t)bin[u];
            }
        }
        __syncthreads();
        const uint32_t sp0 = s_spill;
        for (uint32_t pos = threadIdx.x; pos < len; pos += blockDim.x) {
            const uint32_t r = rid[pos];
            const uint32_t i = pos - toff[r];
            if (i < keep[r]) part[base[r] + i] = staged[pos];
            else spill[sp0 + sbase[r] + (i - keep[r])] = staged[pos];
        }
        __syncthreads();  // staged, rid and the bin arrays are rewritten by the next tile
#pragma unroll
        for (uint32_t u = 0; u < S::kPer; ++u) p[u] = pn[u];
    }
}

// Dynamic LDS of build_bounded_regroup_kernel for bins of `group` regions.
template <int SS>
constexpr size_t build_bounded_regroup_lds(uint32_t group)
{
    return (size_t)BuildShape<SS>::kTile * SS + (size_t)BuildShape<SS>::kTile * 2 + (size_t)group * 16 +
           (size_t)(kBuildMaxSlices + 1) * 4;
}

// Regroup of bin b = blockIdx.x (the pairs of its G slices, read in slice
// order) by region into part2 (region r at r * capR), reading each pair once.
// Writes the apply's range of every region of the bin: offsets[r] = r * capR,
// apply_end[r] = offsets[r] + the pairs kept.  Dynamic LDS:
// build_bounded_regroup_lds, plus G + 1 words of slice starts.
template <int KS, int VS>
__global__ __launch_bounds__(kBuildThreads) void build_bounded_regroup_kernel(BuildArgs a, SpillArgs sp,
                                                                              const uint8_t* part1, uint32_t group)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    using S = BuildShape<L::SS>;
    extern __shared__ __align__(16) uint8_t smem[];
    Raw* staged = (Raw*)smem;                          // kTile
    uint32_t* rcur = (uint32_t*)(staged + S::kTile);   // group: pairs kept so far in each region
    uint32_t* tcnt = rcur + group;                     // group: pairs of each region in the tile
    uint32_t* toff = tcnt + group;                     // group: first of each region in the sorted tile
    uint32_t* pre = toff + group;                      // G + 1: first pair of each slice in the bin
    uint16_t* keep = (uint16_t*)(pre + kBuildMaxSlices + 1);  // group: pairs of the run that fit the region
    uint16_t* sbase = keep + group;                    // group: start of its spilled pairs in the tile's spill run
    uint16_t* rid = sbase + group;                     // kTile
    __shared__ uint32_t s_spill;
    const BatchTable t = a.tables[0];
    const uint32_t b = blockIdx.x;
    const uint32_t rfirst = b * group;
    const uint32_t nr = min(group, a.nreg - rfirst);
    const uint32_t G = sp.slices;
    const Raw* src = (const Raw*)part1 + (uint64_t)b * G * sp.cap_slice;
    Raw* dst = (Raw*)a.part;
    Raw* spill = (Raw*)sp.spill;
    for (uint32_t j = threadIdx.x; j < nr; j += blockDim.x) rcur[j] = 0;
    if (threadIdx.x < 64) {
        // slice starts: a wavefront's running sum over G counts
        uint32_t run = 0;
        for (uint32_t c0 = 0; c0 < G; c0 += 64) {
            const uint32_t sl = c0 + threadIdx.x;
            const uint32_t v = sl < G ? min(sp.fill[(uint64_t)b * G + sl], sp.cap_slice) : 0u;
            uint32_t incl = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(incl, o);
                if (threadIdx.x >= (uint32_t)o) incl += y;
            }
            if (sl < G) pre[sl] = run + incl - v;
            run += __shfl(incl, 63);
        }
        if (threadIdx.x == 0) pre[G] = run;
    }
    __syncthreads();
    const uint32_t n = pre[G];
    Raw p[S::kPer], pn[S::kPer];
    // pair j of the bin: the slice whose start is the last one <= j
    auto load_tile = [&](uint32_t t0, Raw (&q)[S::kPer]) {
#pragma unroll
        for (uint32_t u = 0; u < S::kPer; ++u) {
            const uint32_t j = t0 + threadIdx.x + u * kBuildThreads;
            if (j < n) {
                uint32_t lo = 0, hi = G;  // pre[lo] <= j < pre[hi]
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) / 2;
                    if (pre[mid] <= j) lo = mid;
                    else hi = mid;
                }
                q[u] = src[(uint64_t)lo * sp.cap_slice + (j - pre[lo])];
            }
        }
    };
    load_tile(0, p);
    for (uint32_t t0 = 0; t0 < n; t0 += S::kTile) {
        const uint32_t tlen = min(S::kTile, n - t0);
        for (uint32_t j = threadIdx.x; j < nr; j += blockDim.x) tcnt[j] = 0;
        if (threadIdx.x == 0) s_spill = 0;
        __syncthreads();
        uint32_t bin[S::kPer], rank[S::kPer];
#pragma unroll
        for (uint32_t u = 0; u < S::kPer; ++u) {
            const uint32_t j = threadIdx.x + u * kBuildThreads;
            if (j < tlen) {
                bin[u] = first_slot<KS>(a, t.nbuckets, L::key(p[u])) / S::kRegion - rfirst;
                rank[u] = atomicAdd(&tcnt[bin[u]], 1u);
            }
        }
        load_tile(t0 + S::kTile, pn);
        __syncthreads();
        if (threadIdx.x < 64) wave0_offsets(tcnt, toff, nr);
        for (uint32_t j = threadIdx.x; j < nr; j += blockDim.x) {
            const uint32_t c = tcnt[j];
            const uint32_t k = min(c, sp.cap_region - rcur[j]);
            keep[j] = (uint16_t)k;
            rcur[j] += k;
            if (k < c) sbase[j] = (uint16_t)atomicAdd(&s_spill, c - k);
        }
        __syncthreads();
        if (threadIdx.x == 0 && s_spill) s_spill = atomicAdd(sp.nspill, s_spill);
#pragma unroll
        for (uint32_t u = 0; u < S::kPer; ++u) {
            const uint32_t j = threadIdx.x + u * kBuildThreads;
            if (j < tlen) {
                const uint32_t pos = toff[bin[u]] + rank[u];
                staged[pos] = p[u];
                rid[pos] = (uint16_t)bin[u];
            }
        }
        __syncthreads();
        const uint32_t sp0 = s_spill;
        for (uint32_t pos = threadIdx.x; pos < tlen; pos += blockDim.x) {
            const uint32_t r = rid[pos];
            const uint32_t i = pos - toff[r];
            if (i < keep[r]) dst[(rfirst + r) * sp.cap_region + (rcur[r] - keep[r]) + i] = staged[pos];
            else spill[sp0 + sbase[r] + (i - keep[r])] = staged[pos];
        }
        __syncthreads();
#pragma unroll
        for (uint32_t u = 0; u < S::kPer; ++u) p[u] = pn[u];
    }
    for (uint32_t j = threadIdx.x; j < nr; j += blockDim.x) {
        const uint32_t r = rfirst + j;
        a.offsets[r] = r * sp.cap_region;
        a.apply_end[r] = r * sp.cap_region + rcur[j];
    }
}

// The spilled pairs: the CAS walk from their first slot, after every region
// is written and the region walks' overflow is inserted.
template <int KS, int VS>
__global__ __launch_bounds__(256) void build_spill_kernel(BuildArgs a, SpillArgs sp)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    const BatchTable t = a.tables[0];
    const uint32_t n = *sp.nspill;
    const Raw* spill = (const Raw*)sp.spill;
    uint32_t fails = 0;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const Raw p = spill[j];
        const uint64_t key = L::key(p);
        if (key == a.empty_key) {
            ++fails;
            continue;
        }
        fails += cas_walk<KS, VS>(a, t, p, key, first_slot<KS>(a, t.nbuckets, key), t.cap);
    }
    for (int o = 32; o > 0; o >>= 1) fails += __shfl_xor(fails, o);
    if ((threadIdx.x & 63) == 0) count_failures(a, 0, fails);
}

}  // namespace mxco
