// BloomFilter (include/bloom_filter.h:37-133) for gfx950: the split-block
// filter of BloomFilterPolicy's defaults (include/bloom_filter_policy.h:54-181):
// 256-bit blocks of eight 32-bit words, one pattern bit per word, xxhash_64.
//
//   h      = XXH64(key bytes, seed)
//   block  = (hi32(h) * numBlocks) >> 32              (BlockIndex, :102-107)
//   word w = 1 << ((salt[w] * lo32(h)) >> 27)        (WordPatternImpl, :157-168)
//
// Add has two paths; OR is idempotent and order-free, so both leave the same
// words whatever the schedule.
//  * direct (small batches): read the 32-byte block once, then 64-bit atomic
//    ORs only for the word pairs that still miss bits;
//  * partitioned (at least half a key per block): the filter is cut into
//    ranges of 2048 blocks (64 KiB).  A count pass builds per-segment range
//    histograms, a scan turns them into offsets, a scatter pass writes each
//    key as (block within range, low hash word) into its range's bin, and an
//    apply pass loads each touched range into LDS, ORs its bin there and
//    writes the range back once.  Per key this streams 32 bytes instead of
//    four random global atomics.
// Contains reads the block with two 16-byte loads.  Merge / Intersect are
// word-wise OR / AND streams.
//
// Other policies (BloomFilterPolicy<Key, xxhash_64, uint32_t, W, P>, W = 1..32
// words per block, P pattern bits spread P / W per word over the 64-salt
// table, include/bloom_filter_policy.h:54-168) run a generic kernel pair:
// one key per lane, the block read once, atomic ORs only for words that
// miss bits.
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <new>
#include <string>

#include <hipcub/hipcub.hpp>

#include "mxco_common.h"
#include "mxco/detail/bloom_salts.h"
#include "mxco/detail/hash.h"

using namespace mxco;

struct mxco_bloom {
    uint32_t* words;
    uint64_t num_blocks;
    uint32_t key_size;
    uint64_t seed;
    uint32_t wpb;           // words per block
    uint32_t pattern_bits;  // bits set per key
    bool owns_words;        // false: the caller's storage
    uint8_t* ws;  // partitioned-Add workspace (grown on demand)
    size_t ws_bytes;
};

namespace {

// Pattern salts of BloomFilterPolicy (include/mxco/detail/bloom_salts.h).
using mxco::detail::kBloomSalts;

template <int KS>
__device__ __forceinline__ uint64_t key_at(const uint8_t* keys, uint64_t i)
{
    if constexpr (KS == 4) return ((const uint32_t*)keys)[i];
    else return ((const uint64_t*)keys)[i];
}

__device__ __forceinline__ void locate(uint64_t h, uint32_t nb, uint32_t& blk, uint64_t m[4])
{
    blk = (uint32_t)(((uint64_t)(uint32_t)(h >> 32) * nb) >> 32);
    const uint32_t lo = (uint32_t)h;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const uint32_t a = 1u << ((kBloomSalts[2 * p] * lo) >> 27);
        const uint32_t b = 1u << ((kBloomSalts[2 * p + 1] * lo) >> 27);
        m[p] = (uint64_t)a | ((uint64_t)b << 32);
    }
}

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

template <int KS>
__global__ __launch_bounds__(256) void bloom_add_kernel(uint32_t* words, uint32_t nb, const uint8_t* keys,
                                                        const uint8_t* stencil, uint64_t n, uint64_t seed)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (stencil && stencil[i] == 0) continue;
        const uint64_t h = xxhash64<KS>(key_at<KS>(keys, i), seed);
        uint32_t blk;
        uint64_t m[4];
        locate(h, nb, blk, m);
        u64x2* b = (u64x2*)(words + (uint64_t)blk * 8);
        const u64x2 c0 = b[0], c1 = b[1];
        unsigned long long* w = (unsigned long long*)b;
        if ((c0.x & m[0]) != m[0]) atomicOr(w + 0, (unsigned long long)m[0]);
        if ((c0.y & m[1]) != m[1]) atomicOr(w + 1, (unsigned long long)m[1]);
        if ((c1.x & m[2]) != m[2]) atomicOr(w + 2, (unsigned long long)m[2]);
        if ((c1.y & m[3]) != m[3]) atomicOr(w + 3, (unsigned long long)m[3]);
    }
}

template <int KS>
__global__ __launch_bounds__(256) void bloom_contains_kernel(const uint32_t* words, uint32_t nb, const uint8_t* keys,
                                                             const uint8_t* stencil, uint8_t* out, uint64_t n,
                                                             uint64_t seed)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (stencil && stencil[i] == 0) {
            out[i] = 0;
            continue;
        }
        const uint64_t h = xxhash64<KS>(key_at<KS>(keys, i), seed);
        uint32_t blk;
        uint64_t m[4];
        locate(h, nb, blk, m);
        const u64x2* b = (const u64x2*)(words + (uint64_t)blk * 8);
        const u64x2 c0 = b[0], c1 = b[1];
        const bool hit = ((c0.x & m[0]) == m[0]) & ((c0.y & m[1]) == m[1]) & ((c1.x & m[2]) == m[2]) &
                         ((c1.y & m[3]) == m[3]);
        out[i] = hit ? 1 : 0;
    }
}

// Word w's pattern under a policy of `bpw` bits per word: bit
// (salt[w * bpw + b] * lo) >> 27 for b < bpw (WordPatternImpl,
// bloom_filter_policy.h:157-168).
__device__ __forceinline__ uint32_t word_pattern(uint32_t w, uint32_t bpw, uint32_t lo)
{
    uint32_t m = 0;
    for (uint32_t b = 0; b < bpw; ++b) m |= 1u << ((kBloomSalts[w * bpw + b] * lo) >> 27);
    return m;
}

// One key per lane: the W words of its block are read once; a word gets an
// atomic OR only when it misses bits.
template <int KS, int W>
__global__ __launch_bounds__(256) void bloom_generic_add_kernel(uint32_t* words, uint32_t nb, uint32_t bpw,
                                                                const uint8_t* keys, const uint8_t* stencil,
                                                                uint64_t n, uint64_t seed)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (stencil && stencil[i] == 0) continue;
        const uint64_t h = xxhash64<KS>(key_at<KS>(keys, i), seed);
        const uint32_t blk = (uint32_t)(((uint64_t)(uint32_t)(h >> 32) * nb) >> 32);
        uint32_t* b = words + (uint64_t)blk * W;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint32_t m = word_pattern((uint32_t)w, bpw, (uint32_t)h);
            if ((b[w] & m) != m) atomicOr(b + w, m);
        }
    }
}

template <int KS, int W>
__global__ __launch_bounds__(256) void bloom_generic_contains_kernel(const uint32_t* words, uint32_t nb, uint32_t bpw,
                                                                     const uint8_t* keys, const uint8_t* stencil,
                                                                     uint8_t* out, uint64_t n, uint64_t seed)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (stencil && stencil[i] == 0) {
            out[i] = 0;
            continue;
        }
        const uint64_t h = xxhash64<KS>(key_at<KS>(keys, i), seed);
        const uint32_t blk = (uint32_t)(((uint64_t)(uint32_t)(h >> 32) * nb) >> 32);
        const uint32_t* b = words + (uint64_t)blk * W;
        bool hit = true;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint32_t m = word_pattern((uint32_t)w, bpw, (uint32_t)h);
            hit &= (b[w] & m) == m;
        }
        out[i] = hit ? 1 : 0;
    }
}

template <int KS>
hipError_t launch_generic(bool add, const mxco_bloom* f, const uint8_t* keys, const uint8_t* stencil, uint8_t* out,
                          uint64_t n, hipStream_t s)
{
    const unsigned grid = grid_for(n, 8);
    const uint32_t nb = (uint32_t)f->num_blocks, bpw = f->pattern_bits / f->wpb;
#define MXCO_BLOOM_GENERIC(W)                                                                                  \
    case W:                                                                                                    \
        if (add) bloom_generic_add_kernel<KS, W><<<grid, 256, 0, s>>>(f->words, nb, bpw, keys, stencil, n, f->seed); \
        else bloom_generic_contains_kernel<KS, W><<<grid, 256, 0, s>>>(f->words, nb, bpw, keys, stencil, out, n,     \
                                                                        f->seed);                                   \
        break;
    switch (f->wpb) {
        MXCO_BLOOM_GENERIC(1)
        MXCO_BLOOM_GENERIC(2)
        MXCO_BLOOM_GENERIC(4)
        MXCO_BLOOM_GENERIC(8)
        MXCO_BLOOM_GENERIC(16)
        MXCO_BLOOM_GENERIC(32)
    default: return hipErrorInvalidValue;
    }
#undef MXCO_BLOOM_GENERIC
    return hipGetLastError();
}

// The default policy (8 words, 8 pattern bits) has the tuned kernels.
bool default_policy(const mxco_bloom* f) { return f->wpb == 8 && f->pattern_bits == 8; }

constexpr uint32_t kRangeShift = 11;  // 2048 blocks x 32 B = 64 KiB of LDS per range
constexpr uint32_t kRangeBlocks = 1u << kRangeShift;
constexpr uint32_t kMaxRanges = 8192;             // range histogram in LDS: 32 KiB
constexpr uint64_t kPassKeys = 1ull << 27;        // keys per partition pass (1 GiB of bin entries)
constexpr int kUnroll = 4;                        // key loads in flight per lane

// Visits the keys of this workgroup's contiguous segment that pass the
// stencil, kUnroll loads in flight per lane: f(block, low hash word).
template <int KS, typename F>
__device__ __forceinline__ void for_segment(const uint8_t* keys, const uint8_t* stencil, uint64_t n, uint32_t nb,
                                            uint64_t seed, F&& f)
{
    const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t begin = (uint64_t)blockIdx.x * per;
    const uint64_t end = min(n, begin + per);
    for (uint64_t i0 = begin + threadIdx.x; i0 < end; i0 += (uint64_t)kUnroll * blockDim.x) {
        uint64_t k[kUnroll];
        bool live[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const uint64_t i = i0 + (uint64_t)u * blockDim.x;
            live[u] = i < end;
            k[u] = live[u] ? key_at<KS>(keys, i) : 0;
            if (live[u] && stencil) live[u] = stencil[i] != 0;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            if (!live[u]) continue;
            const uint64_t h = xxhash64<KS>(k[u], seed);
            f((uint32_t)(((uint64_t)(uint32_t)(h >> 32) * nb) >> 32), (uint32_t)h);
        }
    }
}

// counts[r * segments + segment] = keys of the segment that fall in range r.
template <int KS>
__global__ __launch_bounds__(256) void bloom_count_kernel(uint32_t nb, const uint8_t* keys, const uint8_t* stencil,
                                                          uint64_t n, uint64_t seed, uint32_t nranges,
                                                          uint32_t* counts)
{
    extern __shared__ uint32_t hist[];
    for (uint32_t j = threadIdx.x; j < nranges; j += blockDim.x) hist[j] = 0;
    __syncthreads();
    for_segment<KS>(keys, stencil, n, nb, seed,
                    [&](uint32_t blk, uint32_t) { atomicAdd(&hist[blk >> kRangeShift], 1u); });
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < nranges; j += blockDim.x) counts[(uint64_t)j * gridDim.x + blockIdx.x] = hist[j];
}

// Writes each key's (block within range, low hash word) at its range's cursor.
template <int KS>
__global__ __launch_bounds__(256) void bloom_scatter_kernel(uint32_t nb, const uint8_t* keys, const uint8_t* stencil,
                                                            uint64_t n, uint64_t seed, uint32_t nranges,
                                                            const uint32_t* offsets, uint64_t* entries)
{
    extern __shared__ uint32_t cur[];
    for (uint32_t j = threadIdx.x; j < nranges; j += blockDim.x) cur[j] = offsets[(uint64_t)j * gridDim.x + blockIdx.x];
    __syncthreads();
    for_segment<KS>(keys, stencil, n, nb, seed, [&](uint32_t blk, uint32_t lo) {
        const uint32_t pos = atomicAdd(&cur[blk >> kRangeShift], 1u);
        entries[pos] = ((uint64_t)(blk & (kRangeBlocks - 1)) << 32) | lo;
    });
}

// The same scatter with the segment's keys staged in LDS tiles of kTileKeys:
// a tile is sorted by range in LDS (counting sort), then each range's run of
// the tile is written to its bin as one contiguous piece, so the bins fill
// in multi-entry runs instead of one 8-byte write per key.  Dynamic LDS:
// cursors (4 B), tile counts (2 B, two per 32-bit word) and tile offsets
// (2 B) per range, and 10 B per tile key.
constexpr uint32_t kTileKeys = 8192;
constexpr uint32_t kScatterThreads = 1024;
constexpr uint32_t kTilePerThread = kTileKeys / kScatterThreads;

size_t staged_scatter_lds(uint32_t nranges)
{
    return (size_t)nranges * 4 + ((size_t)nranges + 1) / 2 * 4 + (size_t)nranges * 2 + (size_t)kTileKeys * 10;
}

template <int KS>
__global__ __launch_bounds__(kScatterThreads) void bloom_staged_scatter_kernel(
    uint32_t nb, const uint8_t* keys, const uint8_t* stencil, uint64_t n, uint64_t seed, uint32_t nranges,
    const uint32_t* offsets, uint64_t* entries)
{
    extern __shared__ __align__(16) uint8_t smem[];
    uint64_t* ent = (uint64_t*)smem;                             // kTileKeys
    uint32_t* cur = (uint32_t*)(ent + kTileKeys);                // nranges
    uint32_t* tcnt = cur + nranges;                              // (nranges + 1) / 2 words of two u16
    uint16_t* toff = (uint16_t*)(tcnt + (nranges + 1) / 2);      // nranges
    uint16_t* rid = toff + nranges;                              // kTileKeys
    typedef hipcub::BlockScan<uint32_t, kScatterThreads> Scan;
    __shared__ typename Scan::TempStorage scan_tmp;

    for (uint32_t j = threadIdx.x; j < nranges; j += blockDim.x) cur[j] = offsets[(uint64_t)j * gridDim.x + blockIdx.x];
    // the count kernel's segment of this workgroup (same grid, same split)
    const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t begin = (uint64_t)blockIdx.x * per;
    const uint64_t end = min(n, begin + per);
    const uint32_t span = (nranges + kScatterThreads - 1) / kScatterThreads;  // ranges per thread in the scan
    // the next tile's keys are loaded while the current tile is sorted
    uint64_t kc[kTilePerThread], kn[kTilePerThread];
    auto load_tile = [&](uint64_t t, uint64_t (&k)[kTilePerThread]) {
#pragma unroll
        for (uint32_t u = 0; u < kTilePerThread; ++u) {
            const uint64_t i = t + threadIdx.x + u * kScatterThreads;
            k[u] = i < end ? key_at<KS>(keys, i) : 0;
        }
    };
    load_tile(begin, kc);
    for (uint64_t t0 = begin; t0 < end; t0 += kTileKeys) {
        const uint32_t tlen = (uint32_t)min<uint64_t>(kTileKeys, end - t0);
        if (t0 + kTileKeys < end) load_tile(t0 + kTileKeys, kn);
        for (uint32_t j = threadIdx.x; j < (nranges + 1) / 2; j += blockDim.x) tcnt[j] = 0;
        __syncthreads();
        uint32_t rr[kTilePerThread], rank[kTilePerThread];
        uint64_t ev[kTilePerThread];
#pragma unroll
        for (uint32_t u = 0; u < kTilePerThread; ++u) {
            const uint32_t j = threadIdx.x + u * kScatterThreads;
            rr[u] = 0xffffffffu;
            if (j < tlen && (!stencil || stencil[t0 + j] != 0)) {
                const uint64_t h = xxhash64<KS>(kc[u], seed);
                const uint32_t blk = (uint32_t)(((uint64_t)(uint32_t)(h >> 32) * nb) >> 32);
                rr[u] = blk >> kRangeShift;
                ev[u] = ((uint64_t)(blk & (kRangeBlocks - 1)) << 32) | (uint32_t)h;
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kTilePerThread; ++u) {
            if (rr[u] == 0xffffffffu) continue;
            const uint32_t sh = (rr[u] & 1u) * 16u;
            rank[u] = (atomicAdd(&tcnt[rr[u] >> 1], 1u << sh) >> sh) & 0xffffu;
        }
        __syncthreads();
        // exclusive scan of the tile counts: `span` consecutive ranges per thread
        uint32_t local = 0;
        for (uint32_t q = 0; q < span; ++q) {
            const uint32_t r = threadIdx.x * span + q;
            if (r < nranges) local += (tcnt[r >> 1] >> ((r & 1u) * 16u)) & 0xffffu;
        }
        uint32_t excl;
        Scan(scan_tmp).ExclusiveSum(local, excl);
        for (uint32_t q = 0; q < span; ++q) {
            const uint32_t r = threadIdx.x * span + q;
            if (r < nranges) {
                toff[r] = (uint16_t)excl;
                excl += (tcnt[r >> 1] >> ((r & 1u) * 16u)) & 0xffffu;
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t u = 0; u < kTilePerThread; ++u) {
            if (rr[u] == 0xffffffffu) continue;
            const uint32_t pos = toff[rr[u]] + rank[u];
            ent[pos] = ev[u];
            rid[pos] = (uint16_t)rr[u];
        }
        __syncthreads();
        // the tile's selected keys occupy positions [0, selected)
        const uint32_t selected = (uint32_t)toff[nranges - 1] +
                                  ((tcnt[(nranges - 1) >> 1] >> (((nranges - 1) & 1u) * 16u)) & 0xffffu);
        for (uint32_t pos = threadIdx.x; pos < selected; pos += blockDim.x) {
            const uint32_t r = rid[pos];
            entries[cur[r] + (pos - toff[r])] = ent[pos];
        }
        __syncthreads();
        for (uint32_t r = threadIdx.x; r < nranges; r += blockDim.x)
            cur[r] += (tcnt[r >> 1] >> ((r & 1u) * 16u)) & 0xffffu;
#pragma unroll
        for (uint32_t u = 0; u < kTilePerThread; ++u) kc[u] = kn[u];
        __syncthreads();
    }
}

// One range at a time: load its words into LDS, OR its bin, write it back.
// offsets[r * segments] is the first entry of range r; offsets[nranges *
// segments] is the total.
__global__ __launch_bounds__(1024) void bloom_apply_kernel(uint32_t* words, uint32_t nb, uint32_t nranges,
                                                           uint32_t segments, const uint32_t* offsets,
                                                           const uint64_t* entries)
{
    __shared__ __align__(16) uint32_t lw[kRangeBlocks * 8];
    for (uint32_t r = blockIdx.x; r < nranges; r += gridDim.x) {
        const uint32_t s = offsets[(uint64_t)r * segments];
        const uint32_t e = offsets[(uint64_t)(r + 1) * segments];
        if (s == e) continue;
        const uint64_t b0 = (uint64_t)r << kRangeShift;
        const uint32_t n16 = (uint32_t)min<uint64_t>(kRangeBlocks, nb - b0) * 2;  // 16-byte vectors
        u64x2* g = (u64x2*)(words + b0 * 8);
        u64x2* l = (u64x2*)lw;
        for (uint32_t j = threadIdx.x; j < n16; j += blockDim.x) l[j] = g[j];
        __syncthreads();
        // 8 entry loads in flight per lane
        for (uint32_t k0 = s + threadIdx.x; k0 < e; k0 += 8 * blockDim.x) {
            uint64_t v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t k = k0 + u * blockDim.x;
                v[u] = k < e ? entries[k] : 0;
            }
            // Lane l visits the block's 64-bit words in the order l, l+1, ...
            // (mod 4): blocks start at multiples of 8 words, so in a fixed
            // order every lane of a wavefront would hit the same few banks.
            const uint32_t rot = threadIdx.x & 3u;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (k0 + u * blockDim.x >= e) break;
                const uint32_t lo = (uint32_t)v[u];
                // the block's 8 words as 4 64-bit words: one LDS atomic per
                // pair of words, none when the bits are already set
                unsigned long long* w = (unsigned long long*)(lw + (uint32_t)(v[u] >> 32) * 8);
                unsigned long long m[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    m[q] = (unsigned long long)(1u << ((kBloomSalts[2 * q] * lo) >> 27)) |
                           (unsigned long long)(1u << ((kBloomSalts[2 * q + 1] * lo) >> 27)) << 32;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t qq = (q + rot) & 3u;
                    const unsigned long long mq = qq == 0 ? m[0] : (qq == 1 ? m[1] : (qq == 2 ? m[2] : m[3]));
                    if ((w[qq] & mq) != mq) atomicOr(w + qq, mq);
                }
            }
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < n16; j += blockDim.x) g[j] = l[j];
        __syncthreads();
    }
}

__global__ void words_zero_tail(uint32_t* p)
{
    *p = 0;
}

// op 0: dst |= src, op 1: dst &= src, op 2: dst = 0, over nwords 32-bit
// words: 16-byte vectors, then the last nwords % 4 words (blocks of one or
// two words leave them).
__global__ __launch_bounds__(256) void bloom_words_kernel(uint32_t* dst, const uint32_t* src, uint64_t nwords, int op)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n16 = nwords / 4;
    u64x2* d = (u64x2*)dst;
    const u64x2* s = (const u64x2*)src;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) {
        if (op == 2) {
            u64x2 z = {0, 0};
            d[i] = z;
        } else {
            const u64x2 a = d[i], b = s[i];
            d[i] = op == 0 ? (a | b) : (a & b);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < nwords % 4) {
        const uint64_t j = n16 * 4 + threadIdx.x;
        dst[j] = op == 2 ? 0u : (op == 0 ? (dst[j] | src[j]) : (dst[j] & src[j]));
    }
}

// 1: partitioned Add (at least half a key per block and 2^16 keys), 0: direct.
// MXCO_BLOOM_ADD_PATH=0/1 forces a path.
int add_path(uint64_t n, uint64_t num_blocks)
{
    static const int forced = [] {
        const char* e = getenv("MXCO_BLOOM_ADD_PATH");
        return e ? atoi(e) : -1;
    }();
    if (forced == 0 || forced == 1) return forced;
    return (n >= (1u << 16) && 2 * n >= num_blocks) ? 1 : 0;
}

uint32_t segments_for(uint64_t n)
{
    return (uint32_t)std::min<uint64_t>(1024, std::max<uint64_t>(64, n / 16384));
}

mxco_status ensure_ws(mxco_bloom* f, size_t bytes)
{
    if (bytes <= f->ws_bytes) return MXCO_OK;
    if (f->ws) MXCO_HIP(hipFree(f->ws), "BloomFilter::Add: workspace");
    f->ws = nullptr;
    f->ws_bytes = 0;
    MXCO_HIP(hipMalloc((void**)&f->ws, bytes), "BloomFilter::Add: workspace");
    f->ws_bytes = bytes;
    return MXCO_OK;
}

// Workspace of one partitioned pass of n keys: bin entries (8 B per key),
// the range-by-segment counts (+ the total), their offsets and the scan's
// temporary storage, each 256-byte aligned.
struct PassLayout {
    uint32_t nranges, segs;
    uint64_t ncnt;
    size_t off_counts, off_offsets, off_temp, temp, total;
};

mxco_status pass_layout(const mxco_bloom* f, uint64_t n, hipStream_t stream, PassLayout& L)
{
    L.nranges = (uint32_t)((f->num_blocks + kRangeBlocks - 1) >> kRangeShift);
    L.segs = segments_for(n);
    L.ncnt = (uint64_t)L.nranges * L.segs + 1;  // + the total
    L.temp = 0;
    MXCO_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, L.temp, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                              (int64_t)L.ncnt, stream),
             "BloomFilter::Add: scan size");
    L.off_counts = ((n * 8) + 255) & ~size_t(255);
    L.off_offsets = L.off_counts + ((L.ncnt * 4 + 255) & ~size_t(255));
    L.off_temp = L.off_offsets + ((L.ncnt * 4 + 255) & ~size_t(255));
    L.total = L.off_temp + L.temp;
    return MXCO_OK;
}

// Partitioned Add of n <= kPassKeys keys (see the file comment) in the
// workspace `ws` laid out by L.
mxco_status add_partitioned(mxco_bloom* f, const uint8_t* keys, const uint8_t* stencil, uint64_t n,
                            hipStream_t stream, uint8_t* ws, const PassLayout& L)
{
    const uint32_t nb = (uint32_t)f->num_blocks;
    const uint32_t nranges = L.nranges, segs = L.segs;
    const uint64_t ncnt = L.ncnt;
    size_t temp = L.temp;
    uint64_t* entries = (uint64_t*)ws;
    uint32_t* counts = (uint32_t*)(ws + L.off_counts);
    uint32_t* offsets = (uint32_t*)(ws + L.off_offsets);
    const size_t lds = (size_t)nranges * 4;
    // the slot after the last count must read 0 so the scan's last output is the total
    words_zero_tail<<<1, 1, 0, stream>>>(counts + ncnt - 1);
    if (f->key_size == 4)
        bloom_count_kernel<4><<<segs, 256, lds, stream>>>(nb, keys, stencil, n, f->seed, nranges, counts);
    else
        bloom_count_kernel<8><<<segs, 256, lds, stream>>>(nb, keys, stencil, n, f->seed, nranges, counts);
    MXCO_HIP(hipGetLastError(), "BloomFilter::Add: count");
    MXCO_HIP(hipcub::DeviceScan::ExclusiveSum(ws + L.off_temp, temp, counts, offsets, (int64_t)ncnt, stream),
             "BloomFilter::Add: scan");
    // LDS-staged scatter; the per-key scatter remains for MXCO_BLOOM_STAGED=0
    static const int staged = env_int("MXCO_BLOOM_STAGED", 1);
    static const hipError_t attr = [] {
        const int bytes = (int)staged_scatter_lds(kMaxRanges);
        hipError_t e = hipFuncSetAttribute((const void*)bloom_staged_scatter_kernel<4>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)bloom_staged_scatter_kernel<8>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        return e;
    }();
    if (staged && attr == hipSuccess) {
        const size_t slds = staged_scatter_lds(nranges);
        if (f->key_size == 4)
            bloom_staged_scatter_kernel<4><<<segs, kScatterThreads, slds, stream>>>(nb, keys, stencil, n, f->seed,
                                                                                    nranges, offsets, entries);
        else
            bloom_staged_scatter_kernel<8><<<segs, kScatterThreads, slds, stream>>>(nb, keys, stencil, n, f->seed,
                                                                                    nranges, offsets, entries);
    } else {
        if (f->key_size == 4)
            bloom_scatter_kernel<4><<<segs, 256, lds, stream>>>(nb, keys, stencil, n, f->seed, nranges, offsets,
                                                                entries);
        else
            bloom_scatter_kernel<8><<<segs, 256, lds, stream>>>(nb, keys, stencil, n, f->seed, nranges, offsets,
                                                                entries);
    }
    MXCO_HIP(hipGetLastError(), "BloomFilter::Add: scatter");
    const unsigned grid = std::min<unsigned>(nranges, (unsigned)cu_count() * 2);
    bloom_apply_kernel<<<grid, 1024, 0, stream>>>(f->words, nb, nranges, segs, offsets, entries);
    MXCO_HIP(hipGetLastError(), "BloomFilter::Add: apply");
    return MXCO_OK;
}

mxco_status add_direct(mxco_bloom* f, const void* keys, const uint8_t* stencil, uint64_t n, hipStream_t stream)
{
    if (!default_policy(f)) {
        hipError_t e = f->key_size == 4
                           ? launch_generic<4>(true, f, (const uint8_t*)keys, stencil, nullptr, n, stream)
                           : launch_generic<8>(true, f, (const uint8_t*)keys, stencil, nullptr, n, stream);
        MXCO_HIP(e, "BloomFilter::Add");
        return MXCO_OK;
    }
    const unsigned grid = grid_for(n, 8);
    if (f->key_size == 4)
        bloom_add_kernel<4><<<grid, 256, 0, stream>>>(f->words, (uint32_t)f->num_blocks, (const uint8_t*)keys,
                                                     stencil, n, f->seed);
    else
        bloom_add_kernel<8><<<grid, 256, 0, stream>>>(f->words, (uint32_t)f->num_blocks, (const uint8_t*)keys,
                                                     stencil, n, f->seed);
    MXCO_HIP(hipGetLastError(), "BloomFilter::Add");
    return MXCO_OK;
}

// Fixed scratch of mxco_bloom_add_noalloc: one partitioned pass of up to
// kScratchKeys keys on a filter of up to kMaxRanges ranges (512 MiB): 32 MiB
// of bin entries, two count matrices of 8 MiB (kMaxRanges x 256 segments)
// and 4 MiB for the scan.  Every GPU the library runs on holds its own copy
// until the process ends; calls on one GPU share it.
constexpr uint64_t kScratchKeys = 1ull << 22;
constexpr size_t kScratchBytes = (kScratchKeys * 8) + 2 * (((size_t)kMaxRanges * 256 * 4 + 4 + 255) & ~size_t(255)) +
                                 ((size_t)4 << 20);
__device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices];

// Device address of the current device's scratch, looked up once per device
// with that device current (the host-side name of a __device__ array is no
// device address).
mxco_status scratch_on_device(uint8_t** out)
{
    int dev = 0;
    MXCO_HIP(hipGetDevice(&dev), "BloomFilter::Add (no-alloc): get device");
    if (dev < 0 || dev >= kMaxDevices)
        return fail(MXCO_UNSUPPORTED, "BloomFilter::Add (no-alloc): device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        MXCO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "BloomFilter::Add (no-alloc): scratch address");
        g_scratch_ptr[dev] = (uint8_t*)p;
    }
    *out = g_scratch_ptr[dev];
    return MXCO_OK;
}

mxco_status check_filter(const mxco_bloom* f, const char* fn)
{
    if (!f || !f->words) return fail(MXCO_LOGIC_ERROR, std::string("BloomFilter::") + fn + ": filter has no storage");
    return MXCO_OK;
}

mxco_status words_op(mxco_bloom* dst, const mxco_bloom* src, int op, hipStream_t stream, int sync, const char* fn)
{
    const uint64_t nwords = dst->num_blocks * dst->wpb;
    bloom_words_kernel<<<std::max(1u, grid_for(nwords / 4, 8)), 256, 0, stream>>>(dst->words, src ? src->words : nullptr,
                                                                                 nwords, op);
    MXCO_HIP(hipGetLastError(), fn);
    if (sync) MXCO_HIP(hipStreamSynchronize(stream), fn);
    return MXCO_OK;
}

mxco_status compatible(const mxco_bloom* a, const mxco_bloom* b, const char* fn)
{
    if (!a || !b || !a->words || !b->words)
        return fail(MXCO_LOGIC_ERROR, std::string("BloomFilter::") + fn + ": a moved-from filter cannot participate");
    if (a->num_blocks != b->num_blocks)
        return fail(MXCO_INVALID_ARGUMENT, std::string("BloomFilter::") + fn + ": filters have different block extents");
    if (a->seed != b->seed || a->key_size != b->key_size || a->wpb != b->wpb || a->pattern_bits != b->pattern_bits)
        return fail(MXCO_INVALID_ARGUMENT, std::string("BloomFilter::") + fn + ": filters have different policies");
    return MXCO_OK;
}

}  // namespace

extern "C" {

MXCO_API mxco_status mxco_bloom_create_with_policy(uint64_t num_blocks, uint32_t key_size,
                                                   const mxco_bloom_policy* policy, uint32_t* storage,
                                                   hipStream_t stream, mxco_bloom** out)
{
    if (!out || !policy) return fail(MXCO_INVALID_ARGUMENT, "BloomFilter: null argument");
    *out = nullptr;
    if (key_size != 4 && key_size != 8) return fail(MXCO_UNSUPPORTED, "BloomFilter: keys must be 4 or 8 bytes");
    const uint32_t w = policy->words_per_block, p = policy->pattern_bits;
    // BloomFilterPolicy's static checks (bloom_filter_policy.h:80-98)
    if (w == 0 || w > 32 || (w & (w - 1)) != 0)
        return fail(MXCO_INVALID_ARGUMENT, "BloomFilterPolicy: WordsPerBlock must be a power of two of at most 32");
    if (p < w || p > 64 || p > 32 * w || p % w != 0)
        return fail(MXCO_INVALID_ARGUMENT,
                    "BloomFilterPolicy: PatternBits must be a multiple of WordsPerBlock of at most 64");
    if (num_blocks == 0) return fail(MXCO_INVALID_ARGUMENT, "BloomFilter: numBlocks must be greater than zero");
    if (num_blocks > 0xffffffffull)
        return fail(MXCO_LENGTH_ERROR, "BloomFilter: numBlocks exceeds the supported block-index range");
    mxco_bloom* f = new (std::nothrow) mxco_bloom();
    if (!f) return fail(MXCO_OUT_OF_MEMORY, "BloomFilter: host allocation");
    f->num_blocks = num_blocks;
    f->key_size = key_size;
    f->seed = policy->seed;
    f->wpb = w;
    f->pattern_bits = p;
    f->owns_words = storage == nullptr;
    if (storage) {
        f->words = storage;
    } else {
        hipError_t e = hipMalloc((void**)&f->words, num_blocks * w * 4);
        if (e != hipSuccess) {
            delete f;
            return hip_fail(e, "BloomFilter: allocation");
        }
    }
    mxco_status st = words_op(f, nullptr, 2, stream, 1, "BloomFilter::Clear");
    if (st != MXCO_OK) {
        if (f->owns_words) (void)hipFree(f->words);
        delete f;
        return st;
    }
    *out = f;
    return ok();
}

MXCO_API mxco_status mxco_bloom_create(uint64_t num_blocks, uint32_t key_size, uint64_t seed, hipStream_t stream,
                                       mxco_bloom** out)
{
    const mxco_bloom_policy policy = {8, 8, seed};
    return mxco_bloom_create_with_policy(num_blocks, key_size, &policy, nullptr, stream, out);
}

MXCO_API mxco_status mxco_bloom_destroy(mxco_bloom* f)
{
    if (!f) return ok();
    if (f->words && f->owns_words) (void)hipFree(f->words);
    if (f->ws) (void)hipFree(f->ws);
    delete f;
    return ok();
}

MXCO_API mxco_status mxco_bloom_data(const mxco_bloom* f, uint32_t** words, uint64_t* num_blocks)
{
    if (mxco_status s = check_filter(f, "Data")) return s;
    if (words) *words = f->words;
    if (num_blocks) *num_blocks = f->num_blocks;
    return ok();
}

MXCO_API mxco_status mxco_bloom_clear(mxco_bloom* f, hipStream_t stream, int sync)
{
    if (mxco_status s = check_filter(f, "Clear")) return s;
    if (mxco_status s = words_op(f, nullptr, 2, stream, sync, "BloomFilter::Clear")) return s;
    return ok();
}

MXCO_API mxco_status mxco_bloom_add(mxco_bloom* f, const void* keys, const uint8_t* stencil, uint64_t n,
                                    hipStream_t stream, int sync)
{
    if (n == 0) return ok();
    if (mxco_status s = check_filter(f, "Add")) return s;
    if (!keys) return fail(MXCO_INVALID_ARGUMENT, "BloomFilter::Add: keys must not be null when keyNum is non-zero");
    const uint64_t nranges = (f->num_blocks + kRangeBlocks - 1) >> kRangeShift;
    if (default_policy(f) && add_path(n, f->num_blocks) == 1 && nranges <= kMaxRanges) {
        for (uint64_t done = 0; done < n; done += kPassKeys) {
            const uint64_t m = std::min(kPassKeys, n - done);
            PassLayout L;
            if (mxco_status e = pass_layout(f, m, stream, L)) return e;
            if (mxco_status e = ensure_ws(f, L.total)) return e;
            if (mxco_status e = add_partitioned(f, (const uint8_t*)keys + done * f->key_size,
                                                stencil ? stencil + done : nullptr, m, stream, f->ws, L))
                return e;
        }
    } else if (mxco_status e = add_direct(f, keys, stencil, n, stream)) {
        return e;
    }
    if (sync) MXCO_HIP(hipStreamSynchronize(stream), "BloomFilter::Add");
    return ok();
}

MXCO_API mxco_status mxco_bloom_add_noalloc(mxco_bloom* f, const void* keys, const uint8_t* stencil, uint64_t n,
                                            hipStream_t stream, int sync)
{
    if (n == 0) return ok();
    if (mxco_status s = check_filter(f, "Add")) return s;
    if (!keys) return fail(MXCO_INVALID_ARGUMENT, "BloomFilter::Add: keys must not be null when keyNum is non-zero");
    if (n > kScratchKeys)
        return fail(MXCO_LENGTH_ERROR,
                    "BloomFilter::Add (no-alloc): at most 4,194,304 keys per call (the library's fixed scratch)");
    const uint64_t nranges = (f->num_blocks + kRangeBlocks - 1) >> kRangeShift;
    if (default_policy(f) && add_path(n, f->num_blocks) == 1 && nranges <= kMaxRanges) {
        PassLayout L;
        if (mxco_status e = pass_layout(f, n, stream, L)) return e;
        if (L.total > kScratchBytes)
            return fail(MXCO_LENGTH_ERROR, "BloomFilter::Add (no-alloc): the pass needs more than the fixed scratch");
        uint8_t* ws = nullptr;
        if (mxco_status e = scratch_on_device(&ws)) return e;
        if (mxco_status e = add_partitioned(f, (const uint8_t*)keys, stencil, n, stream, ws, L)) return e;
    } else if (mxco_status e = add_direct(f, keys, stencil, n, stream)) {
        return e;
    }
    if (sync) MXCO_HIP(hipStreamSynchronize(stream), "BloomFilter::Add");
    return ok();
}

MXCO_API mxco_status mxco_bloom_contains(const mxco_bloom* f, const void* keys, const uint8_t* stencil, uint8_t* out,
                                         uint64_t n, hipStream_t stream, int sync)
{
    if (n == 0) return ok();
    if (mxco_status s = check_filter(f, "Contains")) return s;
    if (!keys)
        return fail(MXCO_INVALID_ARGUMENT, "BloomFilter::Contains: keys must not be null when keyNum is non-zero");
    if (!out)
        return fail(MXCO_INVALID_ARGUMENT,
                    "BloomFilter::Contains: outputValues must not be null when keyNum is non-zero");
    if (!default_policy(f)) {
        hipError_t e = f->key_size == 4 ? launch_generic<4>(false, f, (const uint8_t*)keys, stencil, out, n, stream)
                                        : launch_generic<8>(false, f, (const uint8_t*)keys, stencil, out, n, stream);
        MXCO_HIP(e, "BloomFilter::Contains");
        if (sync) MXCO_HIP(hipStreamSynchronize(stream), "BloomFilter::Contains");
        return ok();
    }
    const unsigned grid = grid_for(n, 8);
    if (f->key_size == 4)
        bloom_contains_kernel<4><<<grid, 256, 0, stream>>>(f->words, (uint32_t)f->num_blocks, (const uint8_t*)keys,
                                                          stencil, out, n, f->seed);
    else
        bloom_contains_kernel<8><<<grid, 256, 0, stream>>>(f->words, (uint32_t)f->num_blocks, (const uint8_t*)keys,
                                                          stencil, out, n, f->seed);
    MXCO_HIP(hipGetLastError(), "BloomFilter::Contains");
    if (sync) MXCO_HIP(hipStreamSynchronize(stream), "BloomFilter::Contains");
    return ok();
}

MXCO_API mxco_status mxco_bloom_merge(mxco_bloom* dst, const mxco_bloom* src, hipStream_t stream, int sync)
{
    if (mxco_status s = compatible(dst, src, "Merge")) return s;
    if (dst == src) return ok();
    if (mxco_status s = words_op(dst, src, 0, stream, sync, "BloomFilter::Merge")) return s;
    return ok();
}

MXCO_API mxco_status mxco_bloom_intersect(mxco_bloom* dst, const mxco_bloom* src, hipStream_t stream, int sync)
{
    if (mxco_status s = compatible(dst, src, "Intersect")) return s;
    if (dst == src) return ok();
    if (mxco_status s = words_op(dst, src, 1, stream, sync, "BloomFilter::Intersect")) return s;
    return ok();
}

}  // extern "C"
