// HyperLogLog (include/hyperloglog.h:24-59) for gfx950.
//
// Registers: 2^p bytes packed four per 32-bit word, byte i = register i
// (hyperloglog_ref.h:30-47).  For a key, h = XXH64(key, 0):
//   index = h >> (64 - p),  rank = clz((h << p) | 2^(p-1)) + 1,
// and the register keeps the maximum rank.
//
// Add (p <= 17): every workgroup keeps a private sketch in LDS (<= 128 KiB of
// the 160 KiB per CU), updates it with 32-bit LDS compare-and-swap only when
// the new rank is larger, and writes it out once; a reduce kernel takes the
// byte-wise maximum of the private sketches and the resident registers.
// p = 18 (256 KiB) does not fit LDS and updates the HBM registers directly
// with read-then-CAS.  Max is order-free, so every path gives the same
// registers.  Estimate builds the 65-bin rank histogram on the device and
// applies the HLL++ finalizer in FP64 on the host
// (hyperloglog.inl:328-345, finalizer.h:46-151).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>

#include "mxco_common.h"
#include "mxco/detail/hash.h"

using namespace mxco;

#include "hllpp_tables.inc"

struct mxco_hll {
    uint32_t* words;      // 2^p / 4 words
    uint32_t* hist;       // device, 65 bins
    uint32_t* hist_host;  // pinned mirror
    uint8_t* workspace;   // private sketches, ws_blocks x 2^p bytes
    uint32_t ws_blocks;
    uint32_t p;
    uint32_t key_size;
    int device;
};

namespace {

constexpr int kBins = 65;
constexpr int kThreads = 1024;

template <int KS>
__device__ __forceinline__ void hll_index_rank(uint64_t key, uint32_t p, uint32_t& idx, uint32_t& rank)
{
    const uint64_t h = xxhash64<KS>(key, 0);
    idx = (uint32_t)(h >> (64 - p));
    const uint64_t tail = (h << p) | (1ull << (p - 1));
    rank = (uint32_t)__clzll((long long)tail) + 1u;
}

template <int KS>
__device__ __forceinline__ uint64_t key_at(const uint8_t* keys, uint64_t i)
{
    if constexpr (KS == 4) return (uint64_t)(int64_t)((const int32_t*)keys)[i];
    else return ((const uint64_t*)keys)[i];
}

__device__ __forceinline__ uint32_t bytemax(uint32_t a, uint32_t b)
{
    uint32_t r = 0;
#pragma unroll
    for (int s = 0; s < 32; s += 8) r |= max((a >> s) & 255u, (b >> s) & 255u) << s;
    return r;
}

// Raise byte `shift/8` of *w to at least `rank` (LDS or global).
template <typename W>
__device__ __forceinline__ void raise_byte(W* w, uint32_t shift, uint32_t rank)
{
    uint32_t obs = *w;
    while (((obs >> shift) & 255u) < rank) {
        const uint32_t want = (obs & ~(255u << shift)) | (rank << shift);
        const uint32_t prev = atomicCAS((unsigned int*)w, obs, want);
        if (prev == obs) break;
        obs = prev;
    }
}

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

template <int KS>
__global__ __launch_bounds__(kThreads) void hll_private_add_kernel(const uint8_t* keys, uint64_t n, uint32_t p,
                                                                   uint8_t* workspace, int pairs16)
{
    extern __shared__ uint32_t sk[];
    const uint32_t nwords = (1u << p) / 4;
    for (uint32_t j = threadIdx.x; j < nwords; j += blockDim.x) sk[j] = 0;
    __syncthreads();
    auto add = [&](uint64_t key) {
        uint32_t idx, rank;
        hll_index_rank<KS>(key, p, idx, rank);
        raise_byte(&sk[idx >> 2], (idx & 3u) * 8u, rank);
    };
    // Contiguous chunk per workgroup so each input line is read by one CU;
    // an even chunk length keeps every chunk 16-byte aligned for int64 keys.
    const uint64_t per = ((n + gridDim.x - 1) / gridDim.x + 1) & ~1ull;
    const uint64_t begin = min(n, (uint64_t)blockIdx.x * per);
    const uint64_t end = min(n, begin + per);
    constexpr int U = 4;  // loads in flight per lane
    if (KS == 8 && pairs16) {
        // two keys per 16-byte load (keys 16-byte aligned)
        const uint64_t np = (end - begin) / 2;
        const u64x2* kp = (const u64x2*)(keys + begin * 8);
        for (uint64_t j0 = threadIdx.x; j0 < np; j0 += (uint64_t)U * blockDim.x) {
            u64x2 k[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t j = j0 + (uint64_t)u * blockDim.x;
                k[u] = j < np ? kp[j] : u64x2{0, 0};
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (j0 + (uint64_t)u * blockDim.x >= np) break;
                add(k[u].x);
                add(k[u].y);
            }
        }
        if (((end - begin) & 1) && threadIdx.x == 0) add(key_at<KS>(keys, end - 1));
    } else {
        for (uint64_t i0 = begin + threadIdx.x; i0 < end; i0 += (uint64_t)U * blockDim.x) {
            uint64_t k[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t i = i0 + (uint64_t)u * blockDim.x;
                k[u] = i < end ? key_at<KS>(keys, i) : 0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (i0 + (uint64_t)u * blockDim.x >= end) break;
                add(k[u]);
            }
        }
    }
    __syncthreads();
    uint32_t* out = (uint32_t*)(workspace + (uint64_t)blockIdx.x * (1u << p));
    for (uint32_t j = threadIdx.x; j < nwords; j += blockDim.x) out[j] = sk[j];
}

// Byte-max of the nws private sketches into the registers.  A block reduces
// 64 consecutive words: its 4 wavefronts take every 4th sketch (coalesced
// 256-byte rows, 8 independent loads in flight per lane) and combine
// through LDS.  Grid: ceil(words / 64) blocks of 256 lanes.
__global__ __launch_bounds__(256) void hll_reduce_kernel(uint32_t* words, const uint8_t* workspace, uint32_t nws,
                                                         uint32_t p)
{
    __shared__ uint32_t part[4][64];
    const uint32_t nwords = (1u << p) / 4;
    const uint32_t lane = threadIdx.x & 63u, slice = threadIdx.x >> 6;
    const uint32_t j = blockIdx.x * 64u + lane;
    uint32_t r = 0;
    if (j < nwords) {
        const uint32_t* base = (const uint32_t*)workspace + j;
        uint32_t b = slice;
        for (; b + 28u < nws; b += 32u) {
            uint32_t v[8];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) v[u] = base[(uint64_t)(b + 4u * u) * nwords];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) r = bytemax(r, v[u]);
        }
        for (; b < nws; b += 4u) r = bytemax(r, base[(uint64_t)b * nwords]);
    }
    part[slice][lane] = r;
    __syncthreads();
    if (slice == 0 && j < nwords)
        words[j] = bytemax(bytemax(words[j], part[0][lane]), bytemax(bytemax(part[1][lane], part[2][lane]), part[3][lane]));
}

template <int KS>
__global__ __launch_bounds__(256) void hll_global_add_kernel(uint32_t* words, const uint8_t* keys, uint64_t n,
                                                             uint32_t p)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t idx, rank;
        hll_index_rank<KS>(key_at<KS>(keys, i), p, idx, rank);
        raise_byte(&words[idx >> 2], (idx & 3u) * 8u, rank);
    }
}

// op 0: clear, op 1: dst = bytemax(dst, src)
__global__ __launch_bounds__(256) void hll_words_kernel(uint32_t* dst, const uint32_t* src, uint32_t nwords, int op)
{
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nwords; j += gridDim.x * blockDim.x)
        dst[j] = op == 0 ? 0u : bytemax(dst[j], src[j]);
}

__global__ __launch_bounds__(256) void hll_histogram_kernel(const uint32_t* words, uint32_t nwords, uint32_t* hist)
{
    __shared__ uint32_t lh[kBins];
    for (int b = threadIdx.x; b < kBins; b += blockDim.x) lh[b] = 0;
    __syncthreads();
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nwords; j += gridDim.x * blockDim.x) {
        const uint32_t w = words[j];
#pragma unroll
        for (int s = 0; s < 32; s += 8) atomicAdd(&lh[min((w >> s) & 255u, (uint32_t)(kBins - 1))], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kBins; b += blockDim.x)
        if (lh[b]) atomicAdd(&hist[b], lh[b]);
}

__global__ void hll_zero_hist_kernel(uint32_t* hist)
{
    if (threadIdx.x < kBins) hist[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// HLL++ finalizer (finalizer.h:36-151), FP64 on the host.
// ---------------------------------------------------------------------------
struct Tables {
    const double* raw;
    const double* bias;
    int n;
};

Tables tables(uint32_t p)
{
    switch (p) {
    case 13: return {kHllRaw13, kHllBias13, (int)(sizeof(kHllRaw13) / sizeof(double))};
    case 14: return {kHllRaw14, kHllBias14, (int)(sizeof(kHllRaw14) / sizeof(double))};
    case 15: return {kHllRaw15, kHllBias15, (int)(sizeof(kHllRaw15) / sizeof(double))};
    case 16: return {kHllRaw16, kHllBias16, (int)(sizeof(kHllRaw16) / sizeof(double))};
    case 17: return {kHllRaw17, kHllBias17, (int)(sizeof(kHllRaw17) / sizeof(double))};
    default: return {kHllRaw18, kHllBias18, (int)(sizeof(kHllRaw18) / sizeof(double))};
    }
}

double alpha_mm(double m)
{
    if (m == 16) return 0.673 * m * m;
    if (m == 32) return 0.697 * m * m;
    if (m == 64) return 0.709 * m * m;
    return (0.7213 / (1.0 + 1.079 / m)) * m * m;
}

// Mean bias of the k = 6 raw-estimate neighbours of e.
double bias_of(double e, const Tables& t)
{
    const int k = 6;
    int left = 0, right = t.n - 1, anchor = -1;
    bool exact = false;
    while (left <= right) {
        const int mid = left + (right - left) / 2;
        if (t.raw[mid] < e) left = mid + 1;
        else if (t.raw[mid] > e) right = mid - 1;
        else {
            anchor = mid;
            exact = true;
            break;
        }
    }
    if (!exact) anchor = left;
    int low = std::max(std::min(anchor - k + 1, t.n - k), 0);
    int high = std::min(low + k, t.n);
    auto dist = [&](int i) { return (e - t.raw[i]) * (e - t.raw[i]); };
    while (high < t.n && dist(high) < dist(low)) {
        ++low;
        ++high;
    }
    double s = 0.0;
    for (int i = low; i < high; ++i) s += t.bias[i];
    return s / (high - low);
}

uint64_t finalize(uint32_t p, double z, int v)
{
    if (z == 0.0) return UINT64_MAX;
    const double m = (double)(1u << p);
    const Tables t = tables(p);
    double e = alpha_mm(m) / z;
    if (v > 0) {
        const double h = m * log(m / v);
        if (e <= 2.5 * m) return (uint64_t)round(h);
        e = (h <= kHllThreshold[p - 4]) ? h : ((e < 5.0 * m) ? e - bias_of(e, t) : e);
    } else {
        e = (e < 5.0 * m) ? e - bias_of(e, t) : e;
    }
    if (e >= 18446744073709551616.0) return UINT64_MAX;
    return (uint64_t)round(e);
}

mxco_status check_sketch(const mxco_hll* h, const char* fn)
{
    if (!h || !h->words) return fail(MXCO_LOGIC_ERROR, std::string("HyperLogLog: ") + fn + ": moved-from or uninitialized sketch");
    int dev = 0;
    MXCO_HIP(hipGetDevice(&dev), "HyperLogLog: get device");
    if (dev != h->device)
        return fail(MXCO_INVALID_ARGUMENT, "HyperLogLog: current device differs from owner device");
    return MXCO_OK;
}

mxco_status words_op(mxco_hll* dst, const mxco_hll* src, int op, hipStream_t s, const char* fn)
{
    const uint32_t nwords = (1u << dst->p) / 4;
    hll_words_kernel<<<grid_for(nwords, 4), 256, 0, s>>>(dst->words, src ? src->words : nullptr, nwords, op);
    MXCO_HIP(hipGetLastError(), fn);
    MXCO_HIP(hipStreamSynchronize(s), fn);
    return MXCO_OK;
}

// Inputs below this size skip the private sketches (their fixed cost is a
// zero + write-out of 2^p bytes per workgroup plus the reduce).
uint64_t direct_threshold(uint32_t p) { return (uint64_t)1 << (p + 3); }

}  // namespace

extern "C" {

MXCO_API mxco_status mxco_hll_precision_for_sketch_kb(uint32_t kb, uint32_t* precision)
{
    if (!precision) return fail(MXCO_INVALID_ARGUMENT, "HyperLogLog: null output");
    if (kb < 8 || kb > 256 || (kb & (kb - 1)) != 0) return fail(MXCO_INVALID_ARGUMENT, "HyperLogLog: invalid sketch size");
    uint32_t p = 13;
    while ((1u << p) < kb * 1024u) ++p;
    *precision = p;
    return ok();
}

MXCO_API mxco_status mxco_hll_precision_for_stddev(double deviation, uint32_t* precision)
{
    if (!precision) return fail(MXCO_INVALID_ARGUMENT, "HyperLogLog: null output");
    if (!isfinite(deviation) || deviation <= 0) return fail(MXCO_INVALID_ARGUMENT, "HyperLogLog: invalid deviation");
    for (uint32_t p = 13; p <= 18; ++p) {
        if (1.04 / sqrt((double)(1u << p)) <= deviation) {
            *precision = p;
            return ok();
        }
    }
    return fail(MXCO_INVALID_ARGUMENT, "HyperLogLog: deviation requires more than 256 KiB");
}

MXCO_API mxco_status mxco_hll_create(uint32_t precision, uint32_t key_size, hipStream_t stream, mxco_hll** out)
{
    if (!out) return fail(MXCO_INVALID_ARGUMENT, "HyperLogLog: null output");
    *out = nullptr;
    if (precision < 13 || precision > 18) return fail(MXCO_INVALID_ARGUMENT, "HyperLogLog: precision must be in [13, 18]");
    if (key_size != 4 && key_size != 8) return fail(MXCO_UNSUPPORTED, "HyperLogLog supports int32_t and int64_t");
    mxco_hll* h = new (std::nothrow) mxco_hll();
    if (!h) return fail(MXCO_OUT_OF_MEMORY, "HyperLogLog: host allocation");
    h->p = precision;
    h->key_size = key_size;
    (void)hipGetDevice(&h->device);
    // private sketches: two workgroups per CU while two sketches fit in LDS
    h->ws_blocks = precision <= 17 ? (uint32_t)cu_count() * (precision <= 16 ? 2u : 1u) : 0;
    const size_t sketch = (size_t)1 << precision;
    hipError_t e = hipMalloc((void**)&h->words, sketch);
    if (e == hipSuccess) e = hipMalloc((void**)&h->hist, kBins * sizeof(uint32_t));
    if (e == hipSuccess) e = hipHostMalloc((void**)&h->hist_host, kBins * sizeof(uint32_t), 0);
    if (e == hipSuccess && h->ws_blocks) e = hipMalloc((void**)&h->workspace, sketch * h->ws_blocks);
    if (e == hipSuccess && precision <= 17) {
        e = hipFuncSetAttribute((const void*)hll_private_add_kernel<4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)sketch);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)hll_private_add_kernel<8>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)sketch);
    }
    mxco_status st = e == hipSuccess ? MXCO_OK : hip_fail(e, "HyperLogLog: allocate sketch and workspace");
    if (st == MXCO_OK) st = words_op(h, nullptr, 0, stream, "HyperLogLog: clear");
    if (st != MXCO_OK) {
        if (h->words) (void)hipFree(h->words);
        if (h->hist) (void)hipFree(h->hist);
        if (h->hist_host) (void)hipHostFree(h->hist_host);
        if (h->workspace) (void)hipFree(h->workspace);
        delete h;
        return st;
    }
    *out = h;
    return ok();
}

MXCO_API mxco_status mxco_hll_destroy(mxco_hll* h)
{
    if (!h) return ok();
    if (h->words) (void)hipFree(h->words);
    if (h->hist) (void)hipFree(h->hist);
    if (h->hist_host) (void)hipHostFree(h->hist_host);
    if (h->workspace) (void)hipFree(h->workspace);
    delete h;
    return ok();
}

MXCO_API mxco_status mxco_hll_data(const mxco_hll* h, const uint8_t** registers, uint32_t* precision)
{
    if (!h || !h->words) return fail(MXCO_LOGIC_ERROR, "HyperLogLog: moved-from or uninitialized sketch");
    if (registers) *registers = (const uint8_t*)h->words;
    if (precision) *precision = h->p;
    return ok();
}

MXCO_API mxco_status mxco_hll_clear(mxco_hll* h, hipStream_t stream)
{
    if (mxco_status s = check_sketch(h, "Clear")) return s;
    if (mxco_status s = words_op(h, nullptr, 0, stream, "HyperLogLog: clear")) return s;
    return ok();
}

MXCO_API mxco_status mxco_hll_add(mxco_hll* h, const void* keys, uint64_t n, hipStream_t stream)
{
    if (mxco_status s = check_sketch(h, "Add")) return s;
    if (n == 0) {
        MXCO_HIP(hipStreamSynchronize(stream), "HyperLogLog: synchronize stream");
        return ok();
    }
    if (!keys || ((uintptr_t)keys % h->key_size) != 0)
        return fail(MXCO_INVALID_ARGUMENT, "HyperLogLog: invalid input pointer or size");
    const uint8_t* k = (const uint8_t*)keys;
    if (h->p <= 17 && n >= direct_threshold(h->p)) {
        const uint32_t blocks = h->ws_blocks;
        const size_t lds = (size_t)1 << h->p;
        if (h->key_size == 4)
            hll_private_add_kernel<4><<<blocks, kThreads, lds, stream>>>(k, n, h->p, h->workspace, 0);
        else
            hll_private_add_kernel<8><<<blocks, kThreads, lds, stream>>>(k, n, h->p, h->workspace,
                                                                          ((uintptr_t)k & 15u) == 0);
        MXCO_HIP(hipGetLastError(), "HyperLogLog: add");
        const uint32_t nwords = (1u << h->p) / 4;
        hll_reduce_kernel<<<(nwords + 63) / 64, 256, 0, stream>>>(h->words, h->workspace, blocks, h->p);
        MXCO_HIP(hipGetLastError(), "HyperLogLog: reduce");
    } else {
        const unsigned grid = grid_for(n, 8);
        if (h->key_size == 4) hll_global_add_kernel<4><<<grid, 256, 0, stream>>>(h->words, k, n, h->p);
        else hll_global_add_kernel<8><<<grid, 256, 0, stream>>>(h->words, k, n, h->p);
        MXCO_HIP(hipGetLastError(), "HyperLogLog: add");
    }
    MXCO_HIP(hipStreamSynchronize(stream), "HyperLogLog: synchronize stream");
    return ok();
}

MXCO_API mxco_status mxco_hll_merge(mxco_hll* dst, const mxco_hll* src, hipStream_t stream)
{
    if (mxco_status s = check_sketch(dst, "Merge")) return s;
    if (mxco_status s = check_sketch(src, "Merge")) return s;
    if (dst->p != src->p || dst->device != src->device)
        return fail(MXCO_INVALID_ARGUMENT, "HyperLogLog: incompatible sketches");
    if (dst == src) {
        MXCO_HIP(hipStreamSynchronize(stream), "HyperLogLog: synchronize stream");
        return ok();
    }
    if (mxco_status s = words_op(dst, src, 1, stream, "HyperLogLog: merge")) return s;
    return ok();
}

MXCO_API mxco_status mxco_hll_histogram(const mxco_hll* h, hipStream_t stream, uint32_t* hist65)
{
    if (mxco_status s = check_sketch(h, "Estimate")) return s;
    if (!hist65) return fail(MXCO_INVALID_ARGUMENT, "HyperLogLog: null histogram");
    hll_zero_hist_kernel<<<1, 128, 0, stream>>>(h->hist);
    const uint32_t nwords = (1u << h->p) / 4;
    hll_histogram_kernel<<<std::min<unsigned>(grid_for(nwords, 1), 64), 256, 0, stream>>>(h->words, nwords, h->hist);
    MXCO_HIP(hipGetLastError(), "HyperLogLog: histogram");
    MXCO_HIP(hipMemcpyAsync(h->hist_host, h->hist, kBins * sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
             "HyperLogLog: copy histogram");
    MXCO_HIP(hipStreamSynchronize(stream), "HyperLogLog: synchronize stream");
    memcpy(hist65, h->hist_host, kBins * sizeof(uint32_t));
    return ok();
}

MXCO_API mxco_status mxco_hll_estimate(const mxco_hll* h, hipStream_t stream, uint64_t* estimate)
{
    if (!estimate) return fail(MXCO_INVALID_ARGUMENT, "HyperLogLog: null output");
    uint32_t hist[kBins];
    if (mxco_status s = mxco_hll_histogram(h, stream, hist)) return s;
    double z = 0;
    for (int r = 0; r < kBins; ++r) z += ldexp((double)hist[r], -r);
    *estimate = finalize(h->p, z, (int)hist[0]);
    return ok();
}

}  // extern "C"
