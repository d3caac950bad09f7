// RoaringBitmap (include/roaring_bitmap.h:24-143): membership queries of
// uint32 / uint64 keys against a bitmap in the Roaring portable
// serialization (the published RoaringFormatSpec, as the reference reads
// it in detail/roaring_bitmap/format.h):
//   32-bit: cookie 12347 | (containers - 1) << 16 followed by a run-container
//           bitmap, or cookie 12346 followed by a uint32 container count;
//           then (uint16 key, uint16 cardinality - 1) per container; then a
//           uint32 offset per container unless the cookie is 12347 and there
//           are fewer than 4 containers (offsets then follow from the
//           container sizes); then the containers: run (uint16 count, then
//           (start, length - 1) pairs), array (cardinality <= 4096 sorted
//           uint16) or bitset (8192 bytes).
//   64-bit: uint64 bucket count, then per bucket a uint32 high key and a
//           32-bit bitmap for the low halves.
// The host parses and validates the headers once; the serialized bytes and a
// flat container directory go to the device, and each query is a binary
// search of its bucket, then of its container, then inside the container.
#include <string.h>

#include <limits>
#include <new>
#include <string>
#include <vector>

#include "mxco/detail/roaring.h"
#include "mxco_common.h"

using namespace mxco;

namespace {

constexpr uint32_t kCookieNoRun = 12346;
constexpr uint32_t kCookieRun = 12347;
constexpr uint32_t kMaxContainers = 1u << 16;
constexpr uint32_t kMaxArray = 4096;
constexpr uint32_t kBitsetBytes = 8192;
constexpr uint32_t kNoOffsetThreshold = 4;

constexpr uint32_t kArray = detail::kRbArray, kBitset = detail::kRbBitset, kRun = detail::kRbRun;
using detail::RbBucket;
using detail::RbContainer;

// Bounds-checked little-endian reads of the host copy.
struct Reader {
    const uint8_t* p;
    size_t n;  // SIZE_MAX: unbounded (the caller vouches for the length)
    bool ok = true;
    bool has(size_t off, size_t len)
    {
        if (n == SIZE_MAX) return true;
        if (off > n || len > n - off) ok = false;
        return ok;
    }
    uint32_t u16(size_t off) { return has(off, 2) ? (uint32_t)p[off] | (uint32_t)p[off + 1] << 8 : 0; }
    uint32_t u32(size_t off) { return has(off, 4) ? u16(off) | u16(off + 2) << 16 : 0; }
    uint64_t u64(size_t off) { return (uint64_t)u32(off) | (uint64_t)u32(off + 4) << 32; }
};

// Parses one 32-bit bitmap starting at `base`; appends its containers to
// `dir` (offsets relative to the whole buffer) and returns its byte length.
mxco_status parse32(Reader& r, size_t base, std::vector<RbContainer>& dir, uint64_t* keys, size_t* length)
{
    const uint32_t cookie = r.u32(base);
    size_t cur = base + 4;
    uint32_t count;
    bool runs;
    size_t run_bits = 0;
    if ((cookie & 0xffffu) == kCookieRun) {
        runs = true;
        count = (cookie >> 16) + 1;
        run_bits = cur;
        cur += (count + 7) / 8;
    } else if (cookie == kCookieNoRun) {
        runs = false;
        count = r.u32(cur);
        cur += 4;
    } else {
        return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: unsupported serialized cookie");
    }
    if (!r.ok) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: truncated serialized bitmap");
    if (count > kMaxContainers) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: container count is out of range");
    const size_t headers = cur;
    if (!r.has(headers, 4ull * count)) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: truncated serialized bitmap");
    cur += 4ull * count;
    const bool offsets = !runs || count >= kNoOffsetThreshold;
    const size_t offset_table = cur;
    if (offsets) {
        if (!r.has(offset_table, 4ull * count))
            return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: truncated serialized bitmap");
        cur += 4ull * count;
    }
    size_t end = cur;
    uint32_t prev_key = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t key = r.u16(headers + 4ull * i);
        const uint32_t card = r.u16(headers + 4ull * i + 2) + 1;
        if (i && key <= prev_key)
            return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: container keys must be strictly increasing");
        prev_key = key;
        const bool is_run = runs && ((r.p[run_bits + i / 8] >> (i % 8)) & 1u);
        size_t off = end;
        if (offsets) {
            off = base + r.u32(offset_table + 4ull * i);  // offsets count from the bitmap's first byte
            if (off < cur || off < end) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: invalid container offset");
        }
        size_t bytes;
        uint32_t type;
        if (is_run) {
            type = kRun;
            bytes = 2 + 4ull * r.u16(off);
        } else if (card <= kMaxArray) {
            type = kArray;
            bytes = 2ull * card;
        } else {
            type = kBitset;
            bytes = kBitsetBytes;
        }
        if (!r.has(off, bytes)) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: truncated serialized bitmap");
        dir.push_back({off, key, type, card, 0});
        *keys += card;
        end = off + bytes;
    }
    if (!r.ok) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: truncated serialized bitmap");
    *length = end - base;
    return MXCO_OK;
}

template <int KS>
__global__ __launch_bounds__(256) void rb_contains_kernel(const uint8_t* bytes, const RbBucket* buckets,
                                                          uint32_t nbuckets, const RbContainer* dir,
                                                          const uint8_t* keys, uint8_t* out, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t key;
        if constexpr (KS == 4) key = ((const uint32_t*)keys)[i];
        else key = ((const uint64_t*)keys)[i];
        out[i] = detail::rb_contains(bytes, buckets, nbuckets, dir, key) ? 1 : 0;
    }
}

}  // namespace

struct mxco_rb {
    uint32_t ks;
    uint8_t* bytes;  // device copy of the serialized bitmap
    RbBucket* buckets;
    RbContainer* dir;
    uint32_t nbuckets;
    uint64_t num_keys;
    uint64_t size_bytes;
};

extern "C" {

MXCO_API mxco_status mxco_rb_create(uint32_t key_size, const void* bitmap, size_t bitmap_bytes, hipStream_t stream,
                                    mxco_rb** out)
{
    if (!out) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: null output");
    *out = nullptr;
    if (key_size != 4 && key_size != 8)
        return fail(MXCO_UNSUPPORTED, "RoaringBitmap key type must be uint32_t or uint64_t");
    if (!bitmap) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: bitmap must not be null");
    Reader r{(const uint8_t*)bitmap, bitmap_bytes};
    std::vector<RbContainer> dir;
    std::vector<RbBucket> buckets;
    uint64_t keys = 0;
    size_t total = 0;
    if (key_size == 4) {
        size_t len = 0;
        if (mxco_status e = parse32(r, 0, dir, &keys, &len)) return e;
        buckets.push_back({0, 0, (uint32_t)dir.size(), 0});
        total = len;
    } else {
        const uint64_t nb = r.u64(0);
        if (!r.ok) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: truncated serialized bitmap");
        if (bitmap_bytes != SIZE_MAX && nb > (bitmap_bytes - 8) / 12)
            return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: truncated 64-bit bucket table");
        if (nb > 0xffffffffull) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: bucket count is out of range");
        size_t cur = 8;
        for (uint64_t b = 0; b < nb; ++b) {
            const uint32_t hk = r.u32(cur);
            if (!r.ok) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: truncated serialized bitmap");
            if (b && hk <= buckets.back().key)
                return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: bucket keys must be strictly increasing");
            cur += 4;
            const uint32_t first = (uint32_t)dir.size();
            size_t len = 0;
            if (mxco_status e = parse32(r, cur, dir, &keys, &len)) return e;
            buckets.push_back({hk, first, (uint32_t)(dir.size() - first), 0});
            cur += len;
        }
        total = cur;
    }
    mxco_rb* rb = new (std::nothrow) mxco_rb();
    if (!rb) return fail(MXCO_OUT_OF_MEMORY, "RoaringBitmap: host allocation");
    rb->ks = key_size;
    rb->nbuckets = (uint32_t)buckets.size();
    rb->num_keys = keys;
    rb->size_bytes = total;
    hipError_t e = hipMalloc((void**)&rb->bytes, total ? total : 1);
    if (e == hipSuccess) e = hipMalloc((void**)&rb->buckets, (buckets.size() + 1) * sizeof(RbBucket));
    if (e == hipSuccess) e = hipMalloc((void**)&rb->dir, (dir.size() + 1) * sizeof(RbContainer));
    if (e == hipSuccess && total) e = hipMemcpyAsync(rb->bytes, bitmap, total, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && !buckets.empty())
        e = hipMemcpyAsync(rb->buckets, buckets.data(), buckets.size() * sizeof(RbBucket), hipMemcpyHostToDevice,
                           stream);
    if (e == hipSuccess && !dir.empty())
        e = hipMemcpyAsync(rb->dir, dir.data(), dir.size() * sizeof(RbContainer), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        (void)mxco_rb_destroy(rb);
        return hip_fail(e, "RoaringBitmap: device copy");
    }
    *out = rb;
    return ok();
}

MXCO_API mxco_status mxco_rb_destroy(mxco_rb* rb)
{
    if (!rb) return ok();
    if (rb->bytes) (void)hipFree(rb->bytes);
    if (rb->buckets) (void)hipFree(rb->buckets);
    if (rb->dir) (void)hipFree(rb->dir);
    delete rb;
    return ok();
}

MXCO_API mxco_status mxco_rb_info(const mxco_rb* rb, mxco_rb_view* view)
{
    if (!rb) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: bitmap storage is not initialized");
    if (!view) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: null output");
    view->bytes = rb->bytes;
    view->buckets = rb->buckets;
    view->containers = rb->dir;
    view->num_buckets = rb->nbuckets;
    view->reserved = 0;
    view->num_keys = rb->num_keys;
    view->size_bytes = rb->size_bytes;
    return ok();
}

MXCO_API mxco_status mxco_rb_contains(const mxco_rb* rb, const void* keys, uint8_t* out, uint64_t n,
                                      hipStream_t stream, int sync)
{
    if (n == 0) return ok();
    if (!keys || !out)
        return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap::ContainsAsync keys and outputValues must not be null");
    if (!rb) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap::ContainsAsync bitmap storage is not initialized");
    const unsigned grid = grid_for(n, 8);
    if (rb->ks == 4)
        rb_contains_kernel<4><<<grid, 256, 0, stream>>>(rb->bytes, rb->buckets, rb->nbuckets, rb->dir,
                                                        (const uint8_t*)keys, out, n);
    else
        rb_contains_kernel<8><<<grid, 256, 0, stream>>>(rb->bytes, rb->buckets, rb->nbuckets, rb->dir,
                                                        (const uint8_t*)keys, out, n);
    MXCO_HIP(hipGetLastError(), "RoaringBitmap: contains");
    if (sync) MXCO_HIP(hipStreamSynchronize(stream), "RoaringBitmap: contains");
    return ok();
}

}  // extern "C"
