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
// The host parses and validates the headers once, keeps a device copy of the
// serialized bytes (Data()), and builds the line-aligned search image of
// mxco/detail/roaring.h: dense per-bucket container tables, fenced array and
// run containers, bitsets.  Each query is then a bucket lookup, a container
// lookup and a search inside one line of container data.
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
constexpr uint32_t kDenseMinContainers = 16;       // smaller buckets search a key list ...
constexpr uint32_t kDenseMinSingle = 2;            // ... unless the bitmap has one bucket
constexpr uint64_t kDenseBudget = 128ull << 20;    // bytes of dense tables per bitmap (256 tables)
constexpr uint64_t kBitsetImageMax = 1ull << 30;   // all containers as bitsets up to this many bytes
using detail::RbBucket;

struct HostContainer {  // one container of the parsed serialization
    uint64_t offset;    // byte offset of its data in the serialized bytes
    uint32_t key;       // bits 16..31 of the keys it holds
    uint32_t type;
    uint32_t cardinality;
};

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
mxco_status parse32(Reader& r, size_t base, std::vector<HostContainer>& dir, uint64_t* keys, size_t* length)
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
        dir.push_back({off, key, type, card});
        *keys += card;
        end = off + bytes;
    }
    if (!r.ok) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: truncated serialized bitmap");
    *length = end - base;
    return MXCO_OK;
}

struct Search {  // device pointers of the search image
    const RbBucket* buckets;
    const uint64_t* dense;
    const uint16_t* ckeys;
    const uint64_t* cdesc;
    const uint8_t* image;
    uint32_t nbuckets;
};

template <int KS, bool BitsetsOnly>
__global__ __launch_bounds__(256) void rb_contains_kernel(Search s, const uint8_t* keys, uint8_t* out, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t key;
        if constexpr (KS == 4) key = ((const uint32_t*)keys)[i];
        else key = ((const uint64_t*)keys)[i];
        out[i] = detail::rb_contains<BitsetsOnly>(s.buckets, s.nbuckets, s.dense, s.ckeys, s.cdesc, s.image, key) ? 1 : 0;
    }
}

// One-bucket bitmap with a dense container table over a bitset image: a
// query is a descriptor load and a byte load, and each lane keeps U queries
// in flight.  An absent container has descriptor 0, whose byte load reads
// the first bitset (the image holds at least two) and is masked out.
template <int KS>
__global__ __launch_bounds__(256) void rb_dense_bits_kernel(const uint64_t* dense, const uint8_t* image,
                                                            uint32_t bucket_key, const uint8_t* keys, uint8_t* out,
                                                            uint64_t n)
{
    constexpr int U = 4;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += stride * U) {
        uint64_t key[U], d[U];
        uint32_t b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t i = i0 + u * stride;
            key[u] = 0;
            if (i < n) key[u] = KS == 4 ? ((const uint32_t*)keys)[i] : ((const uint64_t*)keys)[i];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) d[u] = dense[(key[u] >> 16) & 0xffffu];
#pragma unroll
        for (int u = 0; u < U; ++u) b[u] = image[(d[u] & ((1ull << 40) - 1)) * 16 + ((key[u] & 0xffffu) >> 3)];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t i = i0 + u * stride;
            const bool hit = (d[u] >> 63) && (uint32_t)(key[u] >> 32) == bucket_key && ((b[u] >> (key[u] & 7)) & 1u);
            if (i < n) out[i] = hit ? 1 : 0;
        }
    }
}

// Small uint32 bitmaps (one bucket, at most kLdsContainers bitset
// containers): one persistent workgroup per CU copies the bitsets into LDS
// once, keeps the container keys in registers, and streams the queries four
// keys per 16-byte load, four loads in flight per lane, writing four result
// bytes per store.  A query is then register compares and one LDS byte read.
constexpr uint32_t kLdsContainers = 12;  // 96 KiB of LDS
constexpr uint32_t kLdsThreads = 1024;

__global__ __launch_bounds__(kLdsThreads) void rb_lds_bits_kernel(const uint8_t* image, const uint16_t* ckeys,
                                                                  uint32_t nc, const uint32_t* keys, uint8_t* out,
                                                                  uint64_t n)
{
    __shared__ uint4 lds[kLdsContainers * kBitsetBytes / 16];
    const uint4* src = (const uint4*)image;
    for (uint32_t j = threadIdx.x; j < nc * (kBitsetBytes / 16); j += blockDim.x) lds[j] = src[j];
    uint32_t kk[kLdsContainers];
#pragma unroll
    for (uint32_t j = 0; j < kLdsContainers; ++j) kk[j] = j < nc ? ckeys[j] : 0x10000u;  // 0x10000: never a key
    __syncthreads();
    const uint8_t* bits = (const uint8_t*)lds;
    auto test = [&](uint32_t key) -> uint32_t {
        const uint32_t k16 = key >> 16, v = key & 0xffffu;
        int c = -1;
#pragma unroll
        for (uint32_t j = 0; j < kLdsContainers; ++j) c = k16 == kk[j] ? (int)j : c;
        const uint32_t b = bits[(uint32_t)max(c, 0) * kBitsetBytes + (v >> 3)];
        return c >= 0 ? (b >> (v & 7)) & 1u : 0u;
    };
    constexpr int U = 4;
    const uint64_t n4 = n / 4, stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q0 < n4; q0 += stride * U) {
        uint4 kv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t q = q0 + u * stride;
            kv[u] = q < n4 ? ((const uint4*)keys)[q] : uint4{0, 0, 0, 0};
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t q = q0 + u * stride;
            if (q < n4)
                ((uint32_t*)out)[q] =
                    test(kv[u].x) | test(kv[u].y) << 8 | test(kv[u].z) << 16 | test(kv[u].w) << 24;
        }
    }
    for (uint64_t i = n4 * 4 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (uint8_t)test(keys[i]);
}

// Element count of a container in the image: values, or runs (the run
// count leads the serialized run container).
uint32_t elements(const uint8_t* ser, const HostContainer& c)
{
    if (c.type != kRun) return c.cardinality;
    return (uint32_t)ser[c.offset] | (uint32_t)ser[c.offset + 1] << 8;
}

uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// Image bytes of a container, and its alignment.
uint64_t image_bytes(const HostContainer& c, uint32_t n, bool as_bitset, uint64_t* align)
{
    if (c.type == kBitset || as_bitset) {
        *align = detail::kRbLine;
        return kBitsetBytes;
    }
    const uint32_t esz = c.type == kRun ? 4 : 2;
    const uint32_t fb = detail::rb_fence_bytes(n, esz);
    *align = fb ? detail::kRbLine : 16;
    return fb + align_up((uint64_t)n * esz, 16);
}

// Writes container c into the image at dst: fences (the first uint16 of every
// 128-byte line of elements), then the elements.
void write_image(uint8_t* dst, const uint8_t* ser, const HostContainer& c, uint32_t n, bool as_bitset)
{
    if (c.type == kBitset) {
        memcpy(dst, ser + c.offset, kBitsetBytes);
        return;
    }
    const uint32_t esz = c.type == kRun ? 4 : 2;
    const uint8_t* src = ser + c.offset + (c.type == kRun ? 2 : 0);  // runs: skip the count
    if (as_bitset) {  // dst holds 8192 zero bytes
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t start = (uint32_t)src[esz * i] | (uint32_t)src[esz * i + 1] << 8;
            const uint32_t len = c.type == kRun ? ((uint32_t)src[4 * i + 2] | (uint32_t)src[4 * i + 3] << 8) : 0;
            for (uint32_t v = start; v <= start + len && v < 65536; ++v) dst[v >> 3] |= (uint8_t)(1u << (v & 7));
        }
        return;
    }
    const uint32_t nf = detail::rb_fences(n, esz), per = detail::kRbLine / esz;
    for (uint32_t f = 0; f < nf; ++f) memcpy(dst + 2 * f, src + (uint64_t)f * per * esz, 2);
    memcpy(dst + detail::rb_fence_bytes(n, esz), src, (size_t)n * esz);
}

}  // namespace

struct mxco_rb {
    uint32_t ks;
    uint8_t* bytes = nullptr;  // device copy of the serialized bitmap
    RbBucket* buckets = nullptr;
    uint64_t* dense = nullptr;
    uint16_t* ckeys = nullptr;
    uint64_t* cdesc = nullptr;
    uint8_t* image = nullptr;
    uint32_t nbuckets = 0;
    bool bitsets_only = false;  // every container of the image is a bitset
    bool dense_bits = false;    // ... and the bitmap is one bucket with a dense table
    uint32_t bucket_key = 0;
    uint32_t bucket_containers = 0;  // containers of the first bucket
    uint64_t num_keys = 0;
    uint64_t size_bytes = 0;
};

namespace {

template <class T>
hipError_t upload(T** dst, const T* src, size_t count, hipStream_t stream)
{
    hipError_t e = hipMalloc((void**)dst, (count ? count : 1) * sizeof(T));
    if (e == hipSuccess && count) e = hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, stream);
    return e;
}

// Parses a whole 32-bit (key_size 4) or 64-bit bitmap: containers, buckets,
// key count and serialized length.
mxco_status parse_all(const uint8_t* ser, size_t bitmap_bytes, uint32_t key_size, std::vector<HostContainer>& dir,
                      std::vector<RbBucket>& buckets, uint64_t* keys, size_t* total)
{
    Reader r{ser, bitmap_bytes};
    if (key_size == 4) {
        size_t len = 0;
        if (mxco_status e = parse32(r, 0, dir, keys, &len)) return e;
        buckets.push_back({0, 0, (uint32_t)dir.size(), detail::kRbNoDense});
        *total = len;
        return MXCO_OK;
    }
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
        if (mxco_status e = parse32(r, cur, dir, keys, &len)) return e;
        buckets.push_back({hk, first, (uint32_t)(dir.size() - first), detail::kRbNoDense});
        cur += len;
    }
    *total = cur;
    return MXCO_OK;
}

// The search image of mxco/detail/roaring.h for parsed containers `dir`:
// every container expanded to an 8 KiB bitset while they all fit
// MXCO_RB_BITSET_MAX_MB (read at each construction; default 1024: one byte
// read per query), else
// aligned container copies; per bucket its container keys and descriptors
// from a 16-byte aligned index (buckets[].first is rewritten to it); a dense
// table for each bucket of at least kDenseMinContainers containers (of
// kDenseMinSingle in a one-bucket bitmap) within the budget; 128 bytes of
// padding after the key list and the image.
mxco_status build_search(const uint8_t* ser, const std::vector<HostContainer>& dir, std::vector<RbBucket>& buckets,
                         std::vector<uint8_t>& image, std::vector<uint64_t>& dense, std::vector<uint16_t>& ckeys,
                         std::vector<uint64_t>& cdesc, bool* bitsets_only)
{
    constexpr uint32_t kPadKeys = detail::kRbLine / sizeof(uint16_t);
    const uint64_t bitset_max = (uint64_t)env_int("MXCO_RB_BITSET_MAX_MB", (int)(kBitsetImageMax >> 20)) << 20;
    const bool as_bitset = (uint64_t)dir.size() * kBitsetBytes <= bitset_max;
    *bitsets_only = as_bitset;
    std::vector<uint64_t> at(dir.size());
    std::vector<uint64_t> desc(dir.size());
    uint64_t image_total = 0;
    try {
        for (size_t i = 0; i < dir.size(); ++i) {
            uint64_t align;
            const uint32_t n = elements(ser, dir[i]);
            const uint64_t bytes = image_bytes(dir[i], n, as_bitset, &align);
            at[i] = align_up(image_total, align);
            image_total = at[i] + bytes;
            desc[i] = as_bitset ? detail::rb_desc(at[i], kBitset, dir[i].cardinality)
                                : detail::rb_desc(at[i], dir[i].type, n);
        }
        uint64_t slots = 0;
        for (const RbBucket& b : buckets) slots = align_up(slots, 8) + b.count;
        ckeys.assign(slots + kPadKeys, 0);
        cdesc.assign(slots + kPadKeys, 0);
        uint64_t pos = 0;
        uint32_t ndense = 0;
        for (RbBucket& b : buckets) {
            pos = align_up(pos, 8);
            for (uint32_t j = 0; j < b.count; ++j) {
                ckeys[pos + j] = (uint16_t)dir[b.first + j].key;
                cdesc[pos + j] = desc[b.first + j];
            }
            b.first = (uint32_t)pos;
            pos += b.count;
            const uint32_t min_count = buckets.size() == 1 ? kDenseMinSingle : kDenseMinContainers;
            if (b.count < min_count || (uint64_t)(ndense + 1) * 65536 * 8 > kDenseBudget) continue;
            b.dense = ndense++;
            dense.resize((size_t)ndense * 65536, 0);
            for (uint32_t c = b.first; c < b.first + b.count; ++c)
                dense[(size_t)b.dense * 65536 + ckeys[c]] = cdesc[c];
        }
        image.assign(image_total + detail::kRbLine, 0);
    } catch (const std::bad_alloc&) {
        return fail(MXCO_OUT_OF_MEMORY, "RoaringBitmap: host allocation");
    }
    for (size_t i = 0; i < dir.size(); ++i)
        write_image(image.data() + at[i], ser, dir[i], elements(ser, dir[i]), as_bitset);
    return MXCO_OK;
}

}  // namespace

extern "C" {

MXCO_API mxco_status mxco_rb_create(uint32_t key_size, const void* bitmap, size_t bitmap_bytes, hipStream_t stream,
                                    mxco_rb** out)
{
    if (!out) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: null output");
    *out = nullptr;
    if (key_size != 4 && key_size != 8)
        return fail(MXCO_UNSUPPORTED, "RoaringBitmap key type must be uint32_t or uint64_t");
    if (!bitmap) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: bitmap must not be null");
    const uint8_t* ser = (const uint8_t*)bitmap;
    std::vector<HostContainer> dir;
    std::vector<RbBucket> buckets;
    uint64_t keys = 0;
    size_t total = 0;
    if (mxco_status e = parse_all(ser, bitmap_bytes, key_size, dir, buckets, &keys, &total)) return e;
    std::vector<uint8_t> image;
    std::vector<uint64_t> dense;
    std::vector<uint16_t> ckeys;
    std::vector<uint64_t> cdesc;
    bool bitsets_only = false;
    if (mxco_status e = build_search(ser, dir, buckets, image, dense, ckeys, cdesc, &bitsets_only)) return e;

    mxco_rb* rb = new (std::nothrow) mxco_rb();
    if (!rb) return fail(MXCO_OUT_OF_MEMORY, "RoaringBitmap: host allocation");
    rb->ks = key_size;
    rb->nbuckets = (uint32_t)buckets.size();
    rb->bitsets_only = bitsets_only;
    rb->dense_bits = bitsets_only && buckets.size() == 1 && buckets[0].dense != detail::kRbNoDense;
    rb->bucket_key = buckets.empty() ? 0 : buckets[0].key;
    rb->bucket_containers = buckets.empty() ? 0 : buckets[0].count;
    rb->num_keys = keys;
    rb->size_bytes = total;
    hipError_t e = upload(&rb->bytes, ser, total, stream);
    if (e == hipSuccess) e = upload(&rb->buckets, buckets.data(), buckets.size(), stream);
    if (e == hipSuccess) e = upload(&rb->dense, dense.data(), dense.size(), stream);
    if (e == hipSuccess) e = upload(&rb->ckeys, ckeys.data(), ckeys.size(), stream);
    if (e == hipSuccess) e = upload(&rb->cdesc, cdesc.data(), cdesc.size(), stream);
    if (e == hipSuccess) e = upload(&rb->image, image.data(), image.size(), stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);  // the host vectors end here
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
    for (void* p : {(void*)rb->bytes, (void*)rb->buckets, (void*)rb->dense, (void*)rb->ckeys, (void*)rb->cdesc,
                    (void*)rb->image})
        if (p) (void)hipFree(p);
    delete rb;
    return ok();
}

MXCO_API mxco_status mxco_rb_info(const mxco_rb* rb, mxco_rb_view* view)
{
    if (!rb) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: bitmap storage is not initialized");
    if (!view) return fail(MXCO_INVALID_ARGUMENT, "RoaringBitmap: null output");
    view->bytes = rb->bytes;
    view->buckets = rb->buckets;
    view->dense = rb->dense;
    view->container_keys = rb->ckeys;
    view->container_descs = rb->cdesc;
    view->image = rb->image;
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
    const Search s{rb->buckets, rb->dense, rb->ckeys, rb->cdesc, rb->image, rb->nbuckets};
    const uint8_t* k = (const uint8_t*)keys;
    if (rb->dense_bits && rb->ks == 4 && rb->bucket_containers <= kLdsContainers && n >= (1u << 16) &&
        ((uintptr_t)keys & 15u) == 0 && ((uintptr_t)out & 3u) == 0) {
        // one bucket of key 0: its containers are the image's first bitsets
        rb_lds_bits_kernel<<<cu_count(), kLdsThreads, 0, stream>>>(rb->image, rb->ckeys, rb->bucket_containers,
                                                                  (const uint32_t*)keys, out, n);
    } else if (rb->dense_bits) {
        const unsigned g = grid_for((n + 3) / 4, 8);
        if (rb->ks == 4) rb_dense_bits_kernel<4><<<g, 256, 0, stream>>>(rb->dense, rb->image, 0, k, out, n);
        else rb_dense_bits_kernel<8><<<g, 256, 0, stream>>>(rb->dense, rb->image, rb->bucket_key, k, out, n);
    } else if (rb->ks == 4) {
        if (rb->bitsets_only) rb_contains_kernel<4, true><<<grid, 256, 0, stream>>>(s, k, out, n);
        else rb_contains_kernel<4, false><<<grid, 256, 0, stream>>>(s, k, out, n);
    } else {
        if (rb->bitsets_only) rb_contains_kernel<8, true><<<grid, 256, 0, stream>>>(s, k, out, n);
        else rb_contains_kernel<8, false><<<grid, 256, 0, stream>>>(s, k, out, n);
    }
    MXCO_HIP(hipGetLastError(), "RoaringBitmap: contains");
    if (sync) MXCO_HIP(hipStreamSynchronize(stream), "RoaringBitmap: contains");
    return ok();
}

}  // extern "C"
