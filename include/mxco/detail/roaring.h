// Device-side search structure of a RoaringBitmap and its membership test,
// shared by libmxco's query kernel and RoaringBitmapRef::Contains
// (include/roaring_bitmap_ref.h:47-160, 176-265 of the reference, which
// searches the serialized bytes directly).
//
// libmxco parses the portable serialization once on the host and builds,
// next to the device copy of the serialized bytes, a search image laid out
// for 128-byte cache lines:
//   - buckets: one per distinct high 32-bit half (a 32-bit bitmap is one
//     bucket of key 0), sorted by key;
//   - per bucket either a dense table of 65,536 container descriptors indexed
//     by bits 16..31 of the key, or a sorted uint16 list of container keys
//     with a parallel descriptor list;
//   - per container, an 8 KiB bitset (the default while all containers fit
//     the library's bitset budget: one byte read per query), or its data
//     copied to an aligned place in the image: bitset, array (sorted uint16
//     values) or run container
//     (sorted uint32 runs, start | (length - 1) << 16).  Arrays of more than
//     64 values and run containers of more than 32 runs start with a fence
//     list, the first element of every 128-byte line of elements, padded to
//     a whole line, so a query reads the fence line(s) and one line of data.
//   - bucket key lists start 16-byte aligned, and the key list and the image
//     are followed by 128 bytes of padding: a query reads a whole 128-byte
//     window with independent 16-byte loads and ignores what lies past the
//     elements it needs.
// A descriptor is a uint64: bit 63 set for a present container, bits 0..39
// the image offset in 16-byte units, bits 40..41 the type, bits 42..58 the
// element count (values for an array or bitset, runs for a run container).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

namespace mxco {
namespace detail {

enum : uint32_t { kRbArray = 0, kRbBitset = 1, kRbRun = 2 };
constexpr uint32_t kRbLine = 128;          // bytes of elements behind one fence
constexpr uint32_t kRbNoDense = 0xffffffffu;

struct RbBucket {
    uint32_t key;    // bits 32..63 of the keys it holds
    uint32_t first;  // first container of the bucket in the key/descriptor lists
    uint32_t count;
    uint32_t dense;  // index of the bucket's dense table, or kRbNoDense
};

inline constexpr uint64_t rb_desc(uint64_t image_offset, uint32_t type, uint32_t count)
{
    return (1ull << 63) | (image_offset / 16) | (uint64_t)type << 40 | (uint64_t)count << 42;
}

// Fences of a container of n elements of esz bytes, and the bytes they take
// in the image (0 when the elements fit one line).
inline constexpr uint32_t rb_fences(uint32_t n, uint32_t esz)
{
    return n * esz > kRbLine ? (n * esz + kRbLine - 1) / kRbLine : 0;
}
inline constexpr uint32_t rb_fence_bytes(uint32_t n, uint32_t esz)
{
    return (rb_fences(n, esz) * 2 + kRbLine - 1) / kRbLine * kRbLine;
}

#if defined(__HIPCC__)
// Number of elements in [lo, hi) whose first uint16 is <= v, plus lo
// (binary search; for fence lists longer than one line).
template <class E>
__host__ __device__ __forceinline__ uint32_t rb_upper(const E* e, uint32_t lo, uint32_t hi, uint32_t v)
{
    uint32_t len = hi - lo;
    while (len > 0) {
        const uint32_t half = len / 2;
        if ((uint32_t)(e[lo + half] & 0xffffu) <= v) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

struct RbScan {
    uint32_t count;  // elements <= v
    uint32_t last;   // the last of them (whole element)
};

// Scans the first m (<= 128 / sizeof(E)) sorted elements of the 128-byte
// window at p (16-byte aligned; the whole window must be readable) with up
// to eight independent 16-byte loads, only as many as the m elements need,
// comparing the low 16 bits of each element with v.
template <class E>
__host__ __device__ __forceinline__ RbScan rb_scan(const uint8_t* p, uint32_t m, uint32_t v)
{
    const uint32_t chunks = (m * (uint32_t)sizeof(E) + 15) / 16;
    uint4 w[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) w[j] = j < chunks ? ((const uint4*)p)[j] : uint4{0, 0, 0, 0};
    RbScan r{0, 0};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t ws[4] = {w[j].x, w[j].y, w[j].z, w[j].w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if constexpr (sizeof(E) == 2) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const uint32_t i = j * 8 + k * 2 + h, e = (ws[k] >> (16 * h)) & 0xffffu;
                    const bool le = i < m && e <= v;
                    r.count += le;
                    r.last = le ? e : r.last;
                }
            } else {
                const uint32_t i = j * 4 + k;
                const bool le = i < m && (ws[k] & 0xffffu) <= v;
                r.count += le;
                r.last = le ? ws[k] : r.last;
            }
        }
    }
    return r;
}

// Array (E = uint16_t) or run (E = uint32_t) container at p with n elements:
// the last element starting at or below v (count 0 for none).  The fence
// window (or, past 64 fences, a binary search of the fences) picks the data
// window, which is scanned.
template <class E>
__host__ __device__ __forceinline__ RbScan rb_search(const uint8_t* p, uint32_t n, uint32_t v)
{
    constexpr uint32_t per = kRbLine / sizeof(E);
    const uint32_t nf = rb_fences(n, sizeof(E));
    if (!nf) return rb_scan<E>(p, n, v);
    const uint32_t f = nf <= 64 ? rb_scan<uint16_t>(p, nf, v).count : rb_upper((const uint16_t*)p, 0, nf, v);
    if (f == 0) return RbScan{0, 0};
    const uint32_t lo = (f - 1) * per, m = n - lo < per ? n - lo : per;
    return rb_scan<E>(p + rb_fence_bytes(n, sizeof(E)) + (uint64_t)lo * sizeof(E), m, v);
}

// BitsetsOnly: the image holds bitset containers only (the library's default
// image), so the array and run paths are compiled out.
template <bool BitsetsOnly = false>
__host__ __device__ __forceinline__ bool rb_contains(const RbBucket* buckets, uint32_t nbuckets,
                                                     const uint64_t* dense, const uint16_t* ckeys,
                                                     const uint64_t* cdesc, const uint8_t* image, uint64_t key)
{
    const uint32_t hi = (uint32_t)(key >> 32);
    uint32_t b = 0;
    if (nbuckets != 1) {
        uint32_t cnt = nbuckets;  // first bucket with key >= hi
        while (cnt > 0) {
            const uint32_t half = cnt / 2;
            if (buckets[b + half].key < hi) {
                b += half + 1;
                cnt -= half + 1;
            } else {
                cnt = half;
            }
        }
        if (b >= nbuckets) return false;
    }
    const RbBucket bk = buckets[b];
    if (bk.key != hi) return false;
    const uint32_t k16 = (uint32_t)(key >> 16) & 0xffffu, v = (uint32_t)key & 0xffffu;
    uint64_t d;
    if (bk.dense != kRbNoDense) {
        d = dense[(uint64_t)bk.dense * 65536u + k16];
    } else {
        uint32_t c;
        if (bk.count <= 64) {  // the key list fits one window
            const RbScan s = rb_scan<uint16_t>((const uint8_t*)(ckeys + bk.first), bk.count, k16);
            if (!s.count || s.last != k16) return false;
            c = bk.first + s.count;
        } else {
            c = rb_upper(ckeys, bk.first, bk.first + bk.count, k16);
            if (c == bk.first || ckeys[c - 1] != k16) return false;
        }
        d = cdesc[c - 1];
    }
    if (!(d >> 63)) return false;
    const uint8_t* p = image + (d & ((1ull << 40) - 1)) * 16;
    const uint32_t type = (uint32_t)(d >> 40) & 3u, n = (uint32_t)(d >> 42) & 0x1ffffu;
    if (BitsetsOnly || type == kRbBitset) return (p[v >> 3] >> (v & 7)) & 1u;
    if (type == kRbArray) {
        const RbScan s = rb_search<uint16_t>(p, n, v);
        return s.count && s.last == v;
    }
    const RbScan s = rb_search<uint32_t>(p, n, v);
    return s.count && v <= (s.last & 0xffffu) + (s.last >> 16);
}
#endif

}  // namespace detail
}  // namespace mxco
