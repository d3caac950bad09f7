// Device-side layout and membership search of a parsed Roaring bitmap,
// shared by libmxco's query kernel and RoaringBitmapRef::Contains
// (include/roaring_bitmap_ref.h:47-160, 176-265 of the reference).
//
// The host parses the portable serialization once (libmxco) and keeps, next
// to the serialized bytes, a directory of containers (RbContainer, ordered by
// 16-bit key within each bucket) and a bucket table (RbBucket, ordered by
// the high 32 bits of a 64-bit key; a 32-bit bitmap is one bucket of key 0).
// A query is a binary search of its bucket, of its container, and inside the
// container: sorted uint16 values (array), 8192-byte bitset, or sorted
// (start, length - 1) runs after a uint16 run count.  The serialized bytes
// are read one byte at a time, so containers need no alignment.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

namespace mxco {
namespace detail {

enum : uint32_t { kRbArray = 0, kRbBitset = 1, kRbRun = 2 };

struct RbContainer {
    uint64_t offset;  // byte offset of the container in the serialized bytes
    uint32_t key;     // bits 16..31 of the keys it holds
    uint32_t type;    // kRbArray, kRbBitset or kRbRun
    uint32_t cardinality;
    uint32_t pad;
};

struct RbBucket {
    uint32_t key;    // bits 32..63 of the keys it holds
    uint32_t first;  // first container of the bucket in the directory
    uint32_t count;
    uint32_t pad;
};

#if defined(__HIPCC__)
__device__ __forceinline__ uint32_t rb_rd16(const uint8_t* p)
{
    return (uint32_t)p[0] | (uint32_t)p[1] << 8;
}

__device__ __forceinline__ bool rb_contains(const uint8_t* bytes, const RbBucket* buckets, uint32_t nbuckets,
                                            const RbContainer* dir, uint64_t key)
{
    const uint32_t hi = (uint32_t)(key >> 32);
    uint32_t lo = 0, cnt = nbuckets;  // first bucket with key >= hi
    while (cnt > 0) {
        const uint32_t half = cnt / 2;
        if (buckets[lo + half].key < hi) {
            lo += half + 1;
            cnt -= half + 1;
        } else {
            cnt = half;
        }
    }
    if (lo >= nbuckets || buckets[lo].key != hi) return false;
    const RbBucket b = buckets[lo];
    const uint32_t k16 = (uint32_t)(key >> 16) & 0xffffu, v = (uint32_t)key & 0xffffu;
    uint32_t c = b.first, m = b.count;  // first container with key >= k16
    while (m > 0) {
        const uint32_t half = m / 2;
        if (dir[c + half].key < k16) {
            c += half + 1;
            m -= half + 1;
        } else {
            m = half;
        }
    }
    if (c >= b.first + b.count || dir[c].key != k16) return false;
    const RbContainer ct = dir[c];
    const uint8_t* p = bytes + ct.offset;
    if (ct.type == kRbBitset) return (p[v >> 3] >> (v & 7)) & 1u;
    if (ct.type == kRbArray) {
        uint32_t a = 0, len = ct.cardinality;  // first value >= v
        while (len > 0) {
            const uint32_t half = len / 2;
            if (rb_rd16(p + 2 * (a + half)) < v) {
                a += half + 1;
                len -= half + 1;
            } else {
                len = half;
            }
        }
        return a < ct.cardinality && rb_rd16(p + 2 * a) == v;
    }
    const uint32_t nruns = rb_rd16(p);
    uint32_t a = 0, len = nruns;  // number of runs starting at or below v
    while (len > 0) {
        const uint32_t half = len / 2;
        if (rb_rd16(p + 2 + 4 * (a + half)) <= v) {
            a += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    if (a == 0) return false;
    return v <= rb_rd16(p + 2 + 4 * (a - 1)) + rb_rd16(p + 4 + 4 * (a - 1));
}
#endif

}  // namespace detail
}  // namespace mxco
