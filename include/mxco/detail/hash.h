// Hash functions on 1-, 2-, 4- and 8-byte keys, callable from host and device.
//
// Families and bit-exact behaviour follow the reference's hashers:
//   murmurhash3_32   include/detail/hash_functions/murmurhash3.h:56-108
//   xxhash_32        include/detail/hash_functions/xxhash.h:53-175
//   xxhash_64        include/detail/hash_functions/xxhash.h:185-326
//   murmurhash3_fmix include/detail/hash_functions/utils.h:42-64
// Only the key widths the containers accept (1, 2, 4 and 8 bytes) are
// written out; each is the published algorithm specialised to that many
// little-endian input bytes: one or two 32-bit blocks, or a tail of one or
// two bytes.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define MXCO_HD __host__ __device__ __forceinline__
#else
#define MXCO_HD inline
#endif

namespace mxco {

enum HashFamily : uint32_t { kMurmur3_32 = 0, kXXHash32 = 1, kXXHash64 = 2, kFmix32 = 3, kFmix64 = 4 };

MXCO_HD uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
MXCO_HD uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

MXCO_HD uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

MXCO_HD uint64_t fmix64(uint64_t h)
{
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

// MurmurHash3_x86_32: one body round per 32-bit block, then length and fmix.
MXCO_HD uint32_t murmur3_block(uint32_t h, uint32_t k)
{
    k *= 0xcc9e2d51u;
    k = rotl32(k, 15);
    k *= 0x1b873593u;
    h ^= k;
    h = rotl32(h, 13);
    return h * 5u + 0xe6546b64u;
}

template <int KS>
MXCO_HD uint32_t murmur3_32(uint64_t key, uint32_t seed)
{
    uint32_t h = seed;
    if (KS >= 4) {
        h = murmur3_block(h, (uint32_t)key);
        if (KS == 8) h = murmur3_block(h, (uint32_t)(key >> 32));
    } else {
        // tail of 1 or 2 bytes: mixed into h without the block rotation
        uint32_t k = (uint32_t)key & (KS == 1 ? 0xffu : 0xffffu);
        k *= 0xcc9e2d51u;
        k = rotl32(k, 15);
        k *= 0x1b873593u;
        h ^= k;
    }
    h ^= (uint32_t)KS;
    return fmix32(h);
}

// XXH32 for inputs shorter than 16 bytes: only the 4-byte tail loop runs.
template <int KS>
MXCO_HD uint32_t xxhash32(uint64_t key, uint32_t seed)
{
    const uint32_t p1 = 0x9e3779b1u, p2 = 0x85ebca77u, p3 = 0xc2b2ae3du, p4 = 0x27d4eb2fu, p5 = 0x165667b1u;
    uint32_t h = seed + p5 + (uint32_t)KS;
    if (KS >= 4) {
        h += (uint32_t)key * p3;
        h = rotl32(h, 17) * p4;
        if (KS == 8) {
            h += (uint32_t)(key >> 32) * p3;
            h = rotl32(h, 17) * p4;
        }
    } else {
        // byte tail
        for (int b = 0; b < KS; ++b) {
            h += (uint32_t)((key >> (8 * b)) & 0xffu) * p5;
            h = rotl32(h, 11) * p1;
        }
    }
    h ^= h >> 15;
    h *= p2;
    h ^= h >> 13;
    h *= p3;
    h ^= h >> 16;
    return h;
}

// XXH64 for 1-, 2-, 4- or 8-byte inputs: the byte, 4-byte or 8-byte tail steps.
template <int KS>
MXCO_HD uint64_t xxhash64(uint64_t key, uint64_t seed)
{
    const uint64_t p1 = 11400714785074694791ull, p2 = 14029467366897019727ull, p3 = 1609587929392839161ull,
                   p4 = 9650029242287828579ull, p5 = 2870177450012600261ull;
    uint64_t h = seed + p5 + (uint64_t)KS;
    if (KS == 8) {
        uint64_t lane = rotl64(key * p2, 31) * p1;
        h ^= lane;
        h = rotl64(h, 27) * p1 + p4;
    } else if (KS == 4) {
        h ^= (uint64_t)(uint32_t)key * p1;
        h = rotl64(h, 23) * p2 + p3;
    } else {
        for (int b = 0; b < KS; ++b) {
            h ^= ((key >> (8 * b)) & 0xffull) * p5;
            h = rotl64(h, 11) * p1;
        }
    }
    h ^= h >> 33;
    h *= p2;
    h ^= h >> 29;
    h *= p3;
    h ^= h >> 32;
    return h;
}

// The probe hash reduced to the 32-bit probe index type
// (BucketStorageRef<..., Extent = uint32_t>, include/detail/storages/bucket_storage_ref.h:21-29;
// SanitizeHash is a plain cast, include/utility/math_utils.h:19-29).
template <int KS>
MXCO_HD uint32_t probe_hash(uint32_t family, uint64_t key, uint32_t seed)
{
    switch (family) {
    case kXXHash32: return xxhash32<KS>(key, seed);
    case kXXHash64: return (uint32_t)xxhash64<KS>(key, seed);
    case kFmix32: return fmix32((uint32_t)key ^ seed);
    case kFmix64: return (uint32_t)fmix64(key ^ (uint64_t)seed);
    default: return murmur3_32<KS>(key, seed);
    }
}

}  // namespace mxco
