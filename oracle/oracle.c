/*
 * oracle.c -- CPU restatement of the ops-collections hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is built into
 * oracle/liboracle.so and loaded only by tests/, by __graft_entry__.smoke()
 * and by the cpu_baseline leg of bench.py.  Nothing in the product
 * (include/, ops-collections_amd/) links, loads or calls it, and the product
 * path fails loudly instead of falling back to it.
 *
 * Every function names the reference file:line whose behaviour it restates.
 * The restatement is written from the semantics, not translated: the
 * reference is AscendC SIMT code and cannot be compiled here.
 *
 * Parity pins (checked by tests/test_oracle.py):
 *   hashes      -- reference KATs tests/utility/hash_test.cpp:17-46 and
 *                  tests/utility/xxhash64_test.cpp:88-103, plus the
 *                  published MurmurHash3 / xxHash test vectors;
 *   primes      -- SHA-256 of the reference table include/utility/prime.h:15
 *                  (614c7e6d...8b43) and KATs tests/utility/extent_test.cpp;
 *   probing     -- closed form idx_k = (start + k*step) mod cap, full-cycle
 *                  coverage (tests/utility/probing_scheme_test.cpp:21-211);
 *   map / set   -- std::unordered_map first-wins goldens
 *                  (tests/common/golden.h:22-150) re-derived in Python;
 *   bloom       -- brute-force bit-by-bit restatement of the Parquet SBBF
 *                  spec and tests/common/bloom_filter_golden.h:67-146;
 *   hll         -- per-key register closed form
 *                  (tests/hyperloglog/add_test.cpp:321-335), linear-counting
 *                  and HLL closed forms for the estimate, and, for the
 *                  bias-corrected regime, a brute-force nearest-six mean over
 *                  the whole raw-estimate table (the method of
 *                  tests/hyperloglog/hll_finalizer_neighbor_count.cpp:36-55).
 */
#define _GNU_SOURCE
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <pthread.h>

#define ORC_API __attribute__((visibility("default")))

/* Planted-mistake builds for tests/test_oracle_mutations.py: compiling with
 * -DORC_PLANT=n (n >= 1) injects one plausible mistake, and that test checks
 * that the pinning tests above reject every such build.  The normal build
 * (ORC_PLANT = 0) contains none of them. */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif
#define PLANT(n) (ORC_PLANT == (n))

/* ------------------------------------------------------------------------ */
/* Hash functions.                                                          */
/* ------------------------------------------------------------------------ */

static inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
static inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

static inline uint32_t load32(const uint8_t* p) { uint32_t v; memcpy(&v, p, 4); return v; }
static inline uint64_t load64(const uint8_t* p) { uint64_t v; memcpy(&v, p, 8); return v; }

/* MurmurHash3 finalizers with a seed xor'ed into the key
 * (include/detail/hash_functions/utils.h:42-64). */
ORC_API uint32_t orc_fmix32(uint32_t key, uint32_t seed)
{
    uint32_t h = key ^ seed;
    h ^= h >> 16; h *= 0x85ebca6bu;
    h ^= h >> 13; h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

ORC_API uint64_t orc_fmix64(uint64_t key, uint64_t seed)
{
    uint64_t h = key ^ seed;
    h ^= h >> (PLANT(1) ? 32 : 33); h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

/* MurmurHash3_x86_32 over the object bytes of the key
 * (include/detail/hash_functions/murmurhash3.h:70-105). */
ORC_API uint32_t orc_murmur3_32(const void* data, size_t len, uint32_t seed)
{
    const uint8_t* p = (const uint8_t*)data;
    const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
    uint32_t h = seed;
    size_t nblocks = len / 4;
    for (size_t i = 0; i < nblocks; ++i) {
        uint32_t k = load32(p + 4 * i);
        k *= c1; k = rotl32(k, 15); k *= c2;
        h ^= k; h = rotl32(h, 13); h = h * 5 + 0xe6546b64u;
    }
    const uint8_t* tail = p + 4 * nblocks;
    uint32_t k = 0;
    switch (len & 3) {
    case 3: k ^= (uint32_t)tail[2] << 16; /* fall through */
    case 2: k ^= (uint32_t)tail[1] << 8;  /* fall through */
    case 1: k ^= (uint32_t)tail[0];
            k *= c1; k = rotl32(k, 15); k *= c2; h ^= k;
    }
    if (!PLANT(3)) h ^= (uint32_t)len;
    h ^= h >> 16; h *= 0x85ebca6bu;
    h ^= h >> 13; h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

/* XXH32 (include/detail/hash_functions/xxhash.h:53-175). */
#define XP1 0x9e3779b1u
#define XP2 0x85ebca77u
#define XP3 0xc2b2ae3du
#define XP4 0x27d4eb2fu
#define XP5 0x165667b1u
ORC_API uint32_t orc_xxh32(const void* data, size_t len, uint32_t seed)
{
    const uint8_t* p = (const uint8_t*)data;
    size_t off = 0;
    uint32_t h;
    if (len >= 16) {
        uint32_t v1 = seed + XP1 + XP2, v2 = seed + XP2, v3 = seed, v4 = seed - XP1;
        while (off + 16 <= len) {
            v1 = rotl32(v1 + load32(p + off) * XP2, PLANT(12) ? 12 : 13) * XP1;
            v2 = rotl32(v2 + load32(p + off + 4) * XP2, 13) * XP1;
            v3 = rotl32(v3 + load32(p + off + 8) * XP2, 13) * XP1;
            v4 = rotl32(v4 + load32(p + off + 12) * XP2, 13) * XP1;
            off += 16;
        }
        h = rotl32(v1, 1) + rotl32(v2, 7) + rotl32(v3, 12) + rotl32(v4, 18);
    } else {
        h = seed + XP5;
    }
    h += (uint32_t)len;
    while (off + 4 <= len) {
        h += load32(p + off) * XP3;
        h = rotl32(h, 17) * XP4;
        off += 4;
    }
    while (off < len) {
        h += (uint32_t)p[off] * XP5;
        h = rotl32(h, 11) * XP1;
        ++off;
    }
    h ^= h >> 15; h *= XP2;
    h ^= h >> 13; h *= XP3;
    h ^= h >> 16;
    return h;
}

/* XXH64 (include/detail/hash_functions/xxhash.h:185-326). */
#define YP1 11400714785074694791ull
#define YP2 14029467366897019727ull
#define YP3 1609587929392839161ull
#define YP4 9650029242287828579ull
#define YP5 2870177450012600261ull
static inline uint64_t xxh64_round(uint64_t acc, uint64_t in)
{
    acc += in * YP2;
    acc = rotl64(acc, 31);
    return acc * YP1;
}
static inline uint64_t xxh64_merge(uint64_t h, uint64_t v)
{
    h ^= xxh64_round(0, v);
    return h * YP1 + YP4;
}
ORC_API uint64_t orc_xxh64(const void* data, size_t len, uint64_t seed)
{
    const uint8_t* p = (const uint8_t*)data;
    size_t off = 0;
    uint64_t h;
    if (len >= 32) {
        uint64_t v1 = seed + YP1 + YP2, v2 = seed + YP2, v3 = seed, v4 = seed - YP1;
        while (off + 32 <= len) {
            v1 = xxh64_round(v1, load64(p + off));
            v2 = xxh64_round(v2, load64(p + off + 8));
            v3 = xxh64_round(v3, load64(p + off + 16));
            v4 = xxh64_round(v4, load64(p + off + 24));
            off += 32;
        }
        h = rotl64(v1, 1) + rotl64(v2, 7) + rotl64(v3, 12) + rotl64(v4, 18);
        h = xxh64_merge(h, v1); h = xxh64_merge(h, v2);
        h = xxh64_merge(h, v3); h = xxh64_merge(h, v4);
    } else {
        h = seed + YP5;
    }
    if (!PLANT(2)) h += (uint64_t)len;
    while (off + 8 <= len) {
        h ^= xxh64_round(0, load64(p + off));
        h = rotl64(h, 27) * YP1 + YP4;
        off += 8;
    }
    if (off + 4 <= len) {
        h ^= (uint64_t)load32(p + off) * YP1;
        h = rotl64(h, 23) * YP2 + YP3;
        off += 4;
    }
    while (off < len) {
        h ^= (uint64_t)p[off] * YP5;
        h = rotl64(h, 11) * YP1;
        ++off;
    }
    h ^= h >> 33; h *= YP2;
    h ^= h >> 29; h *= YP3;
    h ^= h >> 32;
    return h;
}

/* Hash-family identifiers shared with the product's C ABI (include/mxco.h). */
enum { H_MURMUR3_32 = 0, H_XXHASH_32 = 1, H_XXHASH_64 = 2, H_FMIX_32 = 3, H_FMIX_64 = 4 };

/* Hash of a key's object bytes, reduced to the 32-bit probe size type the
 * device kernels use: BucketStorageRef<..., Extent = uint32_t>
 * (include/detail/storages/bucket_storage_ref.h:21-29) and SanitizeHash,
 * a plain cast (include/utility/math_utils.h:19-29). */
static uint32_t probe_hash(int family, const uint8_t* key, uint32_t ks, uint32_t seed)
{
    switch (family) {
    case H_MURMUR3_32: return orc_murmur3_32(key, ks, seed);
    case H_XXHASH_32:  return orc_xxh32(key, ks, seed);
    case H_XXHASH_64:  return (uint32_t)orc_xxh64(key, ks, seed);
    case H_FMIX_32:    return orc_fmix32(load32(key), seed);
    case H_FMIX_64:    return (uint32_t)orc_fmix64(load64(key), seed);
    }
    return 0;
}

ORC_API uint32_t orc_probe_hash(int family, const void* key, uint32_t ks, uint32_t seed)
{
    return probe_hash(family, (const uint8_t*)key, ks, seed);
}

/* ------------------------------------------------------------------------ */
/* Capacity primes.                                                         */
/* ------------------------------------------------------------------------ */
/*
 * The reference ships a 140,741-entry table (include/utility/prime.h:15).
 * It is exactly the set {3, 5, 11, 151} united with the chain
 *     p0 = 2,  p_{k+1} = nextprime(p_k + (p_k < 2^17 ? 5 : 2^17))
 * stopped at 17,177,758,133; the tests compare the SHA-256 of the generated
 * table with the reference's.
 */
static uint64_t mulmod64(uint64_t a, uint64_t b, uint64_t m)
{
    return (uint64_t)(((unsigned __int128)a * b) % m);
}
static uint64_t powmod64(uint64_t a, uint64_t e, uint64_t m)
{
    uint64_t r = 1;
    a %= m;
    while (e) {
        if (e & 1) r = mulmod64(r, a, m);
        a = mulmod64(a, a, m);
        e >>= 1;
    }
    return r;
}
static int is_prime_u64(uint64_t n)
{
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; ++i) {
        if (n % small[i] == 0) return n == small[i];
    }
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) { d >>= 1; ++s; }
    for (int i = 0; i < 12; ++i) {          /* deterministic for n < 3.3e24 */
        uint64_t x = powmod64(small[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int r = 1; r < s; ++r) {
            x = mulmod64(x, x, n);
            if (x == n - 1) { comp = 0; break; }
        }
        if (comp) return 0;
    }
    return 1;
}
static uint64_t next_prime(uint64_t n)
{
    while (!is_prime_u64(n)) ++n;
    return n;
}

#define ORC_PRIME_MAX 17177758133ull
static int64_t* g_primes = NULL;
static size_t g_nprimes = 0;
static pthread_once_t g_primes_once = PTHREAD_ONCE_INIT;

static void build_primes(void)
{
    size_t cap = 150000;
    int64_t* t = (int64_t*)malloc(cap * sizeof(int64_t));
    size_t n = 0;
    uint64_t p = 2;
    /* merge the extra members {3, 5, 11, 151} into the ascending chain */
    static const uint64_t extra[] = {3, 5, 11, 151};
    size_t ei = 0;
    t[n++] = 2;
    for (;;) {
        uint64_t q = next_prime(p + (p < (1ull << 17) ? 5ull : (1ull << 17)));
        if (q > ORC_PRIME_MAX) break;
        while (ei < 4 && extra[ei] < q) t[n++] = (int64_t)extra[ei++];
        t[n++] = (int64_t)q;
        p = q;
    }
    g_primes = t;
    g_nprimes = n;
}

ORC_API size_t orc_primes(const int64_t** out)
{
    pthread_once(&g_primes_once, build_primes);
    if (out) *out = g_primes;
    return g_nprimes;
}

/* MakeValidExtent (include/detail/extent/extent.inl:21-63).
 * probing 0 = LinearProbing, 1 = DoubleHashing.  Returns 0 when the
 * reference throws std::invalid_argument. */
ORC_API uint64_t orc_valid_extent(uint64_t cap, uint32_t bucket, int probing)
{
    uint64_t c = cap ? cap : 1;
    uint64_t size = c / bucket + (c % bucket != 0);
    if (PLANT(9) && probing == 0) size = c / bucket;
    if (probing == 0) {
        return (size + (cap == 0)) * bucket;
    }
    if (size > ORC_PRIME_MAX) return 0;
    const int64_t* t;
    size_t n = orc_primes(&t);
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if ((uint64_t)t[mid] < size) lo = mid + 1; else hi = mid;
    }
    return (uint64_t)t[lo] * bucket;
}

/* ------------------------------------------------------------------------ */
/* Open-addressing table (StaticMap / StaticSet), sequential.               */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint8_t* slots;
    uint32_t cap;           /* device SizeType is uint32_t */
    uint32_t ks, vs, ss, voff;
    uint32_t bucket;
    int probing;            /* 0 linear, 1 double */
    int hash;               /* family, seed 0 (hash2 of DoubleHashing: seed 1) */
    int key_float;          /* KeyEqual is operator== on float/double */
    uint8_t empty_key[8], empty_val[8];
    uint64_t lines_read;    /* distinct 128-byte lines read by the last walk */
} orc_oa;

/* Pair<F,S> alignment: min(16, bit_ceil(sF + sS)) (include/detail/pair/pair.inl:20-25). */
static void pair_layout(uint32_t ks, uint32_t vs, uint32_t* ss, uint32_t* voff)
{
    if (vs == 0) { *ss = ks; *voff = 0; return; }
    uint32_t a = 1;
    while (a < ks + vs) a <<= 1;
    if (a > 16) a = 16;
    uint32_t off = ks;
    if (off % vs && !PLANT(10)) off += vs - off % vs;  /* natural alignment of second */
    uint32_t size = off + vs;
    if (size % a) size += a - size % a;
    *ss = size;
    *voff = off;
}

ORC_API void orc_pair_layout(uint32_t ks, uint32_t vs, uint32_t* ss, uint32_t* voff)
{
    pair_layout(ks, vs, ss, voff);
}

ORC_API orc_oa* orc_oa_create(uint32_t ks, uint32_t vs, uint32_t bucket, int probing, int hash, int key_float,
                              uint64_t requested, const void* empty_key, const void* empty_val)
{
    uint64_t cap = orc_valid_extent(requested, bucket, probing);
    if (cap == 0 || cap > 0xffffffffull) return NULL;
    orc_oa* t = (orc_oa*)calloc(1, sizeof(orc_oa));
    t->ks = ks; t->vs = vs; t->bucket = bucket; t->probing = probing; t->hash = hash;
    t->key_float = key_float;
    pair_layout(ks, vs, &t->ss, &t->voff);
    t->cap = (uint32_t)cap;
    memcpy(t->empty_key, empty_key, ks);
    if (vs) memcpy(t->empty_val, empty_val, vs);
    t->slots = (uint8_t*)calloc((size_t)cap, t->ss);
    for (uint64_t i = 0; i < cap; ++i) {
        memcpy(t->slots + i * t->ss, t->empty_key, ks);
        if (vs) memcpy(t->slots + i * t->ss + t->voff, t->empty_val, vs);
    }
    return t;
}

ORC_API void orc_oa_destroy(orc_oa* t)
{
    if (!t) return;
    free(t->slots);
    free(t);
}
ORC_API uint64_t orc_oa_capacity(const orc_oa* t) { return t->cap; }
ORC_API uint32_t orc_oa_slot_size(const orc_oa* t) { return t->ss; }
ORC_API uint8_t* orc_oa_slots(orc_oa* t) { return t->slots; }

/* Sentinel test is bitwise (include/detail/bitwise_compare.h:91-145). */
static inline int is_empty_key(const orc_oa* t, const uint8_t* k) { return memcmp(k, t->empty_key, t->ks) == 0; }

/* IEEE binary16 ==, by converting both to float (exact for every half). */
static float half_to_float(uint16_t h)
{
    uint32_t sign = (uint32_t)(h >> 15) << 31, exp = (h >> 10) & 31u, man = h & 1023u;
    float f;
    if (exp == 31) {
        uint32_t bits = sign | 0x7f800000u | (man << 13);
        memcpy(&f, &bits, 4);
        return f;
    }
    f = exp ? ldexpf((float)(man | 1024u), (int)exp - 25) : ldexpf((float)man, -24);
    return sign ? -f : f;
}

/* KeyEqual = EqualTo<Key>, operator== (include/utility/math_utils.h:67). */
static inline int key_equal(const orc_oa* t, const uint8_t* a, const uint8_t* b)
{
    if (t->key_float) {
        if (t->ks == 2) { uint16_t x, y; memcpy(&x, a, 2); memcpy(&y, b, 2); return half_to_float(x) == half_to_float(y); }
        if (t->ks == 4) { float x, y; memcpy(&x, a, 4); memcpy(&y, b, 4); return x == y; }
        double x, y; memcpy(&x, a, 8); memcpy(&y, b, 8); return x == y;
    }
    return memcmp(a, b, t->ks) == 0;
}

/* LinearProbing::MakeIterator / DoubleHashing::MakeIterator
 * (include/detail/probing_scheme/probing_scheme_impl.inl:118-128, :177-189),
 * all in uint32_t. */
typedef struct { uint32_t idx, step, cap; } probe_it;

static probe_it make_iter(const orc_oa* t, const uint8_t* key)
{
    probe_it it;
    uint32_t nb = t->cap / t->bucket;
    uint32_t h1 = probe_hash(t->hash, key, t->ks, 0);
    it.idx = ((PLANT(5) ? h1 + 1 : h1) % nb) * t->bucket;
    if (t->probing == 0) {
        it.step = t->bucket;
    } else {
        uint32_t h2 = probe_hash(t->hash, key, t->ks, 1);
        it.step = (PLANT(4) ? h2 % nb + 1 : h2 % (nb - 1) + 1) * t->bucket;
    }
    it.cap = t->cap;
    return it;
}
/* ProbingIterator::operator++ (probing_scheme_impl.inl:76-84). */
static inline void iter_next(probe_it* it)
{
    uint32_t rem = it->cap - it->step;
    it->idx = it->idx >= rem ? it->idx - rem : it->idx + it->step;
}

ORC_API void orc_oa_probe_start(const orc_oa* t, const void* key, uint32_t* start, uint32_t* step)
{
    probe_it it = make_iter(t, (const uint8_t*)key);
    *start = it.idx;
    *step = it.step;
}

enum { R_FAILED = 0, R_SUCCESS = 1, R_DUPLICATE = 2 };

/* Line accounting for the roofline: distinct 128-byte lines covering the
 * slots a walk reads, in walk order (a line counts once per key). */
typedef struct { uint64_t last[4]; int n; uint64_t count; } line_acc;
static inline void line_touch(line_acc* a, uint64_t slot, uint32_t ss)
{
    uint64_t l0 = (slot * ss) / 128, l1 = (slot * ss + ss - 1) / 128;
    for (uint64_t l = l0; l <= l1; ++l) {
        int seen = 0;
        for (int i = 0; i < a->n; ++i) if (a->last[i] == l) { seen = 1; break; }
        if (!seen) {
            a->count++;
            a->last[a->n % 4] = l;
            if (a->n < 4) a->n++; else a->n = 4;
        }
    }
}

/* InsertVerdict<assign> (include/detail/open_addressing/open_addressing_ref_impl.h:490-532),
 * sequential: a CAS on an AVAILABLE slot always succeeds. */
static int insert_verdict(orc_oa* t, const uint8_t* pair, int assign, line_acc* acc)
{
    const uint8_t* key = pair;
    if (is_empty_key(t, key)) return R_FAILED;
    probe_it it = make_iter(t, key);
    uint32_t init = it.idx;
    for (;;) {
        for (uint32_t i = 0; i < t->bucket; ++i) {
            uint64_t s = (uint64_t)it.idx + i;
            uint8_t* slot = t->slots + s * t->ss;
            if (acc) line_touch(acc, s, t->ss);
            if (is_empty_key(t, slot)) {
                /* AVAILABLE: CasDependentWrite / PackCas succeed sequentially */
                memcpy(slot, key, t->ks);
                if (t->vs) memcpy(slot + t->voff, pair + t->voff, t->vs);
                return R_SUCCESS;
            }
            if (key_equal(t, key, slot)) {
                if ((assign || PLANT(6)) && t->vs) memcpy(slot + t->voff, pair + t->voff, t->vs);
                return R_DUPLICATE;
            }
        }
        iter_next(&it);
        if (it.idx == init) return R_FAILED;
    }
}

/* OpenAddressingImpl::Insert / InsertOrAssign counting
 * (include/detail/open_addressing/open_addressing_impl.h:102-201).
 * `pairs` uses the slot layout (stride ss).  Returns failures. */
ORC_API uint64_t orc_oa_insert(orc_oa* t, const void* pairs, uint64_t n, int assign, uint64_t* assigned,
                               uint64_t* lines)
{
    const uint8_t* p = (const uint8_t*)pairs;
    uint64_t failed = 0, asg = 0;
    line_acc acc = {{0}, 0, 0};
    for (uint64_t i = 0; i < n; ++i) {
        acc.n = 0;
        int r = insert_verdict(t, p + i * t->ss, assign, lines ? &acc : NULL);
        if (assign) {
            if (r == R_FAILED) failed++;
            else if (r == R_DUPLICATE) asg++;
        } else if (r != R_SUCCESS) {
            failed++;
        }
    }
    if (assigned) *assigned = asg;
    if (lines) *lines = acc.count;
    return failed;
}

/* InsertIf with a per-element predicate flag (kernels.h:47-83). */
ORC_API uint64_t orc_oa_insert_if(orc_oa* t, const void* pairs, const uint8_t* flags, uint64_t n)
{
    const uint8_t* p = (const uint8_t*)pairs;
    uint64_t failed = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (!flags[i]) continue;
        if (insert_verdict(t, p + i * t->ss, 0, NULL) != R_SUCCESS) failed++;
    }
    return failed;
}

/* AssignIf (kernels.h:183-210 with Assign, open_addressing_ref_impl.h:384-410):
 * for each selected pair, walk the key's probe sequence; an EMPTY slot ends
 * the walk with nothing written, an EQUAL slot gets the pair's payload, a
 * full wrap ends it.  Nothing is inserted.  Returns the pairs assigned. */
ORC_API uint64_t orc_oa_assign_if(orc_oa* t, const void* pairs, const uint8_t* flags, uint64_t n)
{
    const uint8_t* p = (const uint8_t*)pairs;
    uint64_t assigned = 0;
    if (!t->vs) return 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (flags && !flags[i]) continue;
        const uint8_t* pair = p + i * t->ss;
        probe_it it = make_iter(t, pair);
        uint32_t init = it.idx;
        int done = 0;
        while (!done) {
            for (uint32_t b = 0; b < t->bucket && !done; ++b) {
                uint8_t* slot = t->slots + ((uint64_t)it.idx + b) * t->ss;
                if (is_empty_key(t, slot) && !PLANT(17)) {
                    if (PLANT(16)) memcpy(slot, pair, t->ss);
                    done = 1;
                } else if (key_equal(t, pair, slot)) {
                    memcpy(slot + t->voff, pair + t->voff, t->vs);
                    assigned++;
                    done = 1;
                }
            }
            if (done) break;
            iter_next(&it);
            if (it.idx == init) break;
        }
    }
    return assigned;
}

/* InsertAndFind (open_addressing_ref_impl.h:302-330, :544-565). */
ORC_API void orc_oa_insert_and_find(orc_oa* t, const void* pairs, uint64_t n, void* out_found, uint8_t* out_inserted)
{
    const uint8_t* p = (const uint8_t*)pairs;
    uint8_t* of = (uint8_t*)out_found;
    uint32_t os = t->vs ? t->vs : t->ks;
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t* pair = p + i * t->ss;
        const uint8_t* empty_out = t->vs ? t->empty_val : t->empty_key;
        if (is_empty_key(t, pair)) {
            memcpy(of + i * os, empty_out, os);
            out_inserted[i] = 0;
            continue;
        }
        probe_it it = make_iter(t, pair);
        uint32_t init = it.idx;
        int done = 0;
        while (!done) {
            for (uint32_t b = 0; b < t->bucket && !done; ++b) {
                uint8_t* slot = t->slots + ((uint64_t)it.idx + b) * t->ss;
                if (is_empty_key(t, slot)) {
                    memcpy(slot, pair, t->ks);
                    if (t->vs) memcpy(slot + t->voff, pair + t->voff, t->vs);
                    memcpy(of + i * os, t->vs ? pair + t->voff : pair, os);
                    out_inserted[i] = 1;
                    done = 1;
                } else if (key_equal(t, pair, slot)) {
                    memcpy(of + i * os, t->vs ? slot + t->voff : slot, os);
                    out_inserted[i] = 0;
                    done = 1;
                }
            }
            if (done) break;
            iter_next(&it);
            if (it.idx == init) {
                memcpy(of + i * os, empty_out, os);
                out_inserted[i] = 0;
                break;
            }
        }
    }
}

/* FindAndContains (open_addressing_ref_impl.h:568-605).  Writes the found
 * payload (map) or stored key (set) into out_vals and the flag into
 * out_contains; either may be NULL.  Returns the distinct 128-byte lines the
 * walks read, summed over the batch. */
ORC_API uint64_t orc_oa_find(const orc_oa* t, const void* keys, uint64_t n, void* out_vals, uint8_t* out_contains)
{
    const uint8_t* k = (const uint8_t*)keys;
    uint8_t* ov = (uint8_t*)out_vals;
    uint32_t os = t->vs ? t->vs : t->ks;
    line_acc acc = {{0}, 0, 0};
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t* key = k + i * t->ks;
        const uint8_t* res = t->vs ? t->empty_val : t->empty_key;
        int hit = 0;
        acc.n = 0;
        probe_it it = make_iter(t, key);
        uint32_t init = it.idx;
        for (;;) {
            int stop = 0;
            for (uint32_t b = 0; b < t->bucket; ++b) {
                uint64_t s = (uint64_t)it.idx + b;
                const uint8_t* slot = t->slots + s * t->ss;
                line_touch(&acc, s, t->ss);
                if (is_empty_key(t, slot)) { stop = 1; break; }
                if (key_equal(t, key, slot)) {
                    res = t->vs ? slot + t->voff : slot;
                    hit = 1; stop = 1; break;
                }
            }
            if (stop) break;
            iter_next(&it);
            if (it.idx == init) break;
        }
        if (ov) memcpy(ov + i * os, res, os);
        if (out_contains) out_contains[i] = (uint8_t)hit;
    }
    return acc.count;
}

/* Erase with optional tombstone (open_addressing_ref_impl.h:341-379).
 * erased_key == NULL means no tombstone (the slot becomes empty).
 * Returns the number of keys not erased (kernels.h Erase counting). */
ORC_API uint64_t orc_oa_erase(orc_oa* t, const void* keys, uint64_t n, const void* erased_key)
{
    const uint8_t* k = (const uint8_t*)keys;
    uint64_t failed = 0;
    int tomb = erased_key && memcmp(erased_key, t->empty_key, t->ks) != 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t* key = k + i * t->ks;
        probe_it it = make_iter(t, key);
        uint32_t init = it.idx;
        int erased = 0, stop = 0;
        for (;;) {
            for (uint32_t b = 0; b < t->bucket; ++b) {
                uint8_t* slot = t->slots + ((uint64_t)it.idx + b) * t->ss;
                int empty = is_empty_key(t, slot);
                if (!empty && key_equal(t, key, slot)) {
                    memcpy(slot, tomb ? (const uint8_t*)erased_key : t->empty_key, t->ks);
                    if (t->vs) memcpy(slot + t->voff, t->empty_val, t->vs);
                    erased = 1; stop = 1; break;
                }
                if (tomb && empty) { stop = 1; break; }
            }
            if (stop) break;
            iter_next(&it);
            if (it.idx == init) break;
        }
        if (!erased) failed++;
    }
    return failed;
}

/* ------------------------------------------------------------------------ */
/* Multithreaded restatement for the CPU baseline (same algorithm, same     */
/* slot layout, __atomic CAS).  i64->i64 (16-byte slot, key CAS then value  */
/* store) and i32->i32 (8-byte slot, whole-slot CAS).                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    orc_oa* t;
    const uint8_t* in;
    uint8_t* out;
    uint8_t* contains;
    uint64_t begin, end;
    uint64_t failed;
    int op;   /* 0 insert, 1 find */
} mt_job;

static void* mt_worker(void* arg)
{
    mt_job* j = (mt_job*)arg;
    orc_oa* t = j->t;
    uint64_t failed = 0;
    uint32_t nb = t->cap / t->bucket;
    for (uint64_t i = j->begin; i < j->end; ++i) {
        if (j->op == 0) {
            const uint8_t* pair = j->in + i * t->ss;
            if (is_empty_key(t, pair)) { failed++; continue; }
            uint32_t idx = (probe_hash(t->hash, pair, t->ks, 0) % nb) * t->bucket, init = idx;
            int res = R_FAILED, done = 0;
            while (!done) {
                for (uint32_t b = 0; b < t->bucket; ++b) {
                    uint8_t* slot = t->slots + ((uint64_t)idx + b) * t->ss;
                    if (t->ss == 16 && t->ks == 8) {
                        uint64_t cur = __atomic_load_n((uint64_t*)slot, __ATOMIC_RELAXED);
                        uint64_t ek = load64(t->empty_key), key = load64(pair);
                        if (cur == ek) {
                            uint64_t exp = ek;
                            if (__atomic_compare_exchange_n((uint64_t*)slot, &exp, key, 0, __ATOMIC_RELAXED,
                                                            __ATOMIC_RELAXED)) {
                                memcpy(slot + 8, pair + 8, 8);
                                res = R_SUCCESS; done = 1; break;
                            }
                            cur = exp;
                        }
                        if (cur == key) { res = R_DUPLICATE; done = 1; break; }
                    } else {
                        uint64_t cur = __atomic_load_n((uint64_t*)slot, __ATOMIC_RELAXED);
                        uint32_t ek = load32(t->empty_key), key = load32(pair);
                        if ((uint32_t)cur == ek) {
                            uint64_t exp = (uint64_t)ek | ((uint64_t)load32(t->empty_val) << 32);
                            if (__atomic_compare_exchange_n((uint64_t*)slot, &exp, load64(pair), 0, __ATOMIC_RELAXED,
                                                            __ATOMIC_RELAXED)) {
                                res = R_SUCCESS; done = 1; break;
                            }
                            cur = exp;
                        }
                        if ((uint32_t)cur == key) { res = R_DUPLICATE; done = 1; break; }
                    }
                }
                if (done) break;
                uint32_t rem = t->cap - t->bucket;
                idx = idx >= rem ? idx - rem : idx + t->bucket;
                if (idx == init) break;
            }
            if (res != R_SUCCESS) failed++;
        } else {
            const uint8_t* key = j->in + i * t->ks;
            uint32_t idx = (probe_hash(t->hash, key, t->ks, 0) % nb) * t->bucket, init = idx;
            const uint8_t* res = t->empty_val;
            int hit = 0, stop = 0;
            for (;;) {
                for (uint32_t b = 0; b < t->bucket; ++b) {
                    const uint8_t* slot = t->slots + ((uint64_t)idx + b) * t->ss;
                    if (is_empty_key(t, slot)) { stop = 1; break; }
                    if (memcmp(slot, key, t->ks) == 0) { res = slot + t->voff; hit = 1; stop = 1; break; }
                }
                if (stop) break;
                uint32_t rem = t->cap - t->bucket;
                idx = idx >= rem ? idx - rem : idx + t->bucket;
                if (idx == init) break;
            }
            memcpy(j->out + i * t->vs, res, t->vs);
            if (j->contains) j->contains[i] = (uint8_t)hit;
        }
    }
    j->failed = failed;
    return NULL;
}

/* Linear-probing maps only (the baseline workloads). op 0 = Insert (returns
 * failures), op 1 = Find. */
ORC_API uint64_t orc_oa_mt(orc_oa* t, int op, const void* in, uint64_t n, void* out, uint8_t* contains, int threads)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    mt_job jobs[256];
    uint64_t per = (n + threads - 1) / threads;
    for (int i = 0; i < threads; ++i) {
        jobs[i].t = t; jobs[i].in = (const uint8_t*)in; jobs[i].out = (uint8_t*)out; jobs[i].contains = contains;
        jobs[i].begin = per * i < n ? per * i : n;
        jobs[i].end = per * (i + 1) < n ? per * (i + 1) : n;
        jobs[i].op = op; jobs[i].failed = 0;
        pthread_create(&th[i], NULL, mt_worker, &jobs[i]);
    }
    uint64_t failed = 0;
    for (int i = 0; i < threads; ++i) {
        pthread_join(th[i], NULL);
        failed += jobs[i].failed;
    }
    return failed;
}

/* ------------------------------------------------------------------------ */
/* Split-block Bloom filter (BloomFilterPolicy defaults: 8 x u32 words per  */
/* 256-bit block, one bit per word, xxhash_64 seed 0).                      */
/* ------------------------------------------------------------------------ */

/* BloomFilterPolicy's 64 pattern salts (include/bloom_filter_policy.h:170-178;
 * the first eight are the Parquet SBBF salts).  Data: the filter bits
 * depend on every value; tests check the table by its SHA-256. */
static const uint32_t k_salts[64] = {
    0x47b6137bu, 0x44974d91u, 0x8824ad5bu, 0xa2b7289du, 0x705495c7u, 0x2df1424bu, 0x9efc4947u, 0x5c6bfb31u,
    0xb24bcdffu, 0xb6843d6du, 0x6db04543u, 0x3a12efddu, 0xb0ddd463u, 0x8d22f6e7u, 0xb82f1e53u, 0x7db9f86bu,
    0xc7afe639u, 0xfb135cd7u, 0x693256e1u, 0x9466d871u, 0x23d3d02fu, 0x6461d049u, 0x66a91621u, 0xbaa3006fu,
    0x52fb8d99u, 0x3ea88b4fu, 0x0f470cfdu, 0xb1db79a5u, 0x9809fcd1u, 0xbced4445u, 0x2eb7c737u, 0x2cea6803u,
    0x156f1955u, 0x8813c027u, 0xa26819f9u, 0x4c3b57bdu, 0x7df94487u, 0xb975e769u, 0xb8f20cb5u, 0x5c9e2e77u,
    0x5fb1735fu, 0x3a6f759bu, 0x3c090923u, 0xfced424du, 0xa187a6a9u, 0x6f070a41u, 0x2c85233bu, 0x7e62258bu,
    0x2771ef17u, 0x13bbf093u, 0x4ff059e5u, 0xe3ce3d0fu, 0xf1b4789fu, 0x9fbb6173u, 0x6a320cf5u, 0x1be2c481u,
    0x7ba8222bu, 0x6fd619b3u, 0x7b1bbf0du, 0x8b8993adu, 0x448eca95u, 0x82ab09d9u, 0x2ce53909u, 0x4f548685u};

ORC_API const uint32_t* orc_bloom_salts(void) { return k_salts; }

/* BlockIndex = (hi32 * numBlocks) >> 32 (bloom_filter_policy.h:102-107);
 * word w's pattern = OR over b < P/W of 1 << ((salt[w*(P/W) + b] * lo32) >> 27)
 * (WordPatternImpl, :157-168). */
static inline uint32_t bloom_block(uint64_t h, uint64_t num_blocks)
{
    return (uint32_t)(((uint64_t)(uint32_t)(h >> 32) * (uint32_t)num_blocks) >> 32);
}

static inline uint32_t bloom_pattern(uint32_t lo, uint32_t w, uint32_t bpw)
{
    uint32_t m = 0;
    for (uint32_t b = 0; b < bpw; ++b) {
        uint32_t salt = w * bpw + b;
        if (PLANT(7)) salt ^= 1;
        m |= 1u << ((k_salts[salt] * lo) >> 27);
    }
    return m;
}

/* Add / Contains under BloomFilterPolicy<Key, xxhash_64, uint32_t, W, P>. */
ORC_API void orc_bloom_add_policy(uint32_t* words, uint64_t num_blocks, uint32_t wpb, uint32_t pattern_bits,
                                  const void* keys, uint32_t ks, uint64_t n, uint64_t seed)
{
    const uint8_t* k = (const uint8_t*)keys;
    const uint32_t bpw = pattern_bits / wpb;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t h = orc_xxh64(k + i * ks, ks, seed);
        uint64_t base = (uint64_t)bloom_block(h, num_blocks) * wpb;
        for (uint32_t w = 0; w < wpb; ++w) words[base + w] |= bloom_pattern((uint32_t)h, w, bpw);
    }
}

ORC_API void orc_bloom_contains_policy(const uint32_t* words, uint64_t num_blocks, uint32_t wpb,
                                       uint32_t pattern_bits, const void* keys, uint32_t ks, uint64_t n,
                                       uint64_t seed, uint8_t* out)
{
    const uint8_t* k = (const uint8_t*)keys;
    const uint32_t bpw = pattern_bits / wpb;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t h = orc_xxh64(k + i * ks, ks, seed);
        uint64_t base = (uint64_t)bloom_block(h, num_blocks) * wpb;
        int ok = 1;
        for (uint32_t w = 0; w < wpb; ++w) {
            uint32_t m = bloom_pattern((uint32_t)h, w, bpw);
            ok &= (words[base + w] & m) == m;
        }
        out[i] = (uint8_t)ok;
    }
}

/* The default policy: 8 words, 8 pattern bits. */
ORC_API void orc_bloom_add(uint32_t* words, uint64_t num_blocks, const void* keys, uint32_t ks, uint64_t n,
                           uint64_t seed)
{
    orc_bloom_add_policy(words, num_blocks, 8, 8, keys, ks, n, seed);
}

ORC_API void orc_bloom_contains(const uint32_t* words, uint64_t num_blocks, const void* keys, uint32_t ks,
                                uint64_t n, uint64_t seed, uint8_t* out)
{
    orc_bloom_contains_policy(words, num_blocks, 8, 8, keys, ks, n, seed, out);
}

/* Multithreaded Add / Contains for the CPU baseline: contiguous key ranges
 * per thread, Add by __atomic OR on the words (only where bits are missing). */
typedef struct {
    uint32_t* words; uint64_t num_blocks; uint32_t wpb, bpw; const uint8_t* keys; uint32_t ks;
    uint64_t begin, end, seed; uint8_t* out;
} bloom_job;

static void* bloom_worker(void* arg)
{
    bloom_job* j = (bloom_job*)arg;
    for (uint64_t i = j->begin; i < j->end; ++i) {
        uint64_t h = orc_xxh64(j->keys + i * j->ks, j->ks, j->seed);
        uint32_t* b = j->words + (uint64_t)bloom_block(h, j->num_blocks) * j->wpb;
        int ok = 1;
        for (uint32_t w = 0; w < j->wpb; ++w) {
            uint32_t m = bloom_pattern((uint32_t)h, w, j->bpw);
            uint32_t cur = __atomic_load_n(b + w, __ATOMIC_RELAXED);
            if (j->out) ok &= (cur & m) == m;
            else if ((cur & m) != m) __atomic_fetch_or(b + w, m, __ATOMIC_RELAXED);
        }
        if (j->out) j->out[i] = (uint8_t)ok;
    }
    return NULL;
}

/* out == NULL: Add; else Contains into out. */
ORC_API void orc_bloom_mt(uint32_t* words, uint64_t num_blocks, uint32_t wpb, uint32_t pattern_bits,
                          const void* keys, uint32_t ks, uint64_t n, uint64_t seed, uint8_t* out, int threads)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    bloom_job jobs[256];
    uint64_t per = (n + threads - 1) / threads;
    for (int i = 0; i < threads; ++i) {
        bloom_job jb = {words, num_blocks, wpb, pattern_bits / wpb, (const uint8_t*)keys, ks,
                        per * i < n ? per * i : n, per * (i + 1) < n ? per * (i + 1) : n, seed, out};
        jobs[i] = jb;
        pthread_create(&th[i], NULL, bloom_worker, &jobs[i]);
    }
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
}

/* ------------------------------------------------------------------------ */
/* HyperLogLog (hyperloglog_ref.h:30-47, hyperloglog.inl:328-345,           */
/* finalizer.h:46-151).                                                     */
/* ------------------------------------------------------------------------ */

ORC_API void orc_hll_add(uint8_t* regs, uint32_t p, const void* keys, uint32_t ks, uint64_t n)
{
    const uint8_t* k = (const uint8_t*)keys;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t h = orc_xxh64(k + i * ks, ks, 0);
        uint32_t idx = (uint32_t)(h >> (64 - p));
        uint64_t tail = (h << p) | (1ull << (p - 1));
        uint8_t rank = (uint8_t)(__builtin_clzll(tail) + (PLANT(8) ? 0 : 1));
        if (regs[idx] < rank) regs[idx] = rank;
    }
}

/* Multithreaded Add for the CPU baseline: private registers per thread,
 * merged by register-wise maximum. */
typedef struct { uint8_t* regs; uint32_t p; const uint8_t* keys; uint32_t ks; uint64_t begin, end; } hll_job;

static void* hll_worker(void* arg)
{
    hll_job* j = (hll_job*)arg;
    orc_hll_add(j->regs, j->p, j->keys + j->begin * j->ks, j->ks, j->end - j->begin);
    return NULL;
}

ORC_API void orc_hll_add_mt(uint8_t* regs, uint32_t p, const void* keys, uint32_t ks, uint64_t n, int threads)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    const uint64_t m = 1ull << p;
    uint8_t* priv = (uint8_t*)calloc((size_t)threads, m);
    pthread_t th[256];
    hll_job jobs[256];
    uint64_t per = (n + threads - 1) / threads;
    for (int i = 0; i < threads; ++i) {
        hll_job jb = {priv + (uint64_t)i * m, p, (const uint8_t*)keys, ks, per * i < n ? per * i : n,
                      per * (i + 1) < n ? per * (i + 1) : n};
        jobs[i] = jb;
        pthread_create(&th[i], NULL, hll_worker, &jobs[i]);
    }
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    for (int i = 0; i < threads; ++i)
        for (uint64_t r = 0; r < m; ++r)
            if (regs[r] < priv[(uint64_t)i * m + r]) regs[r] = priv[(uint64_t)i * m + r];
    free(priv);
}

/* Rank histogram over the registers, 65 bins (kernels.h:74-92). */
ORC_API void orc_hll_histogram(const uint8_t* regs, uint32_t p, uint32_t* hist65)
{
    memset(hist65, 0, 65 * sizeof(uint32_t));
    for (uint64_t i = 0; i < (1ull << p); ++i) hist65[regs[i]]++;
}

/* HLL++ finalizer.  raw/bias are the precision's interpolation tables and
 * threshold its linear-counting threshold. */
static double hll_alpha_mm(double m)
{
    if (m == 16) return 0.673 * m * m;
    if (m == 32) return 0.697 * m * m;
    if (m == 64) return 0.709 * m * m;
    return (0.7213 / (1.0 + 1.079 / m)) * m * m;
}

static double hll_bias(double e, const double* raw, const double* bias, int n)
{
    const int k = PLANT(13) ? 5 : 6;
    int left = 0, right = n - 1, anchor = -1, found = 0;
    while (left <= right) {
        int mid = left + (right - left) / 2;
        if (raw[mid] < e) left = mid + 1;
        else if (raw[mid] > e) right = mid - 1;
        else { anchor = mid; found = 1; break; }
    }
    if (!found) anchor = left;
    int low = anchor - k + 1;
    if (low > n - k) low = n - k;
    if (low < 0) low = 0;
    int high = low + k < n ? low + k : n;
    while (high < n && (e - raw[high]) * (e - raw[high]) < (e - raw[low]) * (e - raw[low])) {
        low++; high++;
    }
    double s = 0.0;
    for (int i = low; i < high; ++i) s += bias[i];
    return s / (high - low);
}

/* HLL++ bias correction below 5m (finalizer.h:95-98). */
static double hll_corrected(double e, double m, const double* raw, const double* bias, int n)
{
    if (e >= 5.0 * m) return e;
    return PLANT(14) ? e + hll_bias(e, raw, bias, n) : e - hll_bias(e, raw, bias, n);
}

/* finalizer.h:46-74: linear counting up to 2.5m; above it the linear-counting
 * value while it stays at or below the precision's threshold, else the
 * bias-corrected raw estimate. */
ORC_API uint64_t orc_hll_finalize(uint32_t p, double z, int v, const double* raw, const double* bias, int n,
                                  double threshold)
{
    if (z == 0.0) return UINT64_MAX;
    double m = (double)(1u << p);
    double e = hll_alpha_mm(m) / z;
    if (v > 0) {
        double h = m * log(m / (PLANT(11) ? v + 1 : v));
        if (e <= 2.5 * m) return (uint64_t)round(h);
        e = (h <= threshold && !PLANT(15)) ? h : hll_corrected(e, m, raw, bias, n);
    } else {
        e = hll_corrected(e, m, raw, bias, n);
    }
    if (e >= 18446744073709551616.0) return UINT64_MAX;
    return (uint64_t)round(e);
}

/* z = sum_r hist[r] * 2^-r in ascending r (hyperloglog.inl:340-343). */
ORC_API double orc_hll_z(const uint32_t* hist65)
{
    double z = 0;
    for (int r = 0; r < 65; ++r) z += ldexp((double)hist65[r], -r);
    return z;
}
