// Open-addressing bulk kernels (StaticMap / StaticSet) for gfx950.
//
// One kernel template serves every bulk operation.  A key is owned by a
// cooperative group of G lanes (G = 1, 2, 4, 8) of a 64-lane wavefront.
// The group walks the reference's probe sequence one *window* at a time:
//   - LinearProbing: the walk is a contiguous slot scan from the start
//     bucket, wrapping at the capacity (probing_scheme_impl.inl:76-84,
//     118-128), so a window is the rest of the current 128-byte line;
//   - DoubleHashing: a window is (a piece of) the current bucket
//     (probing_scheme_impl.inl:177-189).
// The G lanes load the window's slots with independent vector loads (one
// 128-byte line per round trip at most), classify every slot, and take the
// *first* slot in walk order that ends the scan with a wave ballot -- the
// same slot the reference's sequential per-slot loop stops at
// (open_addressing_ref_impl.h:490-605).  A failed CAS resumes the walk at
// the next slot, as the reference's loop does.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "mxco/detail/hash.h"

namespace mxco {

enum OaOp : int {
    OP_FIND = 0,
    OP_CONTAINS = 1,
    OP_COUNT = 2,
    OP_INSERT = 3,
    OP_INSERT_OR_ASSIGN = 4,
    OP_ASSIGN = 5,
    OP_INSERT_AND_FIND = 6,
    OP_ERASE = 7,
    OP_PROBE_STATS = 8,  // Find's walk, counting 128-byte lines and slots read
    OP_INSERT_STATS = 9, // Insert's CAS walk, counting compare-and-swaps lost to another key and made
};

struct OaArgs {
    uint8_t* table;
    const uint8_t* in;       // slot-layout pairs (insert family) or keys
    const uint8_t* stencil;  // nullptr = AlwaysTrue
    uint8_t* out;            // Find values / Contains flags / InsertAndFind values
    uint8_t* out2;           // InsertAndFind inserted flags
    unsigned long long* counters;  // [0] failed / count, [1] assigned
    uint64_t n;
    uint64_t empty_key;
    uint64_t empty_val;
    uint64_t erased_key;
    uint32_t cap;
    uint32_t bucket;
    uint32_t nbuckets;
    uint32_t probing;  // 0 linear, 1 double
    uint32_t hash;
    uint32_t key_float;
    uint32_t stencil_size;
    uint32_t stencil_pred;
    uint32_t has_tombstone;
    uint32_t store_mode;  // payload store after a key CAS (see store_payload)
    uint32_t probe_flags;  // bit 0: first linear window ends at a 64-byte sector; bit 1: non-temporal probe loads
    uint32_t pad;
};

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

// Slot layouts: Pair<K,V> with alignas(min(16, bit_ceil(sK+sV)))
// (include/pair.h:41, detail/pair/pair.inl:20-25); sets store the key only.
// Keys and payloads are 1, 2, 4 or 8 bytes; the payload sits at its natural
// alignment after the key (offset max(sK, sV)), and the slot is 1 to 16
// bytes.  Bytes that are neither key nor payload are padding.
template <int N>
struct RawOf;
template <>
struct RawOf<1> {
    using T = uint8_t;
};
template <>
struct RawOf<2> {
    using T = uint16_t;
};
template <>
struct RawOf<4> {
    using T = uint32_t;
};
template <>
struct RawOf<8> {
    using T = uint64_t;
};
template <>
struct RawOf<16> {
    using T = u64x2;
};

constexpr int pow2_at_least(int n)
{
    int x = 1;
    while (x < n) x <<= 1;
    return x;
}

constexpr uint64_t low_bytes_mask(int bytes) { return bytes >= 8 ? ~0ull : ((1ull << (8 * bytes)) - 1ull); }

template <int KS, int VS>
struct Lay {
    static_assert(KS == 1 || KS == 2 || KS == 4 || KS == 8, "key width");
    static_assert(VS == 0 || VS == 1 || VS == 2 || VS == 4 || VS == 8, "payload width");
    static constexpr int SS = VS ? (pow2_at_least(KS + VS) < 16 ? pow2_at_least(KS + VS) : 16) : KS;
    static constexpr int VOFF = VS ? (KS > VS ? KS : VS) : 0;  // 8 in every 16-byte slot
    using Raw = typename RawOf<SS>::T;
    static constexpr uint64_t KMASK = low_bytes_mask(KS);
    static constexpr uint64_t VMASK = low_bytes_mask(VS);
    // padding bits of a slot of at most 8 bytes
    static constexpr uint64_t PADMASK =
        SS <= 8 ? (low_bytes_mask(SS) & ~KMASK & ~(VS ? (VMASK << (8 * VOFF)) : 0ull)) : 0ull;
    static __device__ __forceinline__ uint64_t key(Raw r)
    {
        if constexpr (SS == 16) return r.x & KMASK;
        else return (uint64_t)r & KMASK;
    }
    static __device__ __forceinline__ uint64_t val(Raw r)
    {
        if constexpr (VS == 0) return key(r);
        else if constexpr (SS == 16) return r.y & VMASK;
        else return ((uint64_t)r >> (8 * VOFF)) & VMASK;
    }
    // A slot of at most 8 bytes holding k and v over the padding of `pad`.
    static __device__ __forceinline__ uint64_t pack(uint64_t pad, uint64_t k, uint64_t v)
    {
        uint64_t r = (pad & PADMASK) | (k & KMASK);
        if constexpr (VS != 0) r |= (v & VMASK) << (8 * VOFF);
        return r;
    }
};

// Slot layouts whose bulk Find / Contains use the staged kernels: 4- and
// 8-byte keys and payloads.  The others (a 1- or 2-byte key or payload) run
// every operation on the plain kernel.
template <int KS, int VS>
constexpr bool kWideLayout = KS >= 4 && (VS == 0 || VS >= 4);

template <int N>
__device__ __forceinline__ uint64_t load_bits(const uint8_t* p)
{
    if constexpr (N == 1) return *p;
    else if constexpr (N == 2) return *(const uint16_t*)p;
    else if constexpr (N == 4) return *(const uint32_t*)p;
    else return *(const uint64_t*)p;
}

template <int N>
__device__ __forceinline__ void store_bits(uint8_t* p, uint64_t v)
{
    if constexpr (N == 1) *p = (uint8_t)v;
    else if constexpr (N == 2) *(uint16_t*)p = (uint16_t)v;
    else if constexpr (N == 4) *(uint32_t*)p = (uint32_t)v;
    else *(uint64_t*)p = v;
}

// Compare-and-swap of the N = 1 or 2 bytes at p, on the aligned 32-bit word
// that holds them: the other bytes of the word (a neighbouring slot, or
// padding) are kept, and a CAS that fails because only they changed is
// retried (the reference's AtomicCas2Byte, include/utility/atomic_cas_wrap.h:20-56).
// Returns the N bytes found at p (== expected on success).
template <int N>
__device__ __forceinline__ uint64_t cas_narrow(uint8_t* p, uint64_t expected, uint64_t desired)
{
    const uintptr_t addr = (uintptr_t)p;
    unsigned int* word = (unsigned int*)(addr & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)(addr & 3u) * 8u;
    const uint32_t m = (uint32_t)low_bytes_mask(N) << sh;
    uint32_t cur = *word;
    for (;;) {
        const uint32_t mine = (cur & m) >> sh;
        if (mine != (uint32_t)expected) return mine;
        const uint32_t want = (cur & ~m) | (((uint32_t)desired << sh) & m);
        const uint32_t prev = atomicCAS(word, cur, want);
        if (prev == cur) return expected;
        cur = prev;  // a neighbouring byte changed: try again on the new word
    }
}

template <int N>
__device__ __forceinline__ uint64_t cas_bits(uint8_t* p, uint64_t expected, uint64_t desired)
{
    if constexpr (N <= 2) return cas_narrow<N>(p, expected, desired);
    else if constexpr (N == 4) return atomicCAS((unsigned int*)p, (unsigned int)expected, (unsigned int)desired);
    else return atomicCAS((unsigned long long*)p, (unsigned long long)expected, (unsigned long long)desired);
}

template <int N>
__device__ __forceinline__ uint64_t atomic_load_bits(const uint8_t* p)
{
    if constexpr (N == 1)
        return __hip_atomic_load((const uint8_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else if constexpr (N == 2)
        return __hip_atomic_load((const uint16_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else if constexpr (N == 4)
        return __hip_atomic_load((const unsigned int*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        return __hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Agent-scope store: visible to agent-scope loads on every XCD during the
// kernel (a plain store may sit in the writer's XCD-local L2 until the
// kernel ends).  Used where a reader in the same launch waits for the value.
template <int N>
__device__ __forceinline__ void publish_bits(uint8_t* p, uint64_t v)
{
    if constexpr (N == 1)
        __hip_atomic_store((uint8_t*)p, (uint8_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else if constexpr (N == 2)
        __hip_atomic_store((uint16_t*)p, (uint16_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else if constexpr (N == 4)
        __hip_atomic_store((unsigned int*)p, (unsigned int)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        __hip_atomic_store((unsigned long long*)p, (unsigned long long)v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// IEEE binary16 equality on raw bits: NaN equals nothing, +0 equals -0.
__device__ __forceinline__ bool half_eq(uint64_t a, uint64_t b)
{
    const bool nan_a = (a & 0x7c00u) == 0x7c00u && (a & 0x03ffu) != 0;
    const bool nan_b = (b & 0x7c00u) == 0x7c00u && (b & 0x03ffu) != 0;
    if (nan_a || nan_b) return false;
    return a == b || ((a | b) & 0x7fffu) == 0;
}

// KeyEqual = EqualTo<Key> (== on the key type, utility/math_utils.h:67).
template <int KS>
__device__ __forceinline__ bool key_eq(uint64_t a, uint64_t b, uint32_t is_float)
{
    if (is_float) {
        if constexpr (KS == 2) return half_eq(a, b);
        else if constexpr (KS == 4) return __uint_as_float((uint32_t)a) == __uint_as_float((uint32_t)b);
        else if constexpr (KS == 8) return __longlong_as_double((long long)a) == __longlong_as_double((long long)b);
    }
    return a == b;
}

__device__ __forceinline__ bool stencil_ok(const OaArgs& a, uint64_t i)
{
    if (!a.stencil) return true;
    int64_t s;
    if (a.stencil_size == 1) s = (int8_t)a.stencil[i];
    else if (a.stencil_size == 4) s = ((const int32_t*)a.stencil)[i];
    else s = ((const int64_t*)a.stencil)[i];
    switch (a.stencil_pred) {
    case 1: return s != 0;
    case 2: return (s & 1) != 0;
    case 3: return (s & 1) == 0;
    default: return true;
    }
}

// The probe hashes, key equality and sentinel test of a walk.  The compiled
// containers take them from OaArgs at run time (hash family, float keys);
// a header-only instantiation supplies the caller's functor types instead
// (include/mxco/detail/caller_probe.h).  Every hash is reduced to the
// 32-bit probe size type, as SanitizeHash's cast does
// (include/utility/math_utils.h:19-29, bucket_storage_ref.h:21-29).
struct BuiltinFns {
    template <int KS>
    static __device__ __forceinline__ uint32_t hash1(const OaArgs& a, uint64_t key)
    {
        return probe_hash<KS>(a.hash, key, 0u);
    }
    template <int KS>
    static __device__ __forceinline__ uint32_t hash2(const OaArgs& a, uint64_t key)
    {
        return probe_hash<KS>(a.hash, key, 1u);
    }
    // KeyEqual(probe key, slot key)
    template <int KS>
    static __device__ __forceinline__ bool equal(const OaArgs& a, uint64_t slot_key, uint64_t key)
    {
        return key_eq<KS>(slot_key, key, a.key_float);
    }
    // EqualWrapper::IsSentinelValue: bitwise for integers and floating point
    template <int KS>
    static __device__ __forceinline__ bool sentinel(uint64_t key, uint64_t sentinel_key)
    {
        return key == sentinel_key;
    }
};

// Lanes per key of the kernels a caller instantiates with its own functor
// types (the library's grid for them assumes this width).
constexpr int kCallerGroup = 4;

// The probe walk of one key.  All fields are uniform across the key's group.
struct Walk {
    uint32_t pos;     // next slot in walk order
    uint32_t bucket;  // DoubleHashing: current bucket start
    uint32_t init;    // DoubleHashing: first bucket start
    uint32_t step;    // DoubleHashing: bucket stride
    uint32_t left;    // LinearProbing: slots not yet scanned
    bool done;
};

template <int KS, class Fns = BuiltinFns>
__device__ __forceinline__ Walk make_walk(const OaArgs& a, uint64_t key)
{
    Walk w;
    const uint32_t h1 = Fns::template hash1<KS>(a, key);
    const uint32_t start = (h1 % a.nbuckets) * a.bucket;
    w.pos = start;
    w.bucket = start;
    w.init = start;
    w.left = a.cap;
    w.done = false;
    if (a.probing) {
        const uint32_t h2 = Fns::template hash2<KS>(a, key);
        w.step = (h2 % (a.nbuckets - 1u) + 1u) * a.bucket;
    } else {
        w.step = a.bucket;
    }
    return w;
}

// Slots of the next window; never crosses a 128-byte line (linear) nor a
// bucket end (double).
template <int MAXW>
__device__ __forceinline__ uint32_t window_len(const OaArgs& a, const Walk& w)
{
    if (a.probing == 0) {
        // Optionally stop the first window at the end of its 64-byte sector:
        // most hits end there, and the sector is the unit HBM delivers.
        const uint32_t span = ((a.probe_flags & 1u) && w.left == a.cap && MAXW >= 2) ? MAXW / 2 : MAXW;
        uint32_t l = span - (w.pos & (span - 1));
        l = min(l, a.cap - w.pos);
        return min(l, w.left);
    }
    return min(w.bucket + a.bucket - w.pos, (uint32_t)MAXW);
}

__device__ __forceinline__ void consume(const OaArgs& a, Walk& w, uint32_t c)
{
    if (a.probing == 0) {
        w.left -= c;
        w.pos += c;
        if (w.pos == a.cap) w.pos = 0;
        w.done = (w.left == 0);
    } else {
        w.pos += c;
        if (w.pos == w.bucket + a.bucket) {
            // ProbingIterator::operator++ (probing_scheme_impl.inl:76-84)
            const uint32_t rem = a.cap - w.step;
            w.bucket = w.bucket >= rem ? w.bucket - rem : w.bucket + w.step;
            w.pos = w.bucket;
            w.done = (w.bucket == w.init);
        }
    }
}

// Payload store after a won key CAS on a 16-byte slot.  mode 0: plain
// store (written back from the XCD's L2 at the end of the launch); 1: agent-
// scope atomic store (visible to every XCD at once); 2: non-temporal store.
template <int N>
__device__ __forceinline__ void store_payload(uint8_t* p, uint64_t v, uint32_t mode)
{
    if (mode == 1) publish_bits<N>(p, v);
    else if (mode == 2) {
        if constexpr (N == 1) __builtin_nontemporal_store((uint8_t)v, p);
        else if constexpr (N == 2) __builtin_nontemporal_store((uint16_t)v, (uint16_t*)p);
        else if constexpr (N == 4) __builtin_nontemporal_store((uint32_t)v, (uint32_t*)p);
        else __builtin_nontemporal_store(v, (uint64_t*)p);
    } else {
        store_bits<N>(p, v);
    }
}

// Probes keys first, first + stride, ... < end of one table for the calling
// lane group and accumulates the per-lane counters.  Every lane group of the
// wavefront calls it with its own `first`; `end` and `stride` are uniform.
template <int KS, int VS, int G, int OP, int U = 1, class Fns = BuiltinFns>
__device__ __forceinline__ void oa_process(const OaArgs& a, uint64_t first, uint64_t end, uint64_t stride,
                                           unsigned long long& cnt0, unsigned long long& cnt1)
{
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    constexpr int SS = L::SS;
    // slots per window: a 128-byte line; 16 slots (16 or 32 bytes) for 1- and
    // 2-byte slots, one load per slot
    constexpr int SPL = SS >= 4 ? 128 / SS : 16;
    constexpr int K = SPL / G;     // slots per lane per window
    constexpr bool kInsertFamily = OP == OP_INSERT || OP == OP_INSERT_OR_ASSIGN || OP == OP_ASSIGN ||
                                   OP == OP_INSERT_AND_FIND || OP == OP_INSERT_STATS;
    constexpr int OUTS = VS ? VS : KS;  // Find / InsertAndFind output width
    constexpr int VOFF = L::VOFF;
    constexpr uint64_t kGroupMask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);

    const uint32_t lane = __lane_id();
    const uint32_t gl = lane & (G - 1);
    const uint32_t gbase = lane & ~(uint32_t)(G - 1);
    const bool leader = gl == 0;
    const uint64_t ngroups = stride;
    const uint64_t ekey = a.empty_key;
    const uint64_t miss_out = VS ? a.empty_val : ekey;

    // Per-group state: the key being probed and its walk.  The loop below
    // advances every group by one window per iteration; a group whose key
    // is resolved takes its next key at once instead of waiting for the
    // slowest walk of the wavefront.
    //
    // Each group keeps U keys in flight: slot u walks keys
    // gid + u*ngroups, gid + (u+U)*ngroups, ...  The window loads of all U
    // walks are issued before any of them is examined.
    const uint64_t gid = first;
    // RAW: the input pair as one word (slots of at most 8 bytes), padding
    // included: the whole-slot CAS stores it as it is.
    uint64_t I[U], KEY[U], PVAL[U], PREV[U], RAW[U];
    Walk W[U];
    uint32_t WLEN[U];
    Raw S[U][K];
    // Key (and payload) of slot u's next index, loaded one key ahead so a
    // group that finishes a key starts the next walk without a load round
    // trip.
    uint64_t NKEY[U], NPVAL[U], NRAW[U];
    auto prefetch = [&](int u, uint64_t idx) {
        if (idx >= end) return;
        if constexpr (kInsertFamily && SS <= 8) {
            const uint64_t raw = load_bits<SS>(a.in + idx * SS);
            NRAW[u] = raw;
            NKEY[u] = raw & L::KMASK;
            if constexpr (VS) NPVAL[u] = (raw >> (8 * VOFF)) & L::VMASK;
        } else if constexpr (kInsertFamily) {
            NKEY[u] = load_bits<KS>(a.in + idx * SS);
            if constexpr (VS) NPVAL[u] = load_bits<VS>(a.in + idx * SS + VOFF);
        } else {
            NKEY[u] = load_bits<KS>(a.in + idx * KS);
        }
    };

    // Sets up key `idx`.  Returns false when its outcome needs no walk
    // (false predicate, reserved key) -- the outcome is written here.
    auto begin = [&](int u, uint64_t idx) -> bool {
        uint64_t& key = KEY[u];
        uint64_t& pval = PVAL[u];
        key = NKEY[u];
        if constexpr (kInsertFamily && VS) pval = NPVAL[u];
        if constexpr (kInsertFamily && SS <= 8) RAW[u] = NRAW[u];
        prefetch(u, idx + (uint64_t)U * ngroups);
        if (!stencil_ok(a, idx)) {
            if (leader) {
                if constexpr (OP == OP_FIND) store_bits<OUTS>(a.out + idx * OUTS, miss_out);
                if constexpr (OP == OP_CONTAINS) a.out[idx] = 0;
            }
            return false;
        }
        if constexpr (kInsertFamily) {
            // The reserved empty key is never inserted (open_addressing_ref_impl.h:494-496).
            if (Fns::template sentinel<KS>(key, ekey)) {
                if (leader) {
                    if constexpr (OP == OP_INSERT || OP == OP_INSERT_OR_ASSIGN) cnt0++;
                    if constexpr (OP == OP_INSERT_AND_FIND) {
                        store_bits<OUTS>(a.out + idx * OUTS, miss_out);
                        a.out2[idx] = 0;
                    }
                }
                return false;
            }
        }
        W[u] = make_walk<KS, Fns>(a, key);
        PREV[u] = ~0ull;
        return true;
    };
    auto advance = [&](int u, uint64_t from) {
        uint64_t idx = from;
        while (idx < end && !begin(u, idx)) idx += (uint64_t)U * ngroups;
        I[u] = idx;
    };
#pragma unroll
    for (int u = 0; u < U; ++u) {
        W[u].done = true;
        NKEY[u] = 0;
        NPVAL[u] = 0;
        NRAW[u] = 0;
        prefetch(u, gid + (uint64_t)u * ngroups);
        advance(u, gid + (uint64_t)u * ngroups);
    }

    // Window loads of slot u (nothing when its key is exhausted or finished).
    auto load = [&](int u) {
        if (I[u] >= end || W[u].done) return;
        const uint32_t wlen = window_len<SPL>(a, W[u]);
        WLEN[u] = wlen;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t off = gl + k * G;
            if (off < wlen) {
                const Raw* src = (const Raw*)(a.table + (uint64_t)(W[u].pos + off) * SS);
                S[u][k] = (a.probe_flags & 2u) ? __builtin_nontemporal_load(src) : *src;
            }
        }
    };

    // One step of slot u's walk on the loaded window.
    auto step = [&](int u) {
        if (I[u] >= end) return;
        const uint64_t i = I[u];
        const uint64_t key = KEY[u];
        const uint64_t pval = PVAL[u];
        Walk& w = W[u];
        uint64_t& prev_line = PREV[u];
        Raw (&s)[K] = S[u];
        bool key_done = false;
        // Counter increments of this step, added once at its end: separate
        // conditional increments of the two counters get merged by the
        // compiler into one increment through a dynamically indexed (scratch)
        // pair.
        uint32_t add0 = 0, add1 = 0;
        if (w.done) {
            // Walk exhausted: full table / miss after a full cycle.
            if (leader) {
                if constexpr (OP == OP_FIND) store_bits<OUTS>(a.out + i * OUTS, miss_out);
                if constexpr (OP == OP_CONTAINS) a.out[i] = 0;
                if constexpr (OP == OP_INSERT || OP == OP_INSERT_OR_ASSIGN || OP == OP_ERASE) add0 = 1;
                if constexpr (OP == OP_INSERT_AND_FIND) {
                    store_bits<OUTS>(a.out + i * OUTS, miss_out);
                    a.out2[i] = 0;
                }
            }
            key_done = true;
        } else {
            const uint32_t wlen = WLEN[u];
            // First slot (in walk order) of this lane that ends the walk.
            int fk = K;
            bool feq = false;
            Raw ws = s[0];
#pragma unroll
            for (int k = K - 1; k >= 0; --k) {
                const uint32_t off = gl + k * G;
                if (off < wlen) {
                    const uint64_t sk = L::key(s[k]);
                    const bool em = Fns::template sentinel<KS>(sk, ekey);
                    const bool eq = !em && Fns::template equal<KS>(a, sk, key);
                    bool ev;
                    if constexpr (OP == OP_ERASE) ev = eq || (em && a.has_tombstone);
                    else ev = eq || em;
                    if (ev) {
                        fk = k;
                        feq = eq;
                        ws = s[k];
                    }
                }
            }
            // Group-wide first event: lowest round, then lowest lane.
            int wround = K;
            uint32_t wlane = 0;
            if constexpr (G == 1) {
                wround = fk;
            } else {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const uint64_t m = (__ballot(fk == k) >> gbase) & kGroupMask;
                    if (m && wround == K) {
                        wround = k;
                        wlane = __builtin_ctzll(m);
                    }
                }
            }
            const uint32_t woff = wlane + wround * G;
            if constexpr (OP == OP_PROBE_STATS) {
                // Slots the reference's walk reads in this window: up to and
                // including the slot that ends it.
                const uint32_t used = (wround == K) ? wlen : woff + 1;
                const uint64_t first = ((uint64_t)w.pos * SS) / 128;
                const uint64_t last = ((uint64_t)(w.pos + used - 1) * SS + SS - 1) / 128;
                if (leader) {
                    add0 = (uint32_t)(last - first + 1 - (first == prev_line ? 1 : 0));
                    add1 = used;
                }
                prev_line = last;
            }
            if (wround == K) {
                consume(a, w, wlen);
            } else {
                const bool winner = (gl == wlane);
                uint8_t* wslot = a.table + (uint64_t)(w.pos + woff) * SS;
                bool weq;
                if constexpr (G == 1) weq = feq;
                else weq = ((__ballot(winner && feq) >> gbase) & kGroupMask) != 0;

                if constexpr (OP == OP_FIND || OP == OP_CONTAINS || OP == OP_COUNT || OP == OP_PROBE_STATS) {
                    if (winner) {
                        if constexpr (OP == OP_FIND) {
                            store_bits<OUTS>(a.out + i * OUTS, weq ? L::val(ws) : miss_out);
                        } else if constexpr (OP == OP_CONTAINS) {
                            a.out[i] = weq ? 1 : 0;
                        } else if constexpr (OP == OP_COUNT) {
                            add0 = weq ? 1 : 0;
                        }
                    }
                    key_done = true;
                } else if constexpr (OP == OP_ASSIGN) {
                    // Assign (open_addressing_ref_impl.h:384-410): stop at empty, write payload on match.
                    if (winner && weq) store_bits<VS>(wslot + VOFF, pval);
                    key_done = true;
                } else if constexpr (OP == OP_ERASE) {
                    if (!weq) {
                        // empty slot with a tombstone configured: not present
                        w.done = true;
                    } else {
                        // AttemptInsertStable(slot, content, erasedSlot) (ref_impl.h:359).
                        int r = 0;  // 0 failed (continue), 1 erased, 2 already erased
                        if (winner) {
                            if constexpr (SS <= 8) {
                                const uint64_t expect = (uint64_t)ws;
                                const uint64_t desired = L::pack(expect, a.erased_key, a.empty_val);
                                const uint64_t old = cas_bits<SS>(wslot, expect, desired);
                                if (old == expect) r = 1;
                                else if (L::key((Raw)old) == a.erased_key) r = 2;
                            } else {
                                const uint64_t old = cas_bits<KS>(wslot, L::key(ws), a.erased_key);
                                if (old == L::key(ws)) {
                                    store_bits<VS>(wslot + VOFF, a.empty_val);
                                    r = 1;
                                } else if (old == a.erased_key) {
                                    r = 2;
                                }
                            }
                        }
                        if constexpr (G > 1)
                            r = ((__ballot(winner && r == 1) >> gbase) & kGroupMask) ? 1
                                : (((__ballot(winner && r == 2) >> gbase) & kGroupMask) ? 2 : 0);
                        if (r == 1) key_done = true;
                        else if (r == 2) w.done = true;  // counted as not erased next iteration
                        else consume(a, w, woff + 1);
                    }
                } else {
                    // Insert family: EQUAL -> duplicate; AVAILABLE -> CAS attempt
                    // (InsertVerdict, open_addressing_ref_impl.h:490-532).
                    int r = 0;          // 0 lost to another key, 1 inserted, 2 duplicate
                    uint64_t seen = 0;  // stored key (set) / payload (8-byte map slot) on a duplicate
                    bool seen_window = false;  // `seen` came from the window load (16-byte slots)
                    if (weq) {
                        r = 2;
                        seen = VS ? L::val(ws) : L::key(ws);
                        seen_window = true;
                        if constexpr (OP == OP_INSERT_OR_ASSIGN) {
                            // AssignPayload (ref_impl.h:210-220)
                            if (winner) store_bits<VS>(wslot + VOFF, pval);
                        }
                    } else {
                        if (winner) {
                            if constexpr (SS <= 8) {
                                // PackCas / PackCasInsertOrAssign: one CAS of the whole slot
                                // against (emptyKey, emptyValue) over the slot's own padding
                                // (PackCasExpected), storing the input pair as it is; slots
                                // under 4 bytes CAS inside their 32-bit word (NarrowCas)
                                // (ref_impl.h:109-162, 234-257).
                                const uint64_t expect = L::pack((uint64_t)ws, ekey, a.empty_val);
                                const uint64_t desired = RAW[u];
                                const uint64_t old = cas_bits<SS>(wslot, expect, desired);
                                if (old == expect) {
                                    r = 1;
                                } else if (Fns::template equal<KS>(a, L::key((Raw)old), key)) {
                                    r = 2;
                                    seen = VS ? L::val((Raw)old) : L::key((Raw)old);
                                    if constexpr (OP == OP_INSERT_OR_ASSIGN) store_bits<SS>(wslot, desired);
                                }
                            } else {
                                // CasDependentWrite / BackToBackCasInsertOrAssign: CAS the key,
                                // then store the payload (ref_impl.h:165-185, 260-279).
                                const uint64_t old = cas_bits<KS>(wslot, ekey, key);
                                if (old == ekey) {
                                    if constexpr (OP == OP_INSERT_AND_FIND) publish_bits<VS>(wslot + VOFF, pval);
                                    else store_payload<VS>(wslot + VOFF, pval, a.store_mode);
                                    r = 1;
                                } else if (Fns::template equal<KS>(a, old, key)) {
                                    r = 2;
                                    if constexpr (OP == OP_INSERT_OR_ASSIGN) store_bits<VS>(wslot + VOFF, pval);
                                }
                            }
                        }
                        if constexpr (OP == OP_INSERT_STATS) {
                            if (winner) {
                                add0 = r == 0;  // lost to another key: the walk goes on
                                add1 = 1;
                            }
                        }
                        if constexpr (G > 1) {
                            const uint64_t m1 = (__ballot(winner && r == 1) >> gbase) & kGroupMask;
                            const uint64_t m2 = (__ballot(winner && r == 2) >> gbase) & kGroupMask;
                            r = m1 ? 1 : (m2 ? 2 : 0);
                        }
                    }
                    if (r == 0) {
                        consume(a, w, woff + 1);
                    } else {
                        if (leader) {
                            if constexpr (OP == OP_INSERT) add0 = (r == 2);
                            if constexpr (OP == OP_INSERT_OR_ASSIGN) add1 = (r == 2);
                        }
                        if constexpr (OP == OP_INSERT_AND_FIND) {
                            if (winner) {
                                uint64_t v;
                                if (r == 1) {
                                    v = VS ? pval : key;
                                } else if constexpr (VS && SS == 16) {
                                    // WaitForPayload (ref_impl.h:474-486): a racing inserter may
                                    // have stored the key but not yet the payload.  Bounded here.
                                    // The window's 16-byte load already holds the final payload
                                    // unless it still read the empty value.
                                    v = seen_window ? seen : a.empty_val;
                                    if (v == a.empty_val) v = atomic_load_bits<VS>(wslot + VOFF);
                                    for (int spin = 0; spin < (1 << 22) && v == a.empty_val; ++spin)
                                        v = atomic_load_bits<VS>(wslot + VOFF);
                                } else {
                                    v = seen;
                                }
                                store_bits<OUTS>(a.out + i * OUTS, v);
                                a.out2[i] = (r == 1) ? 1 : 0;
                            }
                        }
                        key_done = true;
                    }
                }
            }
        }
        cnt0 += add0;
        cnt1 += add1;
        if (key_done) advance(u, i + (uint64_t)U * ngroups);
    };

    for (;;) {
        bool any = false;
#pragma unroll
        for (int u = 0; u < U; ++u) any |= I[u] < end;
        if (!any) break;
#pragma unroll
        for (int u = 0; u < U; ++u) load(u);
#pragma unroll
        for (int u = 0; u < U; ++u) step(u);
    }

}

// A table's counters: kCounterSlots pairs [failed / count, assigned], each
// workgroup adding into pair blockIdx.x % kCounterSlots and the host summing
// them, so that a batch with many failures (duplicates) does not serialise
// on one address.
constexpr uint32_t kCounterSlots = 64;

__device__ __forceinline__ unsigned long long* counter_slot(unsigned long long* counters)
{
    return counters + 2 * (blockIdx.x % kCounterSlots);
}

// Wave reduction of the two counters, one atomic per wavefront.
template <int OP>
__device__ __forceinline__ void flush_counters(unsigned long long* counters, unsigned long long cnt0,
                                               unsigned long long cnt1)
{
    constexpr bool kTwoCounters = OP == OP_INSERT_OR_ASSIGN || OP == OP_PROBE_STATS || OP == OP_INSERT_STATS;
    if constexpr (OP == OP_INSERT || OP == OP_INSERT_OR_ASSIGN || OP == OP_COUNT || OP == OP_ERASE ||
                  OP == OP_PROBE_STATS || OP == OP_INSERT_STATS) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            cnt0 += __shfl_xor(cnt0, o);
            if constexpr (kTwoCounters) cnt1 += __shfl_xor(cnt1, o);
        }
        if (__lane_id() == 0) {
            if (cnt0) atomicAdd(&counters[0], cnt0);
            if constexpr (kTwoCounters) {
                if (cnt1) atomicAdd(&counters[1], cnt1);
            }
        }
    }
}

template <int KS, int VS, int G, int OP, int U = 1, class Fns = BuiltinFns>
__global__ __launch_bounds__(256) void oa_kernel(OaArgs a)
{
    unsigned long long cnt0 = 0, cnt1 = 0;
    const uint64_t gid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    const uint64_t ngroups = ((uint64_t)gridDim.x * blockDim.x) / G;
    oa_process<KS, VS, G, OP, U, Fns>(a, gid, a.n, ngroups, cnt0, cnt1);
    if (a.counters) flush_counters<OP>(counter_slot(a.counters), cnt0, cnt1);
}

// Chunk-staged variant for the streaming bulk operations.  A workgroup takes
// chunks of kStageKeys consecutive inputs, copies the chunk's inputs into
// LDS with coalesced non-temporal loads, lets its lane groups probe from LDS
// and write results to LDS, then streams the results out with coalesced
// non-temporal stores.  Without it, lane groups that have drifted apart
// write single results into many lines at different times and the L2 writes
// each line back several times.
constexpr uint32_t kStageKeys = 1024;

template <int KS, int VS, int G, int OP>
__global__ __launch_bounds__(256) void oa_staged_kernel(OaArgs a)
{
    constexpr int SS = Lay<KS, VS>::SS;
    constexpr bool kInsertFamily =
        OP == OP_INSERT || OP == OP_INSERT_OR_ASSIGN || OP == OP_ASSIGN || OP == OP_INSERT_AND_FIND;
    constexpr int IN = kInsertFamily ? SS : KS;
    constexpr int OUT = (OP == OP_FIND) ? (VS ? VS : KS) : (OP == OP_CONTAINS ? 1 : 0);
    __shared__ __align__(16) uint8_t sin[kStageKeys * IN];
    __shared__ __align__(16) uint8_t sout[OUT ? kStageKeys * OUT : 16];
    unsigned long long cnt0 = 0, cnt1 = 0;
    const uint32_t gpb = blockDim.x / G;
    const uint32_t gj = threadIdx.x / G;
    const uint64_t nchunks = (a.n + kStageKeys - 1) / kStageKeys;
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint64_t base = c * kStageKeys;
        const uint32_t len = (uint32_t)min((uint64_t)kStageKeys, a.n - base);
        const uint32_t in_bytes = len * IN;
        // 4-byte granules: inputs of 4, 8 and 16 bytes are all multiples of 4
        for (uint32_t t = threadIdx.x; t < in_bytes / 4; t += blockDim.x)
            ((uint32_t*)sin)[t] = __builtin_nontemporal_load((const uint32_t*)(a.in + base * IN) + t);
        __syncthreads();
        OaArgs la = a;
        la.in = sin;
        la.out = OUT ? sout : nullptr;
        la.n = len;
        if (a.stencil) la.stencil = a.stencil + base * a.stencil_size;
        oa_process<KS, VS, G, OP, 1>(la, gj, len, gpb, cnt0, cnt1);
        __syncthreads();
        if constexpr (OUT == 8) {
            for (uint32_t t = threadIdx.x; t < len; t += blockDim.x)
                __builtin_nontemporal_store(((const uint64_t*)sout)[t], (uint64_t*)(a.out + base * 8) + t);
        } else if constexpr (OUT == 4) {
            for (uint32_t t = threadIdx.x; t < len; t += blockDim.x)
                __builtin_nontemporal_store(((const uint32_t*)sout)[t], (uint32_t*)(a.out + base * 4) + t);
        } else if constexpr (OUT == 1) {
            for (uint32_t t = threadIdx.x; t < len; t += blockDim.x) a.out[base + t] = sout[t];
        }
        if constexpr (OUT != 0) __syncthreads();
    }
    if (a.counters) flush_counters<OP>(counter_slot(a.counters), cnt0, cnt1);
}

// Find / Contains with wave-staged outputs.  Each wavefront owns chunks of
// up to kWaveChunk consecutive keys of one table (two chunks open at a time)
// and hands their keys to its lane groups as groups free up, so the flat
// per-group walk is kept.  Results go to the chunk's LDS buffer; when every
// key of a chunk is resolved the wavefront writes the chunk's outputs as
// whole lines.  With plain per-key stores, groups that drifted apart write
// one result at a time into the same output line and the L2 writes that
// line back several times.
constexpr uint32_t kWaveChunk = 128;

// One chunk: its table, and its input / output / stencil already offset to
// the chunk's first key.  len == 0: the source is exhausted.
struct WsChunk {
    uint8_t* table;
    const uint8_t* in;
    uint8_t* out;
    const uint8_t* stencil;
    uint32_t cap;
    uint32_t nbuckets;
    uint32_t len;
};

// Chunks of one table: wavefront w takes chunks w, w + waves, ...
template <int KS, int OUTB>
struct PlainChunks {
    static constexpr bool kOneTable = true;
    const OaArgs* a;
    uint64_t next, step;
    __device__ __forceinline__ WsChunk open()
    {
        WsChunk c;
        const uint64_t base = next * kWaveChunk;
        next += step;
        c.table = a->table;
        c.cap = a->cap;
        c.nbuckets = a->nbuckets;
        c.len = base < a->n ? (uint32_t)min((uint64_t)kWaveChunk, a->n - base) : 0u;
        c.in = a->in + base * KS;
        c.out = a->out + base * OUTB;
        c.stencil = a->stencil ? a->stencil + base * a->stencil_size : nullptr;
        return c;
    }
};

template <int KS, int VS, int G, int OP, typename Src>
__device__ __forceinline__ void wave_staged_find(const OaArgs& a, Src& src, uint8_t* st0, uint8_t* st1,
                                                 uint8_t* kb0, uint8_t* kb1)
{
    static_assert(OP == OP_FIND || OP == OP_CONTAINS, "Find / Contains only");
    using L = Lay<KS, VS>;
    using Raw = typename L::Raw;
    constexpr int SS = L::SS;
    constexpr int SPL = 128 / SS;
    constexpr int K = SPL / G;
    constexpr int OUTB = OP == OP_CONTAINS ? 1 : (VS ? VS : KS);
    constexpr uint64_t kGroupMask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);

    const uint32_t lane = __lane_id();
    const uint32_t gl = lane & (G - 1);
    const uint32_t gbase = lane & ~(uint32_t)(G - 1);
    const bool leader = gl == 0;
    const uint64_t ekey = a.empty_key;
    const uint64_t miss_out = OP == OP_CONTAINS ? 0 : (VS ? a.empty_val : ekey);

    // Wave-uniform state of the two open chunks: keys handed out and keys
    // resolved.  `old` is the chunk opened first.
    // The keys of an open chunk are copied to its LDS key buffer when it
    // opens (one coalesced load round trip per chunk), so a group that takes
    // a new key reads it from LDS and issues its first window load in the
    // same iteration.
    auto stage_keys = [&](uint8_t* kb, const WsChunk& c) {
        static_assert(kWaveChunk == 128, "two keys per lane");
        using KT = typename std::conditional<KS == 8, uint64_t, uint32_t>::type;
        KT v0 = 0, v1 = 0;
        if (lane < c.len) v0 = ((const KT*)c.in)[lane];
        if (lane + 64 < c.len) v1 = ((const KT*)c.in)[lane + 64];
        ((KT*)kb)[lane] = v0;
        ((KT*)kb)[lane + 64] = v1;
    };
    WsChunk c0 = src.open();
    WsChunk c1 = src.open();
    stage_keys(kb0, c0);
    stage_keys(kb1, c1);
    uint32_t ch0 = 0, ch1 = 0, cd0 = 0, cd1 = 0;
    int old = 0;

    // Per-group state (uniform within the group).
    bool active = false;
    int B = 0;       // chunk buffer of the group's key
    uint32_t I = 0;  // key index within the chunk
    uint64_t key = 0;
    OaArgs ga = a;   // the key's table
    Walk w;
    w.done = true;
    Raw s[K];

    auto put = [&](uint64_t v) {
        uint8_t* d = (B ? st1 : st0) + I * OUTB;
        if constexpr (OUTB == 1) *d = (uint8_t)v;
        else if constexpr (OUTB == 4) *(uint32_t*)d = (uint32_t)v;
        else *(uint64_t*)d = v;
    };
    auto flush = [&](const uint8_t* st, const WsChunk& c) {
        for (uint32_t j = lane; j < c.len; j += 64) {
            if constexpr (OUTB == 1) c.out[j] = st[j];
            else if constexpr (OUTB == 4) ((uint32_t*)c.out)[j] = ((const uint32_t*)st)[j];
            else ((uint64_t*)c.out)[j] = ((const uint64_t*)st)[j];
        }
    };

    for (;;) {
        if (c0.len == 0 && c1.len == 0) break;  // every chunk resolved and written
        bool key_done = false;

        // Hand the open chunks' next keys to the groups without one, oldest chunk first.
        const uint64_t need = __ballot(leader && !active);
        const uint32_t av_old = old ? c1.len - ch1 : c0.len - ch0;
        const uint32_t av_new = old ? c0.len - ch0 : c1.len - ch1;
        if (need != 0 && (av_old | av_new) != 0) {
            const uint32_t nneed = (uint32_t)__popcll(need);
            if (!active) {
                const uint32_t rank = (uint32_t)__popcll(need & ((1ull << gbase) - 1ull));
                bool fresh = false;
                if (rank < av_old) {
                    B = old;
                    I = (old ? ch1 : ch0) + rank;
                    fresh = true;
                } else if (rank - av_old < av_new) {
                    B = old ^ 1;
                    I = (old ? ch0 : ch1) + (rank - av_old);
                    fresh = true;
                }
                if (fresh) {
                    active = true;
                    if constexpr (!Src::kOneTable) {  // one table: the fields stay wave-uniform
                        ga.table = B ? c1.table : c0.table;
                        ga.cap = B ? c1.cap : c0.cap;
                        ga.nbuckets = B ? c1.nbuckets : c0.nbuckets;
                    }
                    ga.stencil = B ? c1.stencil : c0.stencil;
                    if (!stencil_ok(ga, I)) {
                        if (leader) put(miss_out);
                        key_done = true;
                    } else {
                        key = load_bits<KS>((B ? kb1 : kb0) + (uint64_t)I * KS);
                        w = make_walk<KS>(ga, key);
                    }
                }
            }
            const uint32_t t_old = min(nneed, av_old);
            const uint32_t t_new = min(nneed - t_old, av_new);
            if (old) {
                ch1 += t_old;
                ch0 += t_new;
            } else {
                ch0 += t_old;
                ch1 += t_new;
            }
        }

        if (active && !key_done) {
            if (w.done) {
                // walk exhausted after a full cycle: a miss
                if (leader) put(miss_out);
                key_done = true;
            } else {
                const uint32_t wlen = window_len<SPL>(ga, w);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const uint32_t off = gl + k * G;
                    if (off < wlen) s[k] = *(const Raw*)(ga.table + (uint64_t)(w.pos + off) * SS);
                }
                int fk = K;
                bool feq = false;
                Raw ws = s[0];
#pragma unroll
                for (int k = K - 1; k >= 0; --k) {
                    const uint32_t off = gl + k * G;
                    if (off < wlen) {
                        const uint64_t sk = L::key(s[k]);
                        const bool em = sk == ekey;
                        const bool eq = !em && key_eq<KS>(sk, key, a.key_float);
                        if (eq || em) {
                            fk = k;
                            feq = eq;
                            ws = s[k];
                        }
                    }
                }
                int wround = K;
                uint32_t wlane = 0;
                if constexpr (G == 1) {
                    wround = fk;
                } else {
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const uint64_t m = (__ballot(fk == k) >> gbase) & kGroupMask;
                        if (m && wround == K) {
                            wround = k;
                            wlane = __builtin_ctzll(m);
                        }
                    }
                }
                if (wround == K) {
                    consume(ga, w, wlen);
                } else {
                    if (gl == wlane) {
                        if constexpr (OP == OP_FIND) put(feq ? L::val(ws) : miss_out);
                        else put(feq ? 1 : 0);
                    }
                    key_done = true;
                }
            }
        }

        cd0 += (uint32_t)__popcll(__ballot(leader && key_done && B == 0));
        cd1 += (uint32_t)__popcll(__ballot(leader && key_done && B == 1));
        if (key_done) active = false;

        if (c0.len != 0 && cd0 == c0.len) {
            flush(st0, c0);
            c0 = src.open();
            stage_keys(kb0, c0);
            ch0 = 0;
            cd0 = 0;
            if (old == 0) old = 1;
        }
        if (c1.len != 0 && cd1 == c1.len) {
            flush(st1, c1);
            c1 = src.open();
            stage_keys(kb1, c1);
            ch1 = 0;
            cd1 = 0;
            if (old == 1) old = 0;
        }
    }
}

template <int KS, int VS, int G, int OP>
__global__ __launch_bounds__(256) void oa_wave_staged_kernel(OaArgs a)
{
    constexpr int OUTB = OP == OP_CONTAINS ? 1 : (VS ? VS : KS);
    __shared__ __align__(16) uint8_t stage[4][2][kWaveChunk * OUTB];
    __shared__ __align__(16) uint8_t keys[4][2][kWaveChunk * KS];
    const uint32_t wib = threadIdx.x >> 6;
    PlainChunks<KS, OUTB> src{&a, (uint64_t)blockIdx.x * (blockDim.x >> 6) + wib,
                              (uint64_t)gridDim.x * (blockDim.x >> 6)};
    wave_staged_find<KS, VS, G, OP>(a, src, stage[wib][0], stage[wib][1], keys[wib][0], keys[wib][1]);
}


// Sentinel fill (Clear / ClearAsync, open_addressing_impl.h:341-452):
// streams the 16-byte replicated slot image over the whole slot array.
// Plain 16-byte stores from a grid of 64 workgroups per CU measured fastest
// on MI355X (5.8 TB/s over 4.3 GB; non-temporal stores and smaller grids
// reached 4.6-5.1 TB/s).
constexpr unsigned kFillBlocksPerCu = 64;

static __global__ __launch_bounds__(256) void oa_fill_kernel(uint8_t* table, uint64_t bytes, u64x2 pattern)
{
    const uint64_t n16 = bytes / 16;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    u64x2* p = (u64x2*)table;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) p[i] = pattern;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the last bytes 15 or fewer, one at a time (1- and 2-byte slots can
        // leave any count)
        const uint8_t* pb = (const uint8_t*)&pattern;
        for (uint64_t b = n16 * 16; b < bytes; ++b) table[b] = pb[b % 16];
    }
}

static __global__ void zero_counters_kernel(unsigned long long* c, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) c[i] = 0;
}

}  // namespace mxco
