// STL-style C++ front end of libmxco.so: the building blocks shared by the
// container headers (extents, hashers, probing schemes, storage, pairs,
// predicates and the status -> exception mapping).
//
// The names and template shapes follow the reference's public headers so
// that code written against them compiles after changing the include and
// namespace (include/extent.h:20-99, include/hash_functions.h:17-30,
// include/probing_scheme.h:37-122, include/storage.h:16-20,
// include/pair.h:41, include/utility/math_utils.h:67).  The library's own
// hashers, EqualTo and the built-in predicates select a configuration of the
// compiled kernels; a caller's KeyEqual, hasher or predicate type has its
// kernels compiled in the caller's translation unit instead
// (detail/caller_probe.h, HIP compiler required).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "../mxco.h"
#include "detail/hash.h"

namespace mxco {

static constexpr std::size_t dynamicExtent = static_cast<std::size_t>(-1);

// Extent<SizeType, N> (include/extent.h:20-60): a static or dynamic count.
template <typename SizeType, std::size_t N = dynamicExtent>
struct Extent {
    static_assert(std::is_integral_v<SizeType> && std::is_unsigned_v<SizeType>,
                  "SizeType must be an unsigned integral type");
    using ValueType = SizeType;
    constexpr Extent() = default;
    constexpr Extent(SizeType) noexcept {}
    constexpr operator SizeType() const noexcept { return N; }
};

template <typename SizeType>
struct Extent<SizeType, dynamicExtent> {
    static_assert(std::is_integral_v<SizeType> && std::is_unsigned_v<SizeType>,
                  "SizeType must be an unsigned integral type");
    using ValueType = SizeType;
    constexpr Extent(SizeType size) noexcept : size_(size) {}
    constexpr operator SizeType() const noexcept { return size_; }

private:
    SizeType size_;
};

// Storage<B> (include/storage.h:16-20): the bucket size.
template <int32_t BucketSize>
struct Storage {
    static constexpr int32_t bucketSize = BucketSize;
};

// Hasher tags (include/hash_functions.h:17-30).  Seeds are fixed by the
// containers, as in the reference (stateful hashers are rejected).
template <typename Key>
struct murmurhash3_32 {
    static constexpr mxco_hash id = MXCO_HASH_MURMUR3_32;
    constexpr bool operator==(murmurhash3_32 const&) const noexcept { return true; }
};
template <typename Key>
struct xxhash_32 {
    static constexpr mxco_hash id = MXCO_HASH_XXHASH_32;
    constexpr bool operator==(xxhash_32 const&) const noexcept { return true; }
};
// xxhash_64 carries a seed: BloomFilterPolicy is built from Hasher{seed}
// (include/bloom_filter_policy.h:78).  Open-addressing containers use seeds 0
// and 1 whatever it holds, as the reference's stateless probing does.
template <typename Key>
struct xxhash_64 {
    using ArgumentType = Key;
    using ResultType = std::uint64_t;
    static constexpr mxco_hash id = MXCO_HASH_XXHASH_64;
    constexpr xxhash_64(std::uint64_t seed = 0) noexcept : seed_{seed} {}
    constexpr std::uint64_t Seed() const noexcept { return seed_; }
    std::uint64_t operator()(Key const& key) const noexcept
    {
        static_assert(sizeof(Key) == 1 || sizeof(Key) == 2 || sizeof(Key) == 4 || sizeof(Key) == 8,
                      "xxhash_64: 1-, 2-, 4- or 8-byte keys");
        std::uint64_t bits = 0;
        std::memcpy(&bits, &key, sizeof(Key));
        return xxhash64<sizeof(Key)>(bits, seed_);
    }
    constexpr bool operator==(xxhash_64 const& o) const noexcept { return seed_ == o.seed_; }
    constexpr bool operator!=(xxhash_64 const& o) const noexcept { return seed_ != o.seed_; }

private:
    std::uint64_t seed_;
};
template <typename Key>
struct murmurhash3_fmix32 {
    static constexpr mxco_hash id = MXCO_HASH_FMIX_32;
    constexpr bool operator==(murmurhash3_fmix32 const&) const noexcept { return true; }
};
template <typename Key>
struct murmurhash3_fmix64 {
    static constexpr mxco_hash id = MXCO_HASH_FMIX_64;
    constexpr bool operator==(murmurhash3_fmix64 const&) const noexcept { return true; }
};

namespace detail {
// A hasher of the library (a compiled hash family) carries its id.
template <typename H, typename = void>
struct HasHashId : std::false_type {};
template <typename H>
struct HasHashId<H, std::void_t<decltype(H::id)>> : std::true_type {};
template <typename H>
constexpr mxco_hash hash_id_or_default()
{
    if constexpr (HasHashId<H>::value) return H::id;
    else return MXCO_HASH_MURMUR3_32;  // unused: the caller's kernels hash
}
}  // namespace detail

// Probing schemes (include/probing_scheme.h:37-122).  Any hasher type works:
// the library's own select the compiled kernels, a caller's runs in kernels
// compiled with it.  `compiled` tells which.
template <typename Hash>
struct LinearProbing {
    using Hasher = Hash;
    static constexpr mxco_probing id = MXCO_PROBING_LINEAR;
    static constexpr bool compiled = detail::HasHashId<Hash>::value;
    static constexpr mxco_hash hash = detail::hash_id_or_default<Hash>();
    constexpr bool operator==(LinearProbing const&) const noexcept { return true; }
};

template <typename Hash1, typename Hash2 = Hash1>
struct DoubleHashing {
    static constexpr mxco_probing id = MXCO_PROBING_DOUBLE;
    // the compiled kernels hash with one family at seeds 0 and 1
    static constexpr bool compiled = detail::HasHashId<Hash1>::value && detail::HasHashId<Hash2>::value &&
                                     detail::hash_id_or_default<Hash1>() == detail::hash_id_or_default<Hash2>();
    static constexpr mxco_hash hash = detail::hash_id_or_default<Hash1>();
    constexpr bool operator==(DoubleHashing const&) const noexcept { return true; }
};

// KeyEqual default (include/utility/math_utils.h:67).
template <typename T>
struct EqualTo {
    constexpr bool operator()(T const& lhs, T const& rhs) const noexcept { return lhs == rhs; }
    constexpr bool operator==(EqualTo const&) const noexcept { return true; }
};

// DefaultAllocator<T> (include/utility/allocator.h:20-50): device memory
// from hipMalloc; Allocate returns nullptr on failure.
template <typename T>
class DefaultAllocator {
public:
    using ValueType = T;
    using value_type = T;
    DefaultAllocator() = default;
    template <class U>
    DefaultAllocator(DefaultAllocator<U> const&) noexcept
    {
    }
    ValueType* Allocate(std::size_t n) const
    {
        if (n > static_cast<std::size_t>(-1) / sizeof(ValueType)) return nullptr;
        void* p = nullptr;
        if (hipMalloc(&p, sizeof(ValueType) * n) != hipSuccess) return nullptr;
        return static_cast<ValueType*>(p);
    }
    void Deallocate(ValueType* p)
    {
        if (p != nullptr) (void)hipFree(p);
    }
};

// Key and payload types compared as floating point: the standard floating
// types, and IEEE binary16 where the compiler has _Float16.
template <typename T>
struct IsFloatKey : std::is_floating_point<T> {};
#if defined(__FLT16_MAX__)
template <>
struct IsFloatKey<_Float16> : std::true_type {};
#endif

// Pair<F, S> (include/pair.h:41): the slot layout of a map.
namespace detail {
constexpr std::size_t bit_ceil(std::size_t n)
{
    std::size_t x = 1;
    while (x < n) x <<= 1;
    return x;
}
constexpr std::size_t pair_alignment(std::size_t a, std::size_t b)
{
    return bit_ceil(a + b) < 16 ? bit_ceil(a + b) : 16;
}
}  // namespace detail

template <typename First, typename Second>
struct alignas(detail::pair_alignment(sizeof(First), sizeof(Second))) Pair {
    using FirstType = First;
    using SecondType = Second;
    First first;
    Second second;
};

// Stencil predicates.  The reference default-constructs a predicate type on
// the device; the compiled kernels provide these four, and any other
// predicate type is evaluated by a kernel compiled with it
// (detail/caller_probe.h).
struct AlwaysTrue {
    static constexpr mxco_pred id = MXCO_PRED_ALWAYS;
};
struct IsNonZero {
    static constexpr mxco_pred id = MXCO_PRED_NONZERO;
};
struct IsOdd {
    static constexpr mxco_pred id = MXCO_PRED_ODD;
};
struct IsEven {
    static constexpr mxco_pred id = MXCO_PRED_EVEN;
};

namespace detail {

template <typename P, typename = void>
struct HasPredId : std::false_type {};
template <typename P>
struct HasPredId<P, std::void_t<decltype(P::id)>> : std::true_type {};

// Throws the exception type the reference throws for the status.
inline void check(mxco_status s)
{
    if (s == MXCO_OK) return;
    const std::string msg = mxco_last_error();
    switch (s) {
    case MXCO_INVALID_ARGUMENT: throw std::invalid_argument(msg);
    case MXCO_LENGTH_ERROR: throw std::length_error(msg);
    case MXCO_OUT_OF_MEMORY: throw std::bad_alloc();
    case MXCO_LOGIC_ERROR: throw std::logic_error(msg);
    case MXCO_OVERFLOW_ERROR: throw std::overflow_error(msg);
    default: throw std::runtime_error(msg);
    }
}

template <typename StencilT, typename Predicate>
mxco_stencil make_stencil(StencilT const* stencil)
{
    static_assert(HasPredId<Predicate>::value, "a built-in predicate: AlwaysTrue, IsNonZero, IsOdd, IsEven");
    static_assert(sizeof(StencilT) == 1 || sizeof(StencilT) == 4 || sizeof(StencilT) == 8,
                  "stencil elements of the built-in predicates are 1, 4 or 8 bytes");
    mxco_stencil s;
    s.data = stencil;
    s.elem_size = sizeof(StencilT);
    s.pred = Predicate::id;
    return s;
}

// The stencil an operation runs with: a built-in predicate goes to the
// compiled kernels as it is; the caller's (detail/caller_probe.h) is
// evaluated first into one byte per input.
template <typename StencilT, typename Predicate, bool Builtin = HasPredId<Predicate>::value>
class Selection;

template <typename StencilT, typename Predicate>
class Selection<StencilT, Predicate, true> {
public:
    Selection(StencilT const* stencil, Predicate const&, uint64_t, hipStream_t)
        : s_(make_stencil<StencilT, Predicate>(stencil))
    {
    }
    const mxco_stencil* get() const noexcept { return &s_; }

private:
    mxco_stencil s_;
};

#if !defined(__HIPCC__)
template <typename StencilT, typename Predicate>
class Selection<StencilT, Predicate, false> {
    static_assert(sizeof(Predicate) == 0,
                  "a caller-defined predicate runs on the device: compile this translation unit with hipcc");
};
#endif

// The reference rejects a KeyEqual or probing scheme that differs from its
// default-constructed value (open_addressing_impl.h:527-532), where the
// type compares.
template <typename T>
auto equals_default(T const& x, int) -> decltype(static_cast<bool>(x == T{}))
{
    return static_cast<bool>(x == T{});
}
template <typename T>
bool equals_default(T const&, long)
{
    return true;
}

// Whether a (KeyEqual, ProbingScheme) pair runs on the compiled kernels.
template <typename Key, typename KeyEqual, typename ProbingScheme>
inline constexpr bool kCompiledProbe = std::is_same_v<KeyEqual, EqualTo<Key>> && ProbingScheme::compiled;

}  // namespace detail
}  // namespace mxco
