// mxco::StaticMultimap -- drop-in for aclco::StaticMultimap
// (include/static_multimap.h:87-556).
//
// Same member functions and argument order; the stream argument is a
// hipStream_t.  Every call forwards to the C ABI of libmxco.so
// (include/mxco.h) and rethrows its status as the reference's exception type:
// std::invalid_argument for a zero capacity, a null pointer that must not be
// null or a selected pair that uses a sentinel; std::length_error when
// Retrieve/RetrieveAll outputs do not fit (outputs untouched).
#pragma once

#include <memory>
#include <utility>

#include "common.h"
#if defined(__HIPCC__)
#include "detail/caller_probe.h"
#endif

namespace mxco {

template <class Key, class Value>
class StaticMultimap {
    static_assert(std::is_integral_v<Key> && std::is_integral_v<Value> && (sizeof(Key) == 4 || sizeof(Key) == 8) &&
                      (sizeof(Value) == 4 || sizeof(Value) == 8),
                  "StaticMultimap holds 32- or 64-bit integral keys and values");

public:
    using SizeType = std::size_t;
    using ExtentType = mxco::Extent<std::size_t>;
    using KeyType = Key;
    using MappedType = Value;

    StaticMultimap(ExtentType capacity, Key emptyKey, Value emptyValue, hipStream_t stream = nullptr)
    {
        mxco_mm* h = nullptr;
        detail::check(mxco_mm_create(sizeof(Key), sizeof(Value), static_cast<SizeType>(capacity), &emptyKey,
                                     &emptyValue, stream, &h));
        h_.reset(h);
    }
    ~StaticMultimap() = default;
    StaticMultimap(StaticMultimap const&) = delete;
    StaticMultimap& operator=(StaticMultimap const&) = delete;
    StaticMultimap(StaticMultimap&&) noexcept = default;
    StaticMultimap& operator=(StaticMultimap&&) noexcept = default;

    // Exact pair capacity; zero for a moved-from object.
    SizeType Capacity() const noexcept
    {
        uint64_t c = 0;
        if (h_) (void)mxco_mm_capacity(h_.get(), &c);
        return c;
    }

    SizeType Size(hipStream_t stream = nullptr) const
    {
        uint64_t s = 0;
        detail::check(mxco_mm_size(get(), stream, &s));
        return s;
    }

    void Clear(hipStream_t stream = nullptr) { detail::check(mxco_mm_clear(get(), stream)); }

    // Admits the input-order prefix that fits; returns the pairs not admitted.
    SizeType Insert(void* pairs, ExtentType n, hipStream_t stream = nullptr)
    {
        uint64_t failed = 0;
        detail::check(mxco_mm_insert(get(), pairs, nullptr, static_cast<SizeType>(n), stream, &failed));
        return failed;
    }

    template <class Stencil, class Predicate>
    SizeType InsertIf(void* pairs, Stencil* stencil, ExtentType n, hipStream_t stream = nullptr)
    {
        detail::Selection<Stencil, Predicate> s(stencil, Predicate{}, static_cast<SizeType>(n), stream);
        uint64_t failed = 0;
        detail::check(mxco_mm_insert(get(), pairs, s.get(), static_cast<SizeType>(n), stream, &failed));
        return failed;
    }

    void Contains(void* keys, void* output, ExtentType n, hipStream_t stream = nullptr) const
    {
        detail::check(
            mxco_mm_contains(get(), keys, nullptr, static_cast<uint8_t*>(output), static_cast<SizeType>(n), stream));
    }

    template <class Stencil, class Predicate>
    void ContainsIf(void* keys, Stencil* stencil, void* output, ExtentType n, hipStream_t stream = nullptr) const
    {
        detail::Selection<Stencil, Predicate> s(stencil, Predicate{}, static_cast<SizeType>(n), stream);
        detail::check(
            mxco_mm_contains(get(), keys, s.get(), static_cast<uint8_t*>(output), static_cast<SizeType>(n), stream));
    }

    // Smallest value stored for each key, or the empty value.
    void Find(void* keys, void* output, ExtentType n, hipStream_t stream = nullptr) const
    {
        detail::check(mxco_mm_find(get(), keys, nullptr, output, static_cast<SizeType>(n), stream));
    }

    template <class Stencil, class Predicate>
    void FindIf(void* keys, Stencil* stencil, void* output, ExtentType n, hipStream_t stream = nullptr) const
    {
        detail::Selection<Stencil, Predicate> s(stencil, Predicate{}, static_cast<SizeType>(n), stream);
        detail::check(mxco_mm_find(get(), keys, s.get(), output, static_cast<SizeType>(n), stream));
    }

    uint64_t Count(void* keys, ExtentType count, hipStream_t stream = nullptr) const
    {
        uint64_t c = 0;
        detail::check(mxco_mm_count(get(), keys, static_cast<SizeType>(count), stream, &c));
        return c;
    }

    // Every match grouped by query, the values of one query ascending.
    SizeType Retrieve(void* keys, ExtentType count, void* probeOut, void* valueOut, ExtentType outputCapacity,
                      hipStream_t stream = nullptr) const
    {
        uint64_t out = 0;
        detail::check(mxco_mm_retrieve(get(), keys, static_cast<SizeType>(count), probeOut, valueOut,
                                       static_cast<SizeType>(outputCapacity), stream, &out));
        return out;
    }

    // All stored pairs, duplicates included; the order is unspecified.
    SizeType RetrieveAll(void* keyOut, void* valueOut, ExtentType outputCapacity, hipStream_t stream = nullptr) const
    {
        uint64_t out = 0;
        detail::check(
            mxco_mm_retrieve_all(get(), keyOut, valueOut, static_cast<SizeType>(outputCapacity), stream, &out));
        return out;
    }

private:
    struct Deleter {
        void operator()(mxco_mm* h) const noexcept { (void)mxco_mm_destroy(h); }
    };

    mxco_mm* get() const
    {
        if (!h_) throw std::logic_error("StaticMultimap: moved-from object");
        return h_.get();
    }

    std::unique_ptr<mxco_mm, Deleter> h_;
};

}  // namespace mxco
