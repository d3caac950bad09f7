// mxco::StaticMultiset -- drop-in for aclco::StaticMultiset
// (include/static_multiset.h:36-575).
//
// Same member functions and argument order (both Extent/pointer orders of
// the query calls); the stream argument is a hipStream_t.  Calls are
// synchronous on the stream, as in the reference.  Statuses are rethrown as
// the reference's exceptions: std::invalid_argument (zero capacity, a null
// pointer for a nonempty operation), std::length_error (Retrieve /
// RetrieveAll outputs too small, nothing written), std::overflow_error
// (Count / Retrieve when n * Size() may exceed 64 bits), std::logic_error
// (moved-from object).  Predicates are the built-in stencil predicates or,
// compiled with hipcc, any predicate object (stateless or not, as the
// reference's InsertIf / ContainsIf / FindIf take it by value).  CountEach /
// CountEachOuter with an explicit equality and hash require them to agree
// with exact key equality and the container's hasher, as the reference's
// do; they give the plain counts.
#pragma once

#include <memory>
#include <utility>

#include "common.h"
#if defined(__HIPCC__)
#include "detail/caller_probe.h"
#endif

namespace mxco {

template <typename Key>
class StaticMultiset {
    static_assert(std::is_same_v<Key, int32_t> || std::is_same_v<Key, int64_t>,
                  "StaticMultiset supports int32_t and int64_t keys");

public:
    using KeyType = Key;
    using SizeType = uint64_t;
    // the probe hash of the slots (include/static_multiset.h:45)
    using Hasher = std::conditional_t<sizeof(Key) == 4, murmurhash3_fmix32<Key>, murmurhash3_fmix64<Key>>;
    using Extent = mxco::Extent<std::size_t>;

    StaticMultiset(Extent capacity, Key emptyKey, hipStream_t stream = nullptr)
    {
        mxco_ms* h = nullptr;
        detail::check(mxco_ms_create(sizeof(Key), static_cast<std::size_t>(capacity), &emptyKey, stream, &h));
        h_.reset(h);
    }
    StaticMultiset(const StaticMultiset&) = delete;
    StaticMultiset& operator=(const StaticMultiset&) = delete;
    StaticMultiset(StaticMultiset&&) noexcept = default;
    StaticMultiset& operator=(StaticMultiset&&) noexcept = default;
    ~StaticMultiset() = default;

    SizeType Capacity() const noexcept
    {
        uint64_t c = 0;
        (void)mxco_ms_capacity(h_.get(), &c);
        return c;
    }
    SizeType Size(hipStream_t stream = nullptr) const
    {
        uint64_t s = 0;
        detail::check(mxco_ms_size(get(), stream, &s));
        return s;
    }
    void Clear(hipStream_t stream = nullptr) { detail::check(mxco_ms_clear(get(), stream)); }

    SizeType Insert(void* keys, Extent n, hipStream_t stream = nullptr)
    {
        uint64_t failed = 0;
        detail::check(mxco_ms_insert(get(), keys, nullptr, static_cast<std::size_t>(n), stream, &failed));
        return failed;
    }
    template <typename Stencil, typename Predicate>
    SizeType InsertIf(void* keys, Stencil* stencil, Extent n, hipStream_t stream = nullptr)
    {
        return InsertIf(keys, n, stencil, Predicate{}, stream);
    }
    template <typename Stencil, typename Predicate>
    SizeType InsertIf(void* keys, Extent n, Stencil* stencil, Predicate pred, hipStream_t stream = nullptr)
    {
        uint64_t failed = 0;
        if (static_cast<std::size_t>(n) && keys && !stencil)
            throw std::invalid_argument("StaticMultiset: null pointer for nonempty operation");
        detail::Selection<Stencil, Predicate> s(stencil, pred, keys ? static_cast<std::size_t>(n) : 0, stream);
        detail::check(mxco_ms_insert(get(), keys, s.get(), static_cast<std::size_t>(n), stream, &failed));
        return failed;
    }

    void Contains(void* keys, void* output, Extent n, hipStream_t stream = nullptr)
    {
        query(MXCO_MS_CONTAINS, keys, nullptr, output, n, stream);
    }
    void Contains(void* keys, Extent n, void* output, hipStream_t stream = nullptr) { Contains(keys, output, n, stream); }
    void Find(void* keys, void* output, Extent n, hipStream_t stream = nullptr)
    {
        query(MXCO_MS_FIND, keys, nullptr, output, n, stream);
    }
    void Find(void* keys, Extent n, void* output, hipStream_t stream = nullptr) { Find(keys, output, n, stream); }
    template <typename Stencil, typename Predicate>
    void ContainsIf(void* keys, Stencil* stencil, void* output, Extent n, hipStream_t stream = nullptr)
    {
        ContainsIf(keys, n, stencil, Predicate{}, output, stream);
    }
    template <typename Stencil, typename Predicate>
    void ContainsIf(void* keys, Extent n, Stencil* stencil, Predicate pred, void* output,
                    hipStream_t stream = nullptr)
    {
        detail::Selection<Stencil, Predicate> s = conditional(stencil, pred, n, stream);
        query(MXCO_MS_CONTAINS, keys, s.get(), output, n, stream);
    }
    template <typename Stencil, typename Predicate>
    void FindIf(void* keys, Stencil* stencil, void* output, Extent n, hipStream_t stream = nullptr)
    {
        FindIf(keys, n, stencil, Predicate{}, output, stream);
    }
    template <typename Stencil, typename Predicate>
    void FindIf(void* keys, Extent n, Stencil* stencil, Predicate pred, void* output, hipStream_t stream = nullptr)
    {
        detail::Selection<Stencil, Predicate> s = conditional(stencil, pred, n, stream);
        query(MXCO_MS_FIND, keys, s.get(), output, n, stream);
    }
    // uint64_t multiplicity per query (CountEachOuter writes 1 for absent keys).
    void CountEach(void* keys, void* output, Extent n, hipStream_t stream = nullptr)
    {
        query(MXCO_MS_COUNT_EACH, keys, nullptr, output, n, stream);
    }
    void CountEach(void* keys, Extent n, void* output, hipStream_t stream = nullptr) { CountEach(keys, output, n, stream); }
    void CountEachOuter(void* keys, void* output, Extent n, hipStream_t stream = nullptr)
    {
        query(MXCO_MS_COUNT_EACH_OUTER, keys, nullptr, output, n, stream);
    }
    void CountEachOuter(void* keys, Extent n, void* output, hipStream_t stream = nullptr)
    {
        CountEachOuter(keys, output, n, stream);
    }
    // With an equality and probe hash compatible with the construction policy
    // (include/static_multiset.h:264-288): the same counts.
    template <typename Equal, typename Hash>
    void CountEach(void* keys, Extent n, Equal, Hash, void* output, hipStream_t stream = nullptr)
    {
        CountEach(keys, output, n, stream);
    }
    template <typename Equal, typename Hash>
    void CountEachOuter(void* keys, Extent n, Equal, Hash, void* output, hipStream_t stream = nullptr)
    {
        CountEachOuter(keys, output, n, stream);
    }

    SizeType Count(void* keys, Extent n, hipStream_t stream = nullptr)
    {
        uint64_t c = 0;
        detail::check(mxco_ms_count(get(), keys, static_cast<std::size_t>(n), stream, &c));
        return c;
    }
    SizeType Retrieve(void* keys, Extent n, void* probeOutput, void* matchOutput, Extent outputCapacity,
                      hipStream_t stream = nullptr)
    {
        uint64_t c = 0;
        detail::check(mxco_ms_retrieve(get(), keys, static_cast<std::size_t>(n), probeOutput, matchOutput,
                                       static_cast<std::size_t>(outputCapacity), stream, &c));
        return c;
    }
    SizeType RetrieveAll(void* output, Extent outputCapacity, hipStream_t stream = nullptr)
    {
        uint64_t c = 0;
        detail::check(mxco_ms_retrieve_all(get(), output, static_cast<std::size_t>(outputCapacity), stream, &c));
        return c;
    }

private:
    struct Deleter {
        void operator()(mxco_ms* h) const noexcept { (void)mxco_ms_destroy(h); }
    };
    mxco_ms* get() const
    {
        if (!h_) throw std::logic_error("StaticMultiset: moved-from object");
        return h_.get();
    }
    template <typename Stencil, typename Predicate>
    static detail::Selection<Stencil, Predicate> conditional(Stencil* stencil, Predicate const& pred, Extent n,
                                                             hipStream_t stream)
    {
        if (static_cast<std::size_t>(n) && !stencil)
            throw std::invalid_argument("StaticMultiset: null pointer for nonempty operation");
        return detail::Selection<Stencil, Predicate>(stencil, pred, static_cast<std::size_t>(n), stream);
    }
    void query(int kind, void* keys, const mxco_stencil* s, void* output, Extent n, hipStream_t stream)
    {
        detail::check(mxco_ms_query(get(), kind, keys, s, output, static_cast<std::size_t>(n), stream));
    }

    std::unique_ptr<mxco_ms, Deleter> h_;
};

}  // namespace mxco
