// mxco::StaticSet -- drop-in for aclco::StaticSet (include/static_set.h:27-525).
// Defaults: DoubleHashing<xxhash_32<Key>>, Storage<5>.  Find returns the
// stored key or the empty key.  A caller-defined KeyEqual, hasher or
// predicate type runs in kernels compiled in the caller's translation unit
// (detail/caller_probe.h, hipcc).
#pragma once

#include <memory>

#include "common.h"
#if defined(__HIPCC__)
#include "detail/caller_probe.h"
#include "for_each.h"
#endif

namespace mxco {

static constexpr std::size_t defaultSetBucketSize = 5;

template <class Key, class Extent = mxco::Extent<std::size_t>, class KeyEqual = mxco::EqualTo<Key>,
          class ProbingScheme = mxco::DoubleHashing<mxco::xxhash_32<Key>>,
          class Storage = mxco::Storage<defaultSetBucketSize>>
class StaticSet {
    static_assert(sizeof(Key) == 1 || sizeof(Key) == 2 || sizeof(Key) == 4 || sizeof(Key) == 8,
                  "keys must be 1, 2, 4 or 8 bytes");
    static constexpr bool kCompiledProbe = detail::kCompiledProbe<Key, KeyEqual, ProbingScheme>;
#if !defined(__HIPCC__)
    static_assert(kCompiledProbe,
                  "a caller-defined KeyEqual or hasher runs in kernels compiled with it: use hipcc for this unit");
#endif

public:
    static constexpr auto bucketSize = Storage::bucketSize;
    using SizeType = typename Extent::ValueType;
    using KeyType = Key;
    using ValueType = Key;

    StaticSet(StaticSet const&) = delete;
    StaticSet& operator=(StaticSet const&) = delete;
    StaticSet(StaticSet&&) = default;
    StaticSet& operator=(StaticSet&&) = default;

    StaticSet(Extent capacity, Key emptyKey, KeyEqual const& pred = {}, ProbingScheme const& probing = {},
              Storage = {}, hipStream_t stream = nullptr)
        : emptyKey_{emptyKey}
    {
        if (!detail::equals_default(pred, 0) || !detail::equals_default(probing, 0))
            throw std::invalid_argument("StaticSet: stateful KeyEqual or probing scheme");
        mxco_oa_desc d{};
        d.key_size = sizeof(Key);
        d.value_size = 0;
        d.key_kind = IsFloatKey<Key>::value ? MXCO_KEY_FLOAT : MXCO_KEY_INTEGRAL;
        d.bucket_size = static_cast<uint32_t>(bucketSize);
        d.probing = ProbingScheme::id;
        d.hash = ProbingScheme::hash;
        mxco_oa* h = nullptr;
        detail::check(mxco_oa_create(&d, static_cast<SizeType>(capacity), &emptyKey, nullptr, stream, &h));
        impl_.reset(h);
#if defined(__HIPCC__)
        if constexpr (!kCompiledProbe)
            detail::check(mxco_oa_set_probe_launcher(
                h, &detail::caller_launch<Key, void, typename detail::CallerFnsOf<Key, KeyEqual, ProbingScheme>::type>));
#endif
    }

    void SetErasedKey(Key erasedKey, hipStream_t stream) { detail::check(mxco_oa_set_erased_key(get(), &erasedKey, stream)); }
    SizeType Insert(void* keys, Extent n, hipStream_t stream)
    {
        uint64_t f = 0;
        detail::check(mxco_oa_insert(get(), keys, nullptr, static_cast<SizeType>(n), stream, 1, &f));
        return static_cast<SizeType>(f);
    }
    void InsertAsync(void* keys, Extent n, hipStream_t stream)
    {
        detail::check(mxco_oa_insert(get(), keys, nullptr, static_cast<SizeType>(n), stream, 0, nullptr));
    }
    template <typename StencilT, typename Predicate>
    SizeType InsertIf(void* keys, StencilT* stencil, Extent n, hipStream_t stream)
    {
        if (static_cast<SizeType>(n) != 0 && stencil == nullptr) return static_cast<SizeType>(n);
        detail::Selection<StencilT, Predicate> s(stencil, Predicate{}, static_cast<SizeType>(n), stream);
        uint64_t f = 0;
        detail::check(mxco_oa_insert(get(), keys, s.get(), static_cast<SizeType>(n), stream, 1, &f));
        return static_cast<SizeType>(f);
    }
    template <typename StencilT, typename Predicate>
    void InsertIfAsync(void* keys, StencilT* stencil, Extent n, hipStream_t stream)
    {
        if (stencil == nullptr) return;
        detail::Selection<StencilT, Predicate> s(stencil, Predicate{}, static_cast<SizeType>(n), stream);
        detail::check(mxco_oa_insert(get(), keys, s.get(), static_cast<SizeType>(n), stream, 0, nullptr));
    }
    void InsertAndFind(void* keys, void* outputFind, void* outputInsert, Extent n, hipStream_t stream)
    {
        detail::check(mxco_oa_insert_and_find(get(), keys, outputFind, static_cast<uint8_t*>(outputInsert),
                                              static_cast<SizeType>(n), stream, 1));
    }
    void InsertAndFindAsync(void* keys, void* outputFind, void* outputInsert, Extent n, hipStream_t stream)
    {
        detail::check(mxco_oa_insert_and_find(get(), keys, outputFind, static_cast<uint8_t*>(outputInsert),
                                              static_cast<SizeType>(n), stream, 0));
    }
    SizeType Erase(void* keys, Extent n, hipStream_t stream)
    {
        uint64_t f = 0;
        detail::check(mxco_oa_erase(get(), keys, static_cast<SizeType>(n), stream, 1, &f));
        return static_cast<SizeType>(f);
    }
    void EraseAsync(void* keys, Extent n, hipStream_t stream)
    {
        detail::check(mxco_oa_erase(get(), keys, static_cast<SizeType>(n), stream, 0, nullptr));
    }
    void Clear(hipStream_t stream) { detail::check(mxco_oa_clear(get(), stream, 1)); }
    void ClearAsync(hipStream_t stream) noexcept { (void)mxco_oa_clear(get(), stream, 0); }
    void Find(void* keys, void* out, Extent n, hipStream_t stream)
    {
        detail::check(mxco_oa_find(get(), keys, nullptr, out, static_cast<SizeType>(n), stream, 1));
    }
    void FindAsync(void* keys, void* out, Extent n, hipStream_t stream)
    {
        detail::check(mxco_oa_find(get(), keys, nullptr, out, static_cast<SizeType>(n), stream, 0));
    }
    template <typename StencilT, typename Predicate>
    void FindIf(void* keys, StencilT* stencil, void* out, Extent n, hipStream_t stream)
    {
        if (stencil == nullptr) return;
        detail::Selection<StencilT, Predicate> s(stencil, Predicate{}, static_cast<SizeType>(n), stream);
        detail::check(mxco_oa_find(get(), keys, s.get(), out, static_cast<SizeType>(n), stream, 1));
    }
    void Contains(void* keys, void* out, Extent n, hipStream_t stream)
    {
        detail::check(
            mxco_oa_contains(get(), keys, nullptr, static_cast<uint8_t*>(out), static_cast<SizeType>(n), stream, 1));
    }
    void ContainsAsync(void* keys, void* out, Extent n, hipStream_t stream)
    {
        detail::check(
            mxco_oa_contains(get(), keys, nullptr, static_cast<uint8_t*>(out), static_cast<SizeType>(n), stream, 0));
    }
    template <typename StencilT, typename Predicate>
    void ContainsIf(void* keys, StencilT* stencil, void* out, Extent n, hipStream_t stream)
    {
        if (stencil == nullptr) return;
        detail::Selection<StencilT, Predicate> s(stencil, Predicate{}, static_cast<SizeType>(n), stream);
        detail::check(
            mxco_oa_contains(get(), keys, s.get(), static_cast<uint8_t*>(out), static_cast<SizeType>(n), stream, 1));
    }
    SizeType Count(void* keys, Extent n, hipStream_t stream)
    {
        uint64_t c = 0;
        detail::check(mxco_oa_count(get(), keys, static_cast<SizeType>(n), stream, &c));
        return static_cast<SizeType>(c);
    }
    auto Capacity() const noexcept
    {
        uint64_t c = 0;
        (void)mxco_oa_capacity(get(), &c);
        return static_cast<SizeType>(c);
    }
    Key* Data() const
    {
        void* p = nullptr;
        detail::check(mxco_oa_data(get(), &p));
        return static_cast<Key*>(p);
    }
    Key EmptyKeySentinel() const noexcept { return emptyKey_; }

#if defined(__HIPCC__)
    // ForEach / ForEachAsync (the callback is device code: callers compile with hipcc).
    template <typename CallbackOp>
    void ForEachAsync(void* keys, Extent keyNum, void* callbackArgs, hipStream_t stream)
    {
        using Fns = std::conditional_t<kCompiledProbe, mxco::BuiltinFns,
                                       typename detail::CallerFnsOf<Key, KeyEqual, ProbingScheme>::type>;
        detail::for_each_launch<Key, Key, CallbackOp, Fns>(Data(), Capacity(), static_cast<uint32_t>(bucketSize),
                                                          ProbingScheme::id, ProbingScheme::hash, keys,
                                                          static_cast<SizeType>(keyNum), emptyKey_, callbackArgs,
                                                          stream);
    }
    template <typename CallbackOp>
    void ForEach(void* keys, Extent keyNum, void* callbackArgs, hipStream_t stream)
    {
        ForEachAsync<CallbackOp>(keys, keyNum, callbackArgs, stream);
        if (hipStreamSynchronize(stream) != hipSuccess) throw std::runtime_error("ForEach: synchronize failed");
    }
#endif

private:
    struct Deleter {
        void operator()(mxco_oa* h) const noexcept { (void)mxco_oa_destroy(h); }
    };
    mxco_oa* get() const noexcept { return impl_.get(); }
    std::unique_ptr<mxco_oa, Deleter> impl_;
    Key emptyKey_;
};

}  // namespace mxco
