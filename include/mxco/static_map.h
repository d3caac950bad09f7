// mxco::StaticMap -- drop-in for aclco::StaticMap (include/static_map.h:29-614).
//
// Same template parameters, same member functions and argument order; the
// stream argument is a hipStream_t.  Every call forwards to the C ABI of
// libmxco.so (include/mxco.h) and rethrows its status as the reference's
// exception type.  A caller-defined KeyEqual, hasher or predicate type runs
// in probe kernels compiled in the caller's translation unit
// (detail/caller_probe.h), which the library calls; that needs hipcc.
#pragma once

#include <memory>
#include <utility>

#include "common.h"
#if defined(__HIPCC__)
#include "detail/caller_probe.h"
#include "for_each.h"
#endif

namespace mxco {

static constexpr std::size_t defaultMapBucketSize = 5;

template <class Key, class T, class Extent = mxco::Extent<std::size_t>, class KeyEqual = mxco::EqualTo<Key>,
          class ProbingScheme = mxco::LinearProbing<mxco::murmurhash3_32<Key>>,
          class Storage = mxco::Storage<defaultMapBucketSize>>
class StaticMap {
    static_assert(sizeof(Key) == 1 || sizeof(Key) == 2 || sizeof(Key) == 4 || sizeof(Key) == 8,
                  "keys must be 1, 2, 4 or 8 bytes");
    static_assert(sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8,
                  "payloads must be 1, 2, 4 or 8 bytes");
    static constexpr bool kCompiledProbe = detail::kCompiledProbe<Key, KeyEqual, ProbingScheme>;
#if !defined(__HIPCC__)
    static_assert(kCompiledProbe,
                  "a caller-defined KeyEqual or hasher runs in kernels compiled with it: use hipcc for this unit");
#endif

public:
    static constexpr auto bucketSize = Storage::bucketSize;
    using SizeType = typename Extent::ValueType;
    using KeyType = Key;
    using MappedType = T;
    using ValueType = Pair<Key, T>;

    StaticMap(StaticMap const&) = delete;
    StaticMap& operator=(StaticMap const&) = delete;
    StaticMap(StaticMap&&) = default;
    StaticMap& operator=(StaticMap&&) = default;
    ~StaticMap() = default;

    StaticMap(Extent capacity, Key emptyKey, T emptyValue, KeyEqual const& pred = {},
              ProbingScheme const& probing = {}, Storage = {}, hipStream_t stream = nullptr)
        : emptyKey_{emptyKey}, emptyValue_{emptyValue}
    {
        if (!detail::equals_default(pred, 0) || !detail::equals_default(probing, 0))
            throw std::invalid_argument("StaticMap: stateful KeyEqual or probing scheme");
        mxco_oa_desc d{};
        d.key_size = sizeof(Key);
        d.value_size = sizeof(T);
        d.key_kind = IsFloatKey<Key>::value ? MXCO_KEY_FLOAT : MXCO_KEY_INTEGRAL;
        d.bucket_size = static_cast<uint32_t>(bucketSize);
        d.probing = ProbingScheme::id;
        d.hash = ProbingScheme::hash;
        d.value_kind = IsFloatKey<T>::value ? MXCO_KEY_FLOAT : MXCO_KEY_INTEGRAL;
        mxco_oa* h = nullptr;
        detail::check(mxco_oa_create(&d, static_cast<SizeType>(capacity), &emptyKey, &emptyValue, stream, &h));
        impl_.reset(h);
#if defined(__HIPCC__)
        if constexpr (!kCompiledProbe)
            detail::check(mxco_oa_set_probe_launcher(
                h, &detail::caller_launch<Key, T, typename detail::CallerFnsOf<Key, KeyEqual, ProbingScheme>::type>));
#endif
    }

    void SetErasedKey(Key erasedKey, hipStream_t stream) { detail::check(mxco_oa_set_erased_key(get(), &erasedKey, stream)); }

    SizeType Insert(void* values, Extent valueNum, hipStream_t stream)
    {
        uint64_t failed = 0;
        detail::check(mxco_oa_insert(get(), values, nullptr, static_cast<SizeType>(valueNum), stream, 1, &failed));
        return static_cast<SizeType>(failed);
    }
    void InsertAsync(void* values, Extent valueNum, hipStream_t stream)
    {
        detail::check(mxco_oa_insert(get(), values, nullptr, static_cast<SizeType>(valueNum), stream, 0, nullptr));
    }
    SizeType InsertOrAssign(void* values, Extent valueNum, hipStream_t stream)
    {
        SizeType assigned = 0;
        return InsertOrAssign(values, valueNum, stream, assigned);
    }
    SizeType InsertOrAssign(void* values, Extent valueNum, hipStream_t stream, SizeType& assignedNum)
    {
        uint64_t f = 0, a = 0;
        detail::check(mxco_oa_insert_or_assign(get(), values, static_cast<SizeType>(valueNum), stream, 1, &f, &a));
        assignedNum = static_cast<SizeType>(a);
        return static_cast<SizeType>(f);
    }
    void InsertOrAssignAsync(void* values, Extent valueNum, hipStream_t stream)
    {
        detail::check(
            mxco_oa_insert_or_assign(get(), values, static_cast<SizeType>(valueNum), stream, 0, nullptr, nullptr));
    }
    template <typename StencilT, typename Predicate>
    void AssignIfAsync(void* values, StencilT* stencil, Extent valueNum, hipStream_t stream)
    {
        detail::Selection<StencilT, Predicate> s(stencil, Predicate{}, static_cast<SizeType>(valueNum), stream);
        detail::check(mxco_oa_assign_if(get(), values, s.get(), static_cast<SizeType>(valueNum), stream, 0));
    }
    template <typename StencilT, typename Predicate>
    SizeType InsertIf(void* values, StencilT* stencil, Extent valueNum, hipStream_t stream)
    {
        if (static_cast<SizeType>(valueNum) != 0 && stencil == nullptr) return static_cast<SizeType>(valueNum);
        detail::Selection<StencilT, Predicate> s(stencil, Predicate{}, static_cast<SizeType>(valueNum), stream);
        uint64_t failed = 0;
        detail::check(mxco_oa_insert(get(), values, s.get(), static_cast<SizeType>(valueNum), stream, 1, &failed));
        return static_cast<SizeType>(failed);
    }
    template <typename StencilT, typename Predicate>
    void InsertIfAsync(void* values, StencilT* stencil, Extent valueNum, hipStream_t stream)
    {
        if (stencil == nullptr) return;
        detail::Selection<StencilT, Predicate> s(stencil, Predicate{}, static_cast<SizeType>(valueNum), stream);
        detail::check(mxco_oa_insert(get(), values, s.get(), static_cast<SizeType>(valueNum), stream, 0, nullptr));
    }
    void InsertAndFind(void* values, void* outputFind, void* outputInsert, Extent valueNum, hipStream_t stream)
    {
        detail::check(mxco_oa_insert_and_find(get(), values, outputFind, static_cast<uint8_t*>(outputInsert),
                                              static_cast<SizeType>(valueNum), stream, 1));
    }
    void InsertAndFindAsync(void* values, void* outputFind, void* outputInsert, Extent valueNum, hipStream_t stream)
    {
        detail::check(mxco_oa_insert_and_find(get(), values, outputFind, static_cast<uint8_t*>(outputInsert),
                                              static_cast<SizeType>(valueNum), stream, 0));
    }
    SizeType Erase(void* keys, Extent keyNum, hipStream_t stream)
    {
        uint64_t failed = 0;
        detail::check(mxco_oa_erase(get(), keys, static_cast<SizeType>(keyNum), stream, 1, &failed));
        return static_cast<SizeType>(failed);
    }
    void EraseAsync(void* keys, Extent keyNum, hipStream_t stream)
    {
        detail::check(mxco_oa_erase(get(), keys, static_cast<SizeType>(keyNum), stream, 0, nullptr));
    }
    void Clear(hipStream_t stream) { detail::check(mxco_oa_clear(get(), stream, 1)); }
    void ClearAsync(hipStream_t stream) noexcept { (void)mxco_oa_clear(get(), stream, 0); }
    void Find(void* keys, void* outputValues, Extent keyNum, hipStream_t stream)
    {
        detail::check(mxco_oa_find(get(), keys, nullptr, outputValues, static_cast<SizeType>(keyNum), stream, 1));
    }
    void FindAsync(void* keys, void* outputValues, Extent keyNum, hipStream_t stream)
    {
        detail::check(mxco_oa_find(get(), keys, nullptr, outputValues, static_cast<SizeType>(keyNum), stream, 0));
    }
    template <typename StencilT, typename Predicate>
    void FindIf(void* keys, StencilT* stencil, void* outputValues, Extent keyNum, hipStream_t stream)
    {
        if (stencil == nullptr) return;
        detail::Selection<StencilT, Predicate> s(stencil, Predicate{}, static_cast<SizeType>(keyNum), stream);
        detail::check(mxco_oa_find(get(), keys, s.get(), outputValues, static_cast<SizeType>(keyNum), stream, 1));
    }
    template <typename StencilT, typename Predicate>
    void FindIfAsync(void* keys, StencilT* stencil, void* outputValues, Extent keyNum, hipStream_t stream)
    {
        if (stencil == nullptr) return;
        detail::Selection<StencilT, Predicate> s(stencil, Predicate{}, static_cast<SizeType>(keyNum), stream);
        detail::check(mxco_oa_find(get(), keys, s.get(), outputValues, static_cast<SizeType>(keyNum), stream, 0));
    }
    void Contains(void* keys, void* outputValues, Extent keyNum, hipStream_t stream)
    {
        detail::check(mxco_oa_contains(get(), keys, nullptr, static_cast<uint8_t*>(outputValues),
                                       static_cast<SizeType>(keyNum), stream, 1));
    }
    void ContainsAsync(void* keys, void* outputValues, Extent keyNum, hipStream_t stream)
    {
        detail::check(mxco_oa_contains(get(), keys, nullptr, static_cast<uint8_t*>(outputValues),
                                       static_cast<SizeType>(keyNum), stream, 0));
    }
    template <typename StencilT, typename Predicate>
    void ContainsIf(void* keys, StencilT* stencil, void* outputValues, Extent keyNum, hipStream_t stream)
    {
        if (stencil == nullptr) return;
        detail::Selection<StencilT, Predicate> s(stencil, Predicate{}, static_cast<SizeType>(keyNum), stream);
        detail::check(mxco_oa_contains(get(), keys, s.get(), static_cast<uint8_t*>(outputValues),
                                       static_cast<SizeType>(keyNum), stream, 1));
    }
    template <typename StencilT, typename Predicate>
    void ContainsIfAsync(void* keys, StencilT* stencil, void* outputValues, Extent keyNum, hipStream_t stream)
    {
        if (stencil == nullptr) return;
        detail::Selection<StencilT, Predicate> s(stencil, Predicate{}, static_cast<SizeType>(keyNum), stream);
        detail::check(mxco_oa_contains(get(), keys, s.get(), static_cast<uint8_t*>(outputValues),
                                       static_cast<SizeType>(keyNum), stream, 0));
    }
    SizeType Count(void* keys, Extent keyNum, hipStream_t stream)
    {
        uint64_t c = 0;
        detail::check(mxco_oa_count(get(), keys, static_cast<SizeType>(keyNum), stream, &c));
        return static_cast<SizeType>(c);
    }
    auto Capacity() const noexcept
    {
        uint64_t c = 0;
        (void)mxco_oa_capacity(get(), &c);
        return static_cast<SizeType>(c);
    }
    KeyEqual GetKeyEqual() const noexcept { return {}; }
    ValueType* Data() const
    {
        void* p = nullptr;
        detail::check(mxco_oa_data(get(), &p));
        return static_cast<ValueType*>(p);
    }
    Key EmptyKeySentinel() const noexcept { return emptyKey_; }

#if defined(__HIPCC__)
    // ForEach / ForEachAsync (the callback is device code: callers compile with hipcc).
    template <typename CallbackOp>
    void ForEachAsync(void* keys, Extent keyNum, void* callbackArgs, hipStream_t stream)
    {
        using Fns = std::conditional_t<kCompiledProbe, mxco::BuiltinFns,
                                       typename detail::CallerFnsOf<Key, KeyEqual, ProbingScheme>::type>;
        detail::for_each_launch<Key, ValueType, CallbackOp, Fns>(Data(), Capacity(), static_cast<uint32_t>(bucketSize),
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
    T EmptyValueSentinel() const noexcept { return emptyValue_; }

private:
    struct Deleter {
        void operator()(mxco_oa* h) const noexcept { (void)mxco_oa_destroy(h); }
    };
    mxco_oa* get() const noexcept { return impl_.get(); }

    std::unique_ptr<mxco_oa, Deleter> impl_;
    Key emptyKey_;
    T emptyValue_;
};

}  // namespace mxco
