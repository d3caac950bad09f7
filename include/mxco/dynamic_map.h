// mxco::DynamicMap -- drop-in for aclco::DynamicMap (include/dynamic_map.h:26-625).
//
// Same template parameters and member functions; the stream argument is a
// hipStream_t.  The growth, deduplication and merge logic runs inside
// libmxco.so (mxco_dm_*), over StaticMap submaps of the same configuration.
#pragma once

#include <memory>
#include <utility>

#include "common.h"
#include "static_map.h"

namespace mxco {

template <class Key, class T, class Extent = mxco::Extent<std::size_t>, class KeyEqual = mxco::EqualTo<Key>,
          class ProbingScheme = mxco::LinearProbing<mxco::murmurhash3_32<Key>>,
          class Storage = mxco::Storage<defaultMapBucketSize>>
class DynamicMap {
public:
    using MapType = StaticMap<Key, T, Extent, KeyEqual, ProbingScheme, Storage>;
    using SizeType = typename MapType::SizeType;
    using KeyType = Key;
    using ValueType = T;

    enum class InsertMode { kExactDedup, kAppendUnique };

    static constexpr SizeType DEFAULT_MIN_INSERT = 10000;
    static constexpr float DEFAULT_MAX_LOAD = 0.60f;
    static constexpr SizeType GROWTH_FACTOR = 2;
    static constexpr SizeType INSERT_SUBBATCH = 800000;

    DynamicMap(const DynamicMap&) = delete;
    DynamicMap& operator=(const DynamicMap&) = delete;
    DynamicMap(DynamicMap&&) = default;
    DynamicMap& operator=(DynamicMap&&) = default;
    ~DynamicMap() = default;

    DynamicMap(Extent initialCapacity, Key emptyKey, T emptyValue, Key erasedKey, KeyEqual const& = {},
               ProbingScheme const& = {}, Storage = {}, hipStream_t stream = nullptr)
    {
        mxco_oa_desc d{};
        d.key_size = sizeof(Key);
        d.value_size = sizeof(T);
        d.key_kind = std::is_floating_point_v<Key> ? MXCO_KEY_FLOAT : MXCO_KEY_INTEGRAL;
        d.bucket_size = static_cast<uint32_t>(Storage::bucketSize);
        d.probing = ProbingScheme::id;
        d.hash = ProbingScheme::hash;
        d.value_kind = std::is_floating_point_v<T> ? MXCO_KEY_FLOAT : MXCO_KEY_INTEGRAL;
        mxco_dm* h = nullptr;
        detail::check(mxco_dm_create(&d, static_cast<SizeType>(initialCapacity), &emptyKey, &emptyValue, &erasedKey,
                                     stream, &h));
        h_.reset(h);
    }

    SizeType Size() const noexcept { return info(0); }
    SizeType Capacity() const noexcept { return info(1); }
    size_t NumSubmaps() const noexcept { return static_cast<size_t>(info(2)); }

    void Reserve(SizeType n, hipStream_t stream) { detail::check(mxco_dm_reserve(get(), n, stream)); }

    SizeType Insert(void* values, Extent valueNum, hipStream_t stream)
    {
        return Insert(values, valueNum, stream, InsertMode::kExactDedup);
    }
    SizeType Insert(void* values, Extent valueNum, hipStream_t stream, InsertMode mode)
    {
        uint64_t ins = 0;
        detail::check(mxco_dm_insert(get(), values, static_cast<SizeType>(valueNum),
                                     mode == InsertMode::kAppendUnique ? MXCO_DM_APPEND_UNIQUE : MXCO_DM_EXACT_DEDUP, 1,
                                     stream, &ins));
        return ins;
    }
    SizeType InsertAsync(void* values, Extent valueNum, hipStream_t stream)
    {
        uint64_t ins = 0;
        detail::check(mxco_dm_insert(get(), values, static_cast<SizeType>(valueNum), MXCO_DM_EXACT_DEDUP, 0, stream,
                                     &ins));
        return ins;
    }
    SizeType InsertOrAssign(void* values, Extent valueNum, hipStream_t stream)
    {
        uint64_t ins = 0;
        detail::check(mxco_dm_insert_or_assign(get(), values, static_cast<SizeType>(valueNum), stream, &ins));
        return ins;
    }
    void Find(void* keys, void* outputValues, Extent keyNum, hipStream_t stream)
    {
        detail::check(mxco_dm_find(get(), keys, outputValues, static_cast<SizeType>(keyNum), stream));
    }
    void Contains(void* keys, void* outputValues, Extent keyNum, hipStream_t stream)
    {
        detail::check(mxco_dm_contains(get(), keys, static_cast<uint8_t*>(outputValues),
                                       static_cast<SizeType>(keyNum), stream));
    }
    SizeType Erase(void* keys, Extent keyNum, hipStream_t stream)
    {
        uint64_t e = 0;
        detail::check(mxco_dm_erase(get(), keys, static_cast<SizeType>(keyNum), 1, stream, &e));
        return e;
    }
    void EraseAsync(void* keys, Extent keyNum, hipStream_t stream)
    {
        detail::check(mxco_dm_erase(get(), keys, static_cast<SizeType>(keyNum), 0, stream, nullptr));
    }
    void Clear(hipStream_t stream) { detail::check(mxco_dm_clear(get(), stream)); }

private:
    struct Deleter {
        void operator()(mxco_dm* h) const noexcept { (void)mxco_dm_destroy(h); }
    };
    mxco_dm* get() const
    {
        if (!h_) throw std::logic_error("DynamicMap: moved-from object");
        return h_.get();
    }
    SizeType info(int which) const noexcept
    {
        uint64_t v[3] = {0, 0, 0};
        if (h_) (void)mxco_dm_info(h_.get(), &v[0], &v[1], &v[2]);
        return v[which];
    }

    std::unique_ptr<mxco_dm, Deleter> h_;
};

}  // namespace mxco
