// mxco::BloomFilter -- drop-in for aclco::BloomFilter with the default
// policy (include/bloom_filter.h:37-133, include/bloom_filter_policy.h:54-181).
#pragma once

#include <memory>

#include "common.h"

namespace mxco {

// BloomFilterPolicy<Key>: xxhash_64<Key> with a seed, 8 words x 32 bits per
// block, one pattern bit per word.
template <class Key>
class BloomFilterPolicy {
public:
    using KeyType = Key;
    using WordType = std::uint32_t;
    static constexpr std::uint32_t wordsPerBlock = 8;
    constexpr explicit BloomFilterPolicy(std::uint64_t seed = 0) noexcept : seed_{seed} {}
    constexpr std::uint64_t Seed() const noexcept { return seed_; }
    constexpr bool operator==(BloomFilterPolicy const& o) const noexcept { return seed_ == o.seed_; }
    constexpr bool operator!=(BloomFilterPolicy const& o) const noexcept { return seed_ != o.seed_; }

private:
    std::uint64_t seed_;
};

template <class Key, class Extent = mxco::Extent<std::size_t>, class Policy = mxco::BloomFilterPolicy<Key>>
class BloomFilter {
    static_assert(std::is_same_v<Key, std::int32_t> || std::is_same_v<Key, std::uint32_t> ||
                      std::is_same_v<Key, std::int64_t> || std::is_same_v<Key, std::uint64_t> ||
                      std::is_same_v<Key, float>,
                  "BloomFilter supports int32_t, uint32_t, int64_t, uint64_t and float keys");

public:
    using KeyType = Key;
    using ExtentType = Extent;
    using SizeType = typename ExtentType::ValueType;
    using PolicyType = Policy;
    using WordType = std::uint32_t;
    static constexpr std::uint32_t wordsPerBlock = 8;

    BloomFilter(BloomFilter const&) = delete;
    BloomFilter& operator=(BloomFilter const&) = delete;
    BloomFilter(BloomFilter&&) noexcept = default;
    BloomFilter& operator=(BloomFilter&&) noexcept = default;

    explicit BloomFilter(ExtentType numBlocks, PolicyType const& policy = PolicyType{}, hipStream_t stream = nullptr)
        : numBlocks_{static_cast<SizeType>(numBlocks)}, policy_{policy}
    {
        mxco_bloom* h = nullptr;
        detail::check(mxco_bloom_create(numBlocks_, sizeof(Key), policy.Seed(), stream, &h));
        impl_.reset(h);
    }

    void Clear(hipStream_t stream = nullptr) { detail::check(mxco_bloom_clear(get(), stream, 1)); }
    void ClearAsync(hipStream_t stream = nullptr) { detail::check(mxco_bloom_clear(get(), stream, 0)); }
    void Add(void* keys, ExtentType n, hipStream_t stream = nullptr)
    {
        detail::check(mxco_bloom_add(get(), keys, nullptr, static_cast<SizeType>(n), stream, 1));
    }
    void AddAsync(void* keys, ExtentType n, hipStream_t stream = nullptr)
    {
        detail::check(mxco_bloom_add(get(), keys, nullptr, static_cast<SizeType>(n), stream, 0));
    }
    void AddIf(void* keys, void* stencil, ExtentType n, hipStream_t stream = nullptr)
    {
        if (static_cast<SizeType>(n) != 0 && stencil == nullptr)
            throw std::invalid_argument("BloomFilter::AddIf: stencil must not be null when keyNum is non-zero");
        detail::check(mxco_bloom_add(get(), keys, static_cast<const uint8_t*>(stencil), static_cast<SizeType>(n),
                                     stream, 1));
    }
    void AddIfAsync(void* keys, void* stencil, ExtentType n, hipStream_t stream = nullptr)
    {
        if (static_cast<SizeType>(n) != 0 && stencil == nullptr)
            throw std::invalid_argument("BloomFilter::AddIfAsync: stencil must not be null when keyNum is non-zero");
        detail::check(mxco_bloom_add(get(), keys, static_cast<const uint8_t*>(stencil), static_cast<SizeType>(n),
                                     stream, 0));
    }
    void Contains(void* keys, ExtentType n, void* out, hipStream_t stream = nullptr) const
    {
        detail::check(mxco_bloom_contains(get(), keys, nullptr, static_cast<uint8_t*>(out),
                                          static_cast<SizeType>(n), stream, 1));
    }
    void Contains(void* keys, void* out, ExtentType n, hipStream_t stream = nullptr) const
    {
        Contains(keys, n, out, stream);
    }
    void ContainsAsync(void* keys, ExtentType n, void* out, hipStream_t stream = nullptr) const
    {
        detail::check(mxco_bloom_contains(get(), keys, nullptr, static_cast<uint8_t*>(out),
                                          static_cast<SizeType>(n), stream, 0));
    }
    void ContainsAsync(void* keys, void* out, ExtentType n, hipStream_t stream = nullptr) const
    {
        ContainsAsync(keys, n, out, stream);
    }
    void ContainsIf(void* keys, void* stencil, void* out, ExtentType n, hipStream_t stream = nullptr) const
    {
        if (static_cast<SizeType>(n) != 0 && stencil == nullptr)
            throw std::invalid_argument("BloomFilter::ContainsIf: stencil must not be null when keyNum is non-zero");
        detail::check(mxco_bloom_contains(get(), keys, static_cast<const uint8_t*>(stencil),
                                          static_cast<uint8_t*>(out), static_cast<SizeType>(n), stream, 1));
    }
    void ContainsIfAsync(void* keys, void* stencil, void* out, ExtentType n, hipStream_t stream = nullptr) const
    {
        if (static_cast<SizeType>(n) != 0 && stencil == nullptr)
            throw std::invalid_argument(
                "BloomFilter::ContainsIfAsync: stencil must not be null when keyNum is non-zero");
        detail::check(mxco_bloom_contains(get(), keys, static_cast<const uint8_t*>(stencil),
                                          static_cast<uint8_t*>(out), static_cast<SizeType>(n), stream, 0));
    }
    void Merge(BloomFilter const& other, hipStream_t stream = nullptr)
    {
        detail::check(mxco_bloom_merge(get(), other.get(), stream, 1));
    }
    void MergeAsync(BloomFilter const& other, hipStream_t stream = nullptr)
    {
        detail::check(mxco_bloom_merge(get(), other.get(), stream, 0));
    }
    void Intersect(BloomFilter const& other, hipStream_t stream = nullptr)
    {
        detail::check(mxco_bloom_intersect(get(), other.get(), stream, 1));
    }
    void IntersectAsync(BloomFilter const& other, hipStream_t stream = nullptr)
    {
        detail::check(mxco_bloom_intersect(get(), other.get(), stream, 0));
    }
    WordType* Data() noexcept
    {
        uint32_t* w = nullptr;
        (void)mxco_bloom_data(get(), &w, nullptr);
        return w;
    }
    WordType const* Data() const noexcept
    {
        uint32_t* w = nullptr;
        (void)mxco_bloom_data(get(), &w, nullptr);
        return w;
    }
    ExtentType BlockExtent() const noexcept { return ExtentType{numBlocks_}; }
    SizeType NumWords() const noexcept { return numBlocks_ * wordsPerBlock; }
    std::size_t SizeBytes() const noexcept { return static_cast<std::size_t>(NumWords()) * sizeof(WordType); }
    PolicyType const& GetPolicy() const noexcept { return policy_; }

private:
    struct Deleter {
        void operator()(mxco_bloom* h) const noexcept { (void)mxco_bloom_destroy(h); }
    };
    mxco_bloom* get() const noexcept { return impl_.get(); }
    std::unique_ptr<mxco_bloom, Deleter> impl_;
    SizeType numBlocks_;
    PolicyType policy_;
};

}  // namespace mxco
