// mxco::BloomFilter -- drop-in for aclco::BloomFilter<Key, Extent, Policy,
// Allocator> (include/bloom_filter.h:37-133) with aclco::BloomFilterPolicy
// (include/bloom_filter_policy.h:54-211).
#pragma once

#include <limits>
#include <memory>
#include <utility>

#include "common.h"
#include "detail/bloom_salts.h"

namespace mxco {

// BloomFilterPolicy (include/bloom_filter_policy.h:54-168): a 64-bit key hash
// whose upper half picks one block of WordsPerBlock 32-bit words and whose
// lower half sets PatternBits / WordsPerBlock bits in every word of it.  The
// layout and early-exit parameters choose how the reference's kernels walk a
// block; they change no bit, and every value is accepted.
template <class Key, class Hash = mxco::xxhash_64<Key>, class Word = std::uint32_t,
          std::uint32_t WordsPerBlock = 32 / sizeof(Word), std::uint32_t PatternBits = WordsPerBlock,
          std::uint32_t AddHorizontalLayout = WordsPerBlock, std::uint32_t AddVerticalLayout = 1,
          std::uint32_t ContainsHorizontalLayout = 1, std::uint32_t ContainsVerticalLayout = WordsPerBlock,
          bool ConditionalAdd = false, bool EarlyExitContains = false>
class BloomFilterPolicy {
public:
    using KeyType = Key;
    using Hasher = Hash;
    using WordType = Word;
    using HashResultType = typename Hasher::ResultType;

    static constexpr std::uint32_t wordsPerBlock = WordsPerBlock;
    static constexpr std::uint32_t patternBits = PatternBits;
    static constexpr std::uint32_t addHorizontalLayout = AddHorizontalLayout;
    static constexpr std::uint32_t addVerticalLayout = AddVerticalLayout;
    static constexpr std::uint32_t containsHorizontalLayout = ContainsHorizontalLayout;
    static constexpr std::uint32_t containsVerticalLayout = ContainsVerticalLayout;
    static constexpr bool conditionalAdd = ConditionalAdd;
    static constexpr bool earlyExitContains = EarlyExitContains;
    static constexpr std::uint32_t wordBits = sizeof(WordType) * 8;
    static constexpr std::uint32_t bitsPerWord = patternBits / wordsPerBlock;
    static constexpr std::uint32_t maxFilterBlocks = std::numeric_limits<std::uint32_t>::max();

    static_assert(std::is_unsigned_v<WordType>, "BloomFilter word type must be unsigned");
    static_assert(sizeof(WordType) == 4, "BloomFilter supports uint32_t words only");
    static_assert(std::is_same_v<HashResultType, std::uint64_t>, "BloomFilter requires a 64-bit hash result");
    static_assert(wordsPerBlock != 0 && (wordsPerBlock & (wordsPerBlock - 1)) == 0,
                  "WordsPerBlock must be a non-zero power of two");
    static_assert(wordsPerBlock <= 32, "WordsPerBlock must not exceed 32");
    static_assert(patternBits >= wordsPerBlock, "PatternBits must provide at least one bit per word");
    static_assert(patternBits <= 64, "PatternBits exceeds the salt table");
    static_assert(patternBits <= wordBits * wordsPerBlock, "PatternBits exceeds the number of bits in a block");
    static_assert(patternBits % wordsPerBlock == 0, "PatternBits must be evenly distributed across all words");
    static_assert(addHorizontalLayout != 0 && addVerticalLayout != 0, "Add layout dimensions must be non-zero");
    static_assert(containsHorizontalLayout != 0 && containsVerticalLayout != 0,
                  "Contains layout dimensions must be non-zero");
    static_assert(wordsPerBlock % (addHorizontalLayout * addVerticalLayout) == 0,
                  "Add layout must evenly divide WordsPerBlock");
    static_assert(wordsPerBlock % (containsHorizontalLayout * containsVerticalLayout) == 0,
                  "Contains layout must evenly divide WordsPerBlock");

    constexpr BloomFilterPolicy(Hasher hash = Hasher{}) noexcept : hash_{hash} {}

    std::uint64_t HashKey(KeyType const& key) const noexcept { return hash_(key); }

    template <typename SizeType>
    constexpr std::uint32_t BlockIndex(std::uint32_t upperHash, SizeType numBlocks) const noexcept
    {
        return static_cast<std::uint32_t>((static_cast<std::uint64_t>(upperHash) *
                                           static_cast<std::uint32_t>(numBlocks)) >> 32);
    }

    template <std::uint32_t WordIndex>
    constexpr WordType WordPattern(std::uint32_t lowerHash) const noexcept
    {
        static_assert(WordIndex < wordsPerBlock, "word index out of range");
        return WordPatternForLane(WordIndex, lowerHash);
    }

    // @pre lane < wordsPerBlock
    constexpr WordType WordPatternForLane(std::uint32_t lane, std::uint32_t lowerHash) const noexcept
    {
        WordType m = 0;
        for (std::uint32_t b = 0; b < bitsPerWord; ++b)
            m |= WordType{1} << ((detail::kBloomSalts[lane * bitsPerWord + b] * lowerHash) >> 27);
        return m;
    }

    constexpr Hasher const& GetHasher() const noexcept { return hash_; }
    constexpr bool operator==(BloomFilterPolicy const& other) const noexcept { return hash_ == other.hash_; }
    constexpr bool operator!=(BloomFilterPolicy const& other) const noexcept { return !(*this == other); }

private:
    Hasher hash_;
};

namespace detail {
template <class Policy>
struct UsesMultiplyHighBlockIndex : std::false_type {};
template <class K, class H, class W, std::uint32_t WPB, std::uint32_t PB, std::uint32_t AH, std::uint32_t AV,
          std::uint32_t CH, std::uint32_t CV, bool CA, bool EE>
struct UsesMultiplyHighBlockIndex<BloomFilterPolicy<K, H, W, WPB, PB, AH, AV, CH, CV, CA, EE>> : std::true_type {};
}  // namespace detail

// DefaultFilterPolicy / ArrowFilterPolicy (include/bloom_filter_policy.h:190-211).
template <class Hash, class Word, std::uint32_t WordsPerBlock>
using DefaultFilterPolicy = BloomFilterPolicy<typename Hash::ArgumentType, Hash, Word, WordsPerBlock>;

template <class Key>
using ArrowFilterPolicy = BloomFilterPolicy<Key, mxco::xxhash_64<Key>, std::uint32_t, 8, 8, 8, 1, 1, 8, false, false>;

template <class Key, class Extent = mxco::Extent<std::size_t>, class Policy = mxco::BloomFilterPolicy<Key>,
          class Allocator = mxco::DefaultAllocator<typename Policy::WordType>>
class BloomFilter {
    static_assert(std::is_same_v<Key, std::int32_t> || std::is_same_v<Key, std::uint32_t> ||
                      std::is_same_v<Key, std::int64_t> || std::is_same_v<Key, std::uint64_t> ||
                      std::is_same_v<Key, float>,
                  "BloomFilter supports int32_t, uint32_t, int64_t, uint64_t and float keys");
    static_assert(std::is_same_v<typename Policy::KeyType, Key>, "BloomFilter policy key type must match Key");
    static_assert(std::is_same_v<typename Policy::Hasher, mxco::xxhash_64<Key>>,
                  "BloomFilter kernels support xxhash_64<Key> policies only");

public:
    using KeyType = Key;
    using ExtentType = Extent;
    using SizeType = typename ExtentType::ValueType;
    using PolicyType = Policy;
    using WordType = typename PolicyType::WordType;
    using AllocatorType = Allocator;
    static constexpr std::uint32_t wordsPerBlock = PolicyType::wordsPerBlock;

    static_assert(std::is_integral_v<SizeType> && std::is_unsigned_v<SizeType>,
                  "BloomFilter extent value type must be an unsigned integer");

    BloomFilter(BloomFilter const&) = delete;
    BloomFilter& operator=(BloomFilter const&) = delete;
    BloomFilter(BloomFilter&& o) noexcept
        : allocator_{std::move(o.allocator_)},
          words_{std::exchange(o.words_, nullptr)},
          impl_{std::move(o.impl_)},
          numBlocks_{std::exchange(o.numBlocks_, 0)},
          policy_{std::move(o.policy_)}
    {
    }
    BloomFilter& operator=(BloomFilter&& o) noexcept
    {
        if (this != &o) {
            Release();
            allocator_ = std::move(o.allocator_);
            words_ = std::exchange(o.words_, nullptr);
            impl_ = std::move(o.impl_);
            numBlocks_ = std::exchange(o.numBlocks_, 0);
            policy_ = std::move(o.policy_);
        }
        return *this;
    }

    // BloomFilter(numBlocks, policy, allocator, stream)
    // (bloom_filter_lifecycle.inl:15-55): the words come from the allocator.
    explicit BloomFilter(ExtentType numBlocks, PolicyType const& policy = PolicyType{},
                         AllocatorType const& allocator = AllocatorType{}, hipStream_t stream = nullptr)
        : allocator_{allocator}, numBlocks_{static_cast<SizeType>(numBlocks)}, policy_{policy}
    {
        if (numBlocks_ == 0) throw std::invalid_argument("BloomFilter: numBlocks must be greater than zero");
        if (static_cast<std::uint64_t>(numBlocks_) > PolicyType::maxFilterBlocks)
            throw std::length_error("BloomFilter: numBlocks exceeds the supported block-index range");
        if (static_cast<std::uint64_t>(numBlocks_) > std::numeric_limits<SizeType>::max() / wordsPerBlock)
            throw std::length_error("BloomFilter: word count overflows Extent::ValueType");
        words_ = allocator_.Allocate(static_cast<std::size_t>(NumWords()));
        if (words_ == nullptr) throw std::bad_alloc();
        mxco_bloom_policy p{PolicyType::wordsPerBlock, PolicyType::patternBits, policy.GetHasher().Seed()};
        mxco_bloom* h = nullptr;
        const mxco_status st = mxco_bloom_create_with_policy(numBlocks_, sizeof(Key), &p, words_, stream, &h);
        if (st != MXCO_OK) {
            allocator_.Deallocate(words_);
            words_ = nullptr;
            detail::check(st);
        }
        impl_.reset(h);
    }

    // BloomFilter(numBlocks, policy, stream) (bloom_filter_lifecycle.inl:57-62)
    explicit BloomFilter(ExtentType numBlocks, PolicyType const& policy, hipStream_t stream)
        : BloomFilter(numBlocks, policy, AllocatorType{}, stream)
    {
    }

    ~BloomFilter() { Release(); }

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
    WordType* Data() noexcept { return words_; }
    WordType const* Data() const noexcept { return words_; }
    ExtentType BlockExtent() const noexcept { return ExtentType{numBlocks_}; }
    SizeType NumWords() const noexcept { return numBlocks_ * wordsPerBlock; }
    std::size_t SizeBytes() const noexcept { return static_cast<std::size_t>(NumWords()) * sizeof(WordType); }
    PolicyType const& GetPolicy() const noexcept { return policy_; }

private:
    struct Deleter {
        void operator()(mxco_bloom* h) const noexcept { (void)mxco_bloom_destroy(h); }
    };
    mxco_bloom* get() const
    {
        if (!impl_) throw std::logic_error("BloomFilter: moved-from filter");
        return impl_.get();
    }
    void Release() noexcept
    {
        impl_.reset();  // the library leaves the words to their allocator
        if (words_) allocator_.Deallocate(words_);
        words_ = nullptr;
    }

    AllocatorType allocator_;
    WordType* words_{nullptr};
    std::unique_ptr<mxco_bloom, Deleter> impl_;
    SizeType numBlocks_;
    PolicyType policy_;
};

}  // namespace mxco
