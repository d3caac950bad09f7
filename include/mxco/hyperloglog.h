// mxco::HyperLogLog -- drop-in for aclco::HyperLogLog (include/hyperloglog.h:24-59).
#pragma once

#include <memory>

#include "common.h"

namespace mxco {

template <class Key>
class HyperLogLog {
    static_assert(std::is_same_v<Key, std::int32_t> || std::is_same_v<Key, std::int64_t>,
                  "HyperLogLog supports int32_t and int64_t");

public:
    using KeyType = Key;
    using ExtentType = mxco::Extent<std::size_t>;

    static HyperLogLog CreateWithSketchSizeKB(std::uint32_t kb, hipStream_t stream = nullptr)
    {
        std::uint32_t p = 0;
        detail::check(mxco_hll_precision_for_sketch_kb(kb, &p));
        return HyperLogLog(p, stream);
    }
    static HyperLogLog CreateWithStandardDeviation(double deviation, hipStream_t stream = nullptr)
    {
        std::uint32_t p = 0;
        detail::check(mxco_hll_precision_for_stddev(deviation, &p));
        return HyperLogLog(p, stream);
    }
    static HyperLogLog CreateWithPrecision(std::uint32_t precision, hipStream_t stream = nullptr)
    {
        return HyperLogLog(precision, stream);
    }

    HyperLogLog(HyperLogLog const&) = delete;
    HyperLogLog& operator=(HyperLogLog const&) = delete;
    HyperLogLog(HyperLogLog&&) noexcept = default;
    HyperLogLog& operator=(HyperLogLog&&) noexcept = default;

    std::uint32_t SketchSizeKB() const noexcept { return impl_ ? (1u << precision_) / 1024u : 0u; }
    void Clear(hipStream_t stream = nullptr) { detail::check(mxco_hll_clear(get(), stream)); }
    void Add(void* keys, ExtentType count, hipStream_t stream = nullptr)
    {
        detail::check(mxco_hll_add(get(), keys, static_cast<std::size_t>(count), stream));
    }
    void Merge(HyperLogLog const& other, hipStream_t stream = nullptr)
    {
        detail::check(mxco_hll_merge(get(), other.get(), stream));
    }
    std::uint64_t Estimate(hipStream_t stream = nullptr) const
    {
        std::uint64_t e = 0;
        detail::check(mxco_hll_estimate(get(), stream, &e));
        return e;
    }
    std::uint8_t const* Data() const noexcept
    {
        const std::uint8_t* r = nullptr;
        (void)mxco_hll_data(get(), &r, nullptr);
        return r;
    }

private:
    explicit HyperLogLog(std::uint32_t precision, hipStream_t stream) : precision_{precision}
    {
        mxco_hll* h = nullptr;
        detail::check(mxco_hll_create(precision, sizeof(Key), stream, &h));
        impl_.reset(h);
    }
    struct Deleter {
        void operator()(mxco_hll* h) const noexcept { (void)mxco_hll_destroy(h); }
    };
    mxco_hll* get() const
    {
        if (!impl_) throw std::logic_error("HyperLogLog: moved-from or uninitialized sketch");
        return impl_.get();
    }
    std::unique_ptr<mxco_hll, Deleter> impl_;
    std::uint32_t precision_;
};

}  // namespace mxco
