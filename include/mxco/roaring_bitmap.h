// mxco::RoaringBitmap -- drop-in for aclco::RoaringBitmap
// (include/roaring_bitmap.h:24-143) and aclco::RoaringBitmapRef
// (include/roaring_bitmap_ref.h:166-265): membership queries against a
// bitmap in the Roaring portable serialization.  The stream argument is a
// hipStream_t; malformed input throws std::invalid_argument.  Device memory
// comes from libmxco (hipMalloc): the Allocator parameter is accepted for
// source compatibility and only DefaultAllocator<uint8_t> is offered.
// RoaringBitmapRef::Contains is device code and needs the HIP compiler.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>

#include "common.h"
#include "detail/roaring.h"

namespace mxco {

template <typename Key>
class RoaringBitmapRef {
    static_assert(std::is_same_v<Key, uint32_t> || std::is_same_v<Key, uint64_t>,
                  "RoaringBitmap key type must be uint32_t or uint64_t");

public:
    using KeyType = Key;

    explicit RoaringBitmapRef(mxco_rb_view const& view) noexcept : v_(view) {}

#if defined(__HIPCC__)
    __device__ bool Contains(Key value) const noexcept
    {
        return detail::rb_contains(static_cast<detail::RbBucket const*>(v_.buckets), v_.num_buckets, v_.dense,
                                   v_.container_keys, v_.container_descs, v_.image, static_cast<uint64_t>(value));
    }
    __host__ __device__ uint64_t Size() const noexcept { return v_.num_keys; }
    __host__ __device__ bool Empty() const noexcept { return v_.num_keys == 0; }
    __host__ __device__ uint8_t const* Data() const noexcept { return v_.bytes; }
    __host__ __device__ uint64_t SizeBytes() const noexcept { return v_.size_bytes; }
#else
    uint64_t Size() const noexcept { return v_.num_keys; }
    bool Empty() const noexcept { return v_.num_keys == 0; }
    uint8_t const* Data() const noexcept { return v_.bytes; }
    uint64_t SizeBytes() const noexcept { return v_.size_bytes; }
#endif

private:
    mxco_rb_view v_;
};

template <typename Key, typename Extent = mxco::Extent<size_t>, typename Allocator = DefaultAllocator<uint8_t>>
class RoaringBitmap {
    static_assert(std::is_same_v<Key, uint32_t> || std::is_same_v<Key, uint64_t>,
                  "RoaringBitmap key type must be uint32_t or uint64_t");
    static_assert(std::is_same_v<typename Allocator::ValueType, uint8_t>,
                  "RoaringBitmap allocator value type must be uint8_t");
    static_assert(std::is_same_v<Allocator, DefaultAllocator<uint8_t>>,
                  "RoaringBitmap device memory comes from libmxco: use DefaultAllocator<uint8_t>");

public:
    using KeyType = Key;
    using SizeType = typename Extent::ValueType;
    using AllocatorType = Allocator;
    using RefType = RoaringBitmapRef<Key>;

    // Length inferred from the format: the caller vouches for complete input.
    explicit RoaringBitmap(void const* bitmap, Allocator const& allocator = {}, hipStream_t stream = nullptr)
        : RoaringBitmap(bitmap, SIZE_MAX, allocator, stream)
    {
    }
    RoaringBitmap(void const* bitmap, hipStream_t stream) : RoaringBitmap(bitmap, SIZE_MAX, Allocator{}, stream) {}
    RoaringBitmap(void const* bitmap, size_t bitmapBytes, Allocator const& allocator = {},
                  hipStream_t stream = nullptr)
        : allocator_(allocator)
    {
        mxco_rb* h = nullptr;
        detail::check(mxco_rb_create(sizeof(Key), bitmap, bitmapBytes, stream, &h));
        h_.reset(h);
    }
    RoaringBitmap(RoaringBitmap const&) = delete;
    RoaringBitmap& operator=(RoaringBitmap const&) = delete;
    RoaringBitmap(RoaringBitmap&&) noexcept = default;
    RoaringBitmap& operator=(RoaringBitmap&&) noexcept = default;
    ~RoaringBitmap() = default;

    void Contains(void const* keys, void* outputValues, Extent keyNum, hipStream_t stream) const
    {
        run(keys, outputValues, keyNum, stream, 1);
    }
    void ContainsAsync(void const* keys, void* outputValues, Extent keyNum, hipStream_t stream) const
    {
        run(keys, outputValues, keyNum, stream, 0);
    }
    [[nodiscard]] uint64_t Size() const noexcept { return view().num_keys; }
    [[nodiscard]] bool Empty() const noexcept { return Size() == 0; }
    [[nodiscard]] uint8_t const* Data() const noexcept { return view().bytes; }
    [[nodiscard]] uint64_t SizeBytes() const noexcept { return view().size_bytes; }
    [[nodiscard]] Allocator GetAllocator() const { return allocator_; }
    [[nodiscard]] RefType Ref() const noexcept { return RefType(view()); }

private:
    struct Deleter {
        void operator()(mxco_rb* h) const noexcept { (void)mxco_rb_destroy(h); }
    };
    void run(void const* keys, void* out, Extent keyNum, hipStream_t stream, int sync) const
    {
        const SizeType n = static_cast<SizeType>(keyNum);
        if constexpr (std::is_signed_v<SizeType>) {
            if (n < 0) throw std::invalid_argument("RoaringBitmap::ContainsAsync keyNum must not be negative");
        }
        detail::check(mxco_rb_contains(h_.get(), keys, static_cast<uint8_t*>(out), static_cast<uint64_t>(n), stream,
                                       sync));
    }
    mxco_rb_view view() const noexcept
    {
        mxco_rb_view v{};
        if (h_) (void)mxco_rb_info(h_.get(), &v);
        return v;
    }
    Allocator allocator_{};
    std::unique_ptr<mxco_rb, Deleter> h_;
};

}  // namespace mxco
