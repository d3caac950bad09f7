// ForEach<CallbackOp> / ForEachAsync<CallbackOp> of StaticMap and StaticSet
// (include/static_map.h:515-558, include/static_set.h:437-480; the walk of
// open_addressing_ref_impl.h:424-454 and the driver of
// open_addressing/kernels.h:378-408, open_addressing_impl.h:325-339).
//
// For every probe key the walk visits the key's probe sequence bucket by
// bucket, calls the functor on each slot whose key equals the probe key, and
// stops at the first EMPTY slot or after a full cycle.  The functor is the
// caller's device code, so it is compiled into the caller's translation unit:
// this header needs the HIP compiler.  CallbackOp provides a device
// constructor taking the callbackArgs device pointer (uint8_t*) and a device
// operator() taking the slot (Pair<Key, T> for a map, Key for a set).  The
// walk hashes and compares as the container's probe kernels do (the
// library's hash families and EqualTo, or the caller's types).
#pragma once

#if !defined(__HIPCC__)
#error "mxco/for_each.h needs the HIP compiler: ForEach runs the caller's device functor"
#endif

#include <hip/hip_runtime.h>

#include <cstring>
#include <stdexcept>

#include "common.h"
#include "detail/oa_kernels.h"

namespace mxco {
namespace detail {

template <class Key>
__device__ __forceinline__ uint64_t key_bits(Key k)
{
    uint64_t b = 0;
    __builtin_memcpy(&b, &k, sizeof(Key));
    return b;
}

template <class Slot>
__device__ __forceinline__ auto slot_key(const Slot& s)
{
    if constexpr (std::is_arithmetic_v<Slot>) return s;
    else return s.first;
}

// `a` carries the table shape, the probing, the hash family and the empty key.
template <class Key, class Slot, class CallbackOp, class Fns>
__global__ __launch_bounds__(256) void for_each_kernel(const Slot* slots, OaArgs a, const Key* keys, uint32_t n,
                                                       uint8_t* args)
{
    constexpr int KS = (int)sizeof(Key);
    CallbackOp op(args);
    const uint32_t cap = a.cap, bucket = a.bucket, nb = a.nbuckets;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const Key key = keys[i];
        const uint64_t kb = key_bits(key);
        // LinearProbing / DoubleHashing::MakeIterator (probing_scheme_impl.inl:118-128, :177-189)
        uint32_t idx = (Fns::template hash1<KS>(a, kb) % nb) * bucket;
        const uint32_t step = a.probing ? (Fns::template hash2<KS>(a, kb) % (nb - 1u) + 1u) * bucket : bucket;
        const uint32_t init = idx;
        bool done = false;
        while (!done) {
            for (uint32_t b = 0; b < bucket; ++b) {
                const Slot s = slots[idx + b];
                const uint64_t sk = key_bits(slot_key(s));
                if (Fns::template sentinel<KS>(sk, a.empty_key)) {  // EMPTY: the walk ends
                    done = true;
                    break;
                }
                if (Fns::template equal<KS>(a, sk, kb)) op(s);  // EQUAL
            }
            if (done) break;
            // ProbingIterator::operator++ (probing_scheme_impl.inl:76-84)
            idx = idx >= cap - step ? idx - (cap - step) : idx + step;
            done = idx == init;
        }
    }
}

template <class Key, class Slot, class CallbackOp, class Fns = BuiltinFns>
void for_each_launch(const Slot* slots, uint64_t cap, uint32_t bucket, uint32_t probing, uint32_t hash,
                     const void* keys, uint64_t n, Key empty, void* args, hipStream_t stream)
{
    if (n == 0 || keys == nullptr) return;
    int cus = 0;
    (void)mxco_device_cu_count(&cus);  // of the current device
    // CheckedBatchSize (open_addressing_impl.h:534-545)
    const uint64_t limit = 0xffffffffull - (uint64_t)(cus > 0 ? cus : 1) * 1024ull;
    if (n > limit) throw std::length_error("ForEach: batch larger than the kernel index range");
    OaArgs a;
    std::memset(&a, 0, sizeof(a));
    a.cap = (uint32_t)cap;
    a.bucket = bucket;
    a.nbuckets = (uint32_t)(cap / bucket);
    a.probing = probing;
    a.hash = hash;
    a.key_float = IsFloatKey<Key>::value ? 1u : 0u;
    std::memcpy(&a.empty_key, &empty, sizeof(Key));
    const uint64_t blocks = (n + 255) / 256;
    const unsigned grid = (unsigned)(blocks < (uint64_t)cus * 8 ? blocks : (uint64_t)cus * 8);
    for_each_kernel<Key, Slot, CallbackOp, Fns><<<grid ? grid : 1, 256, 0, stream>>>(slots, a, (const Key*)keys,
                                                                                     (uint32_t)n, (uint8_t*)args);
    if (hipGetLastError() != hipSuccess) throw std::runtime_error("ForEach: kernel launch failed");
}

}  // namespace detail
}  // namespace mxco
