// Probe kernels instantiated with the caller's functor types.
//
// The reference's device kernels default-construct the container's KeyEqual,
// its probing scheme (and so the hashers) and the stencil predicate type
// (include/detail/open_addressing/kernels.h:64-66), so any stateless type
// the caller writes works there.  Such types exist only in the caller's
// translation unit: the kernels that use them are instantiated here, in that
// unit (this header needs the HIP compiler), from the probe walk the library
// runs (oa_kernels.h), and registered on the table with
// mxco_oa_set_probe_launcher.  The library keeps the argument checks,
// counters and synchronization of every operation.
//   - hashers: LinearProbing<H> uses H{}, DoubleHashing<H1, H2> uses H1{} and
//     H2{1} (the default arguments of DoubleHashing(),
//     include/probing_scheme.h:99); the library's own hash families keep
//     their compiled form (seed 0 and seed 1).  A result is cast to the
//     32-bit probe size type (SanitizeHash, include/utility/math_utils.h:19-29);
//   - KeyEqual{}(probe key, slot key) decides EQUAL; the EMPTY test is bitwise
//     for keys with a unique object representation and for floating point,
//     and KeyEqual otherwise (EqualWrapper::IsSentinelValue,
//     include/utility/equal_wrapper.h:34-41, include/utility/traits.h:19-28);
//   - a stencil predicate of any type, stateless or not, is applied by
//     Selection below: one kernel writes a byte per input (1 when the
//     predicate holds), and the operation runs the built-in nonzero test on
//     those bytes.
#pragma once

#if !defined(__HIPCC__)
#error "caller-defined KeyEqual, hasher and predicate types need the HIP compiler: their kernels are compiled in the caller's translation unit"
#endif

#include <hip/hip_runtime.h>

#include <new>
#include <type_traits>

#include "../common.h"
#include "oa_kernels.h"

namespace mxco {
namespace detail {

template <class Key>
__device__ __forceinline__ Key key_of_bits(uint64_t bits)
{
    Key k;
    __builtin_memcpy(&k, &bits, sizeof(Key));  // the low sizeof(Key) bytes (little-endian)
    return k;
}

template <class Key>
inline constexpr bool kBitwiseSentinel =
    std::has_unique_object_representations_v<Key> || std::is_floating_point_v<Key> || IsFloatKey<Key>::value;

// hash of a key by hasher H with the given default seed: the compiled
// family for the library's hashers, H{} / H{1} for the caller's
template <class Key, class H, bool Second>
__device__ __forceinline__ uint32_t caller_hash(const OaArgs&, uint64_t bits)
{
    if constexpr (HasHashId<H>::value) {
        return probe_hash<(int)sizeof(Key)>(H::id, bits, Second ? 1u : 0u);
    } else if constexpr (Second) {
        return static_cast<uint32_t>(H{1}(key_of_bits<Key>(bits)));
    } else {
        return static_cast<uint32_t>(H{}(key_of_bits<Key>(bits)));
    }
}

// H2 = void: LinearProbing (no step hash)
template <class Key, class KeyEqual, class H1, class H2>
struct CallerFns {
    template <int KS>
    static __device__ __forceinline__ uint32_t hash1(const OaArgs& a, uint64_t key)
    {
        return caller_hash<Key, H1, false>(a, key);
    }
    template <int KS>
    static __device__ __forceinline__ uint32_t hash2(const OaArgs& a, uint64_t key)
    {
        if constexpr (std::is_void_v<H2>) return 0u;  // not called: the walk steps by one bucket
        else return caller_hash<Key, H2, true>(a, key);
    }
    template <int KS>
    static __device__ __forceinline__ bool equal(const OaArgs&, uint64_t slot_key, uint64_t key)
    {
        return KeyEqual{}(key_of_bits<Key>(key), key_of_bits<Key>(slot_key));
    }
    template <int KS>
    static __device__ __forceinline__ bool sentinel(uint64_t key, uint64_t sentinel_key)
    {
        if constexpr (kBitwiseSentinel<Key>) return key == sentinel_key;
        else return KeyEqual{}(key_of_bits<Key>(key), key_of_bits<Key>(sentinel_key));
    }
};

template <class Key, class KeyEqual, class ProbingScheme>
struct CallerFnsOf;
template <class Key, class KeyEqual, class H>
struct CallerFnsOf<Key, KeyEqual, LinearProbing<H>> {
    using type = CallerFns<Key, KeyEqual, H, void>;
};
template <class Key, class KeyEqual, class H1, class H2>
struct CallerFnsOf<Key, KeyEqual, DoubleHashing<H1, H2>> {
    using type = CallerFns<Key, KeyEqual, H1, H2>;
};

// mxco_oa_probe_launcher for a table of Key -> T (T = void: a set).
template <class Key, class T, class Fns>
int caller_launch(int op, const void* kernel_args, uint32_t grid, uint32_t block, hipStream_t stream)
{
    constexpr int KS = (int)sizeof(Key);
    constexpr int VS = std::is_void_v<T> ? 0 : (int)sizeof(std::conditional_t<std::is_void_v<T>, char, T>);
    constexpr int G = kCallerGroup;
    const OaArgs a = *static_cast<const OaArgs*>(kernel_args);
    switch (op) {
    case OP_FIND: oa_kernel<KS, VS, G, OP_FIND, 1, Fns><<<grid, block, 0, stream>>>(a); break;
    case OP_CONTAINS: oa_kernel<KS, VS, G, OP_CONTAINS, 1, Fns><<<grid, block, 0, stream>>>(a); break;
    case OP_COUNT: oa_kernel<KS, VS, G, OP_COUNT, 1, Fns><<<grid, block, 0, stream>>>(a); break;
    case OP_INSERT: oa_kernel<KS, VS, G, OP_INSERT, 1, Fns><<<grid, block, 0, stream>>>(a); break;
    case OP_INSERT_AND_FIND: oa_kernel<KS, VS, G, OP_INSERT_AND_FIND, 1, Fns><<<grid, block, 0, stream>>>(a); break;
    case OP_ERASE: oa_kernel<KS, VS, G, OP_ERASE, 1, Fns><<<grid, block, 0, stream>>>(a); break;
    case OP_INSERT_OR_ASSIGN:
        if constexpr (VS != 0) {
            oa_kernel<KS, VS, G, OP_INSERT_OR_ASSIGN, 1, Fns><<<grid, block, 0, stream>>>(a);
            break;
        } else {
            return (int)hipErrorInvalidValue;
        }
    case OP_ASSIGN:
        if constexpr (VS != 0) {
            oa_kernel<KS, VS, G, OP_ASSIGN, 1, Fns><<<grid, block, 0, stream>>>(a);
            break;
        } else {
            return (int)hipErrorInvalidValue;
        }
    default: return (int)hipErrorNotSupported;  // the diagnostics run on the library's kernels only
    }
    return (int)hipGetLastError();
}

template <class StencilT, class Predicate>
__global__ __launch_bounds__(256) void predicate_flags_kernel(const StencilT* stencil, Predicate pred, uint64_t n,
                                                              uint8_t* flags)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        flags[i] = pred(stencil[i]) ? 1 : 0;
}

// A caller's predicate: its value on every stencil element, as bytes from
// the stream-ordered allocator, freed on the stream after the operation.
template <typename StencilT, typename Predicate>
class Selection<StencilT, Predicate, false> {
public:
    Selection(StencilT const* stencil, Predicate const& pred, uint64_t n, hipStream_t stream) : stream_{stream}
    {
        s_.data = stencil;  // null stays null: the operation's own rule for a null stencil applies
        s_.elem_size = 1;
        s_.pred = MXCO_PRED_NONZERO;
        if (n == 0 || stencil == nullptr) return;
        if (hipMallocAsync((void**)&flags_, n, stream) != hipSuccess) throw std::bad_alloc();
        int cus = 0;
        (void)mxco_device_cu_count(&cus);
        const uint64_t blocks = (n + 255) / 256;
        const uint64_t cap = (uint64_t)(cus > 0 ? cus : 1) * 8;
        predicate_flags_kernel<StencilT, Predicate>
            <<<(unsigned)(blocks < cap ? blocks : cap), 256, 0, stream>>>(stencil, pred, n, flags_);
        if (hipGetLastError() != hipSuccess) {
            (void)hipFreeAsync(flags_, stream);
            throw std::runtime_error("predicate: kernel launch failed");
        }
        s_.data = flags_;
    }
    Selection(Selection const&) = delete;
    Selection& operator=(Selection const&) = delete;
    ~Selection()
    {
        if (flags_) (void)hipFreeAsync(flags_, stream_);
    }
    const mxco_stencil* get() const noexcept { return &s_; }

private:
    mxco_stencil s_{};
    uint8_t* flags_ = nullptr;
    hipStream_t stream_;
};

}  // namespace detail
}  // namespace mxco
