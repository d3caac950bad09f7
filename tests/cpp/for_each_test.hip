// ForEach<CallbackOp> through the C++ headers, compiled with hipcc so the
// callback is device code (include/mxco/for_each.h).  Checks, against sums
// computed on the host: every present probe key calls the functor once per
// occurrence in the query list, absent and erased keys never do, for linear
// and double hashing maps and for a set.  Prints "OK" on success.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <vector>

#include "mxco/aclco.h"

#define CHECK(c)                                                                \
    do {                                                                        \
        if (!(c)) {                                                             \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            std::exit(1);                                                       \
        }                                                                       \
    } while (0)

struct SumValues {  // acc[0] += value, acc[1] += 1
    unsigned long long* acc;
    __device__ explicit SumValues(uint8_t* args) : acc(reinterpret_cast<unsigned long long*>(args)) {}
    __device__ void operator()(aclco::Pair<int64_t, int64_t> slot) const
    {
        atomicAdd(acc, static_cast<unsigned long long>(slot.second));
        atomicAdd(acc + 1, 1ull);
    }
};

struct SumKeys {
    unsigned long long* acc;
    __device__ explicit SumKeys(uint8_t* args) : acc(reinterpret_cast<unsigned long long*>(args)) {}
    __device__ void operator()(int32_t key) const
    {
        atomicAdd(acc, static_cast<unsigned long long>(key));
        atomicAdd(acc + 1, 1ull);
    }
};

template <typename T>
T* upload(std::vector<T> const& v)
{
    T* p = nullptr;
    CHECK(hipMalloc(&p, v.size() * sizeof(T) + 16) == hipSuccess);
    CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess);
    return p;
}

template <class Map>
void map_case(hipStream_t s, uint64_t seed)
{
    std::mt19937_64 rng(seed);
    const size_t n = 100000;
    std::set<int64_t> seen;
    std::vector<int64_t> keys;
    while (keys.size() < n) {
        const int64_t k = static_cast<int64_t>(rng() >> 2);
        if (seen.insert(k).second) keys.push_back(k);
    }
    std::vector<int64_t> pairs;
    for (auto k : keys) {
        pairs.push_back(k);
        pairs.push_back(k * 7 + 1);
    }
    Map map(aclco::Extent<size_t>{2 * n}, -1, -1, {}, {}, {}, s);
    map.SetErasedKey(-2, s);
    auto* dp = upload(pairs);
    CHECK(map.Insert(dp, n, s) == 0);
    std::vector<int64_t> erased(keys.begin(), keys.begin() + n / 10);  // tombstones
    const std::set<int64_t> gone_set(erased.begin(), erased.end());
    auto* de = upload(erased);
    CHECK(map.Erase(de, erased.size(), s) == 0);
    std::vector<int64_t> q;
    unsigned long long want_sum = 0, want_calls = 0;
    for (size_t i = 0; i < 60000; ++i) {  // present (some twice), erased and absent keys
        const int64_t k = keys[rng() % n];
        q.push_back(k);
        const bool gone = gone_set.count(k) != 0;
        if (!gone) {
            want_sum += static_cast<unsigned long long>(k * 7 + 1);
            ++want_calls;
        }
        q.push_back(static_cast<int64_t>(rng() >> 2) | 1ll << 62);  // absent (keys are < 2^62)
    }
    auto* dq = upload(q);
    std::vector<unsigned long long> zero(2, 0);
    auto* acc = upload(zero);
    map.template ForEach<SumValues>(dq, q.size(), acc, s);
    std::vector<unsigned long long> got(2);
    CHECK(hipMemcpy(got.data(), acc, 16, hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(got[0] == want_sum && got[1] == want_calls);
    CHECK(hipFree(dp) == hipSuccess && hipFree(de) == hipSuccess && hipFree(dq) == hipSuccess &&
          hipFree(acc) == hipSuccess);
}

int main()
{
    hipStream_t s;
    CHECK(hipStreamCreate(&s) == hipSuccess);
    map_case<aclco::StaticMap<int64_t, int64_t>>(s, 1);
    map_case<aclco::StaticMap<int64_t, int64_t, aclco::Extent<size_t>, aclco::EqualTo<int64_t>,
                              aclco::DoubleHashing<aclco::murmurhash3_32<int64_t>, aclco::murmurhash3_32<int64_t>>>>(s,
                                                                                                                    2);
    {  // StaticSet<int32_t>: DoubleHashing<xxhash_32>, bucket 5
        aclco::StaticSet<int32_t> set(aclco::Extent<size_t>{5000}, -1);
        std::vector<int32_t> keys;
        for (int32_t k = 1; k <= 3000; ++k) keys.push_back(k * 3);
        auto* dk = upload(keys);
        CHECK(set.Insert(dk, keys.size(), s) == 0);
        std::vector<int32_t> q = {3, 6, 6, 7, 9000, 9001, -5};
        auto* dq = upload(q);
        std::vector<unsigned long long> zero(2, 0);
        auto* acc = upload(zero);
        set.ForEach<SumKeys>(dq, q.size(), acc, s);
        std::vector<unsigned long long> got(2);
        CHECK(hipMemcpy(got.data(), acc, 16, hipMemcpyDeviceToHost) == hipSuccess);
        CHECK(got[0] == 3 + 6 + 6 + 9000 && got[1] == 4);
        CHECK(hipFree(dk) == hipSuccess && hipFree(dq) == hipSuccess && hipFree(acc) == hipSuccess);
    }
    CHECK(hipStreamDestroy(s) == hipSuccess);
    std::printf("OK\n");
    return 0;
}
