// Exercises the C++ drop-in headers (include/mxco/*.h) the way reference
// callers use aclco::StaticMap / StaticSet / StaticMultimap / BloomFilter /
// HyperLogLog (through the aclco alias header):
// device buffers, void* arguments, a stream, exceptions on misuse.
// Prints "OK" and exits 0 when every check holds.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include <algorithm>
#include <limits>
#include <type_traits>
#include <utility>

#include "mxco/aclco.h"

#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            std::exit(1);                                                    \
        }                                                                    \
    } while (0)

template <typename T>
T* to_device(std::vector<T> const& v)
{
    T* p = nullptr;
    CHECK(hipMalloc(&p, v.size() * sizeof(T) + 16) == hipSuccess);
    CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess);
    return p;
}

template <typename T>
std::vector<T> to_host(T const* p, size_t n)
{
    std::vector<T> v(n);
    CHECK(hipMemcpy(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess);
    return v;
}

int main()
{
    hipStream_t s;
    CHECK(hipStreamCreate(&s) == hipSuccess);
    const int n = 10000;

    {  // StaticMap<int32, int32> with the reference's defaults
        mxco::StaticMap<int32_t, int32_t> map(mxco::Extent<size_t>{2 * n}, -1, -1);
        CHECK(map.Capacity() == 20000);
        std::vector<mxco::Pair<int32_t, int32_t>> pairs(n + 10);
        for (int i = 0; i < n; ++i) pairs[i] = {i * 3, i * 7 + 1};
        for (int i = 0; i < 10; ++i) pairs[n + i] = {i * 3, 5};  // duplicates
        auto* dp = to_device(pairs);
        CHECK(map.Insert(dp, pairs.size(), s) == 10);
        std::vector<int32_t> keys(2 * n);
        for (int i = 0; i < 2 * n; ++i) keys[i] = i * 3;
        auto* dk = to_device(keys);
        int32_t* dout;
        bool* dcon;
        CHECK(hipMalloc(&dout, 2 * n * sizeof(int32_t)) == hipSuccess);
        CHECK(hipMalloc(&dcon, 2 * n) == hipSuccess);
        map.Find(dk, dout, keys.size(), s);
        map.Contains(dk, dcon, keys.size(), s);
        auto out = to_host(dout, 2 * n);
        auto con = to_host(reinterpret_cast<uint8_t*>(dcon), 2 * n);
        for (int i = 0; i < 2 * n; ++i) {
            // keys 0..9 were also inserted with value 5: either insert may win
            CHECK(out[i] == (i < n ? i * 7 + 1 : -1) || (i < 10 && out[i] == 5));
            CHECK(con[i] == (i < n));
        }
        CHECK(map.Count(dk, keys.size(), s) == static_cast<size_t>(n));
        size_t assigned = 0;
        CHECK(map.InsertOrAssign(dp, 100, s, assigned) == 0 && assigned == 100);
        CHECK(map.Insert(nullptr, 7, s) == 7);
        bool threw = false;
        try {
            map.Find(dk, dout, size_t{1} << 32, s);
        } catch (std::length_error const&) {
            threw = true;
        }
        CHECK(threw);
        // InsertIf with the IsOdd predicate over an int32 stencil
        mxco::StaticMap<int64_t, int64_t> m64(mxco::Extent<size_t>{1000}, -1, -1);
        std::vector<mxco::Pair<int64_t, int64_t>> p64(100);
        std::vector<int32_t> st(100);
        for (int i = 0; i < 100; ++i) {
            p64[i] = {i + 1000000000000ll, i};
            st[i] = i;
        }
        auto* dp64 = to_device(p64);
        auto* dst = to_device(st);
        CHECK((m64.InsertIf<int32_t, mxco::IsOdd>(dp64, dst, 100, s)) == 0);
        std::vector<int64_t> k64(100);
        for (int i = 0; i < 100; ++i) k64[i] = i + 1000000000000ll;
        auto* dk64 = to_device(k64);
        CHECK(m64.Count(dk64, 100, s) == 50);
        CHECK(hipFree(dp) == hipSuccess && hipFree(dk) == hipSuccess && hipFree(dout) == hipSuccess);
        CHECK(hipFree(dcon) == hipSuccess && hipFree(dp64) == hipSuccess && hipFree(dst) == hipSuccess);
        CHECK(hipFree(dk64) == hipSuccess);
    }
    {  // StaticSet<int64> (DoubleHashing<xxhash_32>, prime capacity)
        mxco::StaticSet<int64_t> set(mxco::Extent<size_t>{20}, -1);
        CHECK(set.Capacity() == 25);
        std::vector<int64_t> keys = {5, 6, 7, 5, -1};
        auto* dk = to_device(keys);
        CHECK(set.Insert(dk, keys.size(), s) == 2);
        int64_t* dout;
        CHECK(hipMalloc(&dout, 5 * sizeof(int64_t)) == hipSuccess);
        set.Find(dk, dout, keys.size(), s);
        auto out = to_host(dout, 5);
        CHECK(out[0] == 5 && out[1] == 6 && out[2] == 7 && out[3] == 5 && out[4] == -1);
        CHECK(hipFree(dk) == hipSuccess && hipFree(dout) == hipSuccess);
    }
    {  // BloomFilter<int64>
        mxco::BloomFilter<int64_t> f(mxco::Extent<size_t>{64});
        CHECK(f.NumWords() == 512 && f.SizeBytes() == 2048);
        std::vector<int64_t> keys(1000);
        for (int i = 0; i < 1000; ++i) keys[i] = i * 11;
        auto* dk = to_device(keys);
        f.Add(dk, keys.size(), s);
        bool* dc;
        CHECK(hipMalloc(&dc, 1000) == hipSuccess);
        f.Contains(dk, keys.size(), dc, s);
        auto c = to_host(reinterpret_cast<uint8_t*>(dc), 1000);
        for (uint8_t b : c) CHECK(b == 1);
        bool threw = false;
        try {
            mxco::BloomFilter<int64_t> bad(mxco::Extent<size_t>{0});
        } catch (std::invalid_argument const&) {
            threw = true;
        }
        CHECK(threw);
        CHECK(hipFree(dk) == hipSuccess && hipFree(dc) == hipSuccess);
    }
    {  // HyperLogLog<int64>
        auto h = mxco::HyperLogLog<int64_t>::CreateWithSketchSizeKB(32);
        CHECK(h.SketchSizeKB() == 32);
        std::vector<int64_t> keys(100000);
        for (int i = 0; i < 100000; ++i) keys[i] = i;
        auto* dk = to_device(keys);
        h.Add(dk, keys.size(), s);
        const double e = static_cast<double>(h.Estimate(s));
        CHECK(e > 100000 * 0.97 && e < 100000 * 1.03);
        bool threw = false;
        try {
            (void)mxco::HyperLogLog<int64_t>::CreateWithPrecision(12);
        } catch (std::invalid_argument const&) {
            threw = true;
        }
        CHECK(threw);
        CHECK(hipFree(dk) == hipSuccess);
    }
    {  // StaticMultimap<int64, int64> through the aclco spelling
        aclco::StaticMultimap<int64_t, int64_t> mm(aclco::Extent<size_t>{64}, -1, -1, s);
        CHECK(mm.Capacity() == 64);
        std::vector<int64_t> pairs;  // key k (1..5) with values 10k+j, j < k
        for (int64_t k = 1; k <= 5; ++k)
            for (int64_t j = k - 1; j >= 0; --j) {
                pairs.push_back(k);
                pairs.push_back(10 * k + j);
            }
        const size_t n = pairs.size() / 2;
        auto* dp = to_device(pairs);
        CHECK(mm.Insert(dp, n, s) == 0);
        CHECK(mm.Size(s) == n);
        std::vector<int64_t> q = {3, 7, 1};
        auto* dq = to_device(q);
        int64_t* dv = nullptr;
        CHECK(hipMalloc(&dv, 64 * sizeof(int64_t)) == hipSuccess);
        mm.Find(dq, dv, q.size(), s);
        auto f = to_host(dv, 3);
        CHECK(f[0] == 30 && f[1] == -1 && f[2] == 10);
        CHECK(mm.Count(dq, q.size(), s) == 4);
        int64_t* dprobe = nullptr;
        CHECK(hipMalloc(&dprobe, 64 * sizeof(int64_t)) == hipSuccess);
        CHECK(mm.Retrieve(dq, q.size(), dprobe, dv, 64, s) == 4);
        auto vals = to_host(dv, 4);
        auto probes = to_host(dprobe, 4);
        CHECK(vals[0] == 30 && vals[1] == 31 && vals[2] == 32 && vals[3] == 10);
        CHECK(probes[0] == 3 && probes[3] == 1);
        bool threw = false;
        try {
            (void)mm.Retrieve(dq, q.size(), dprobe, dv, 2, s);
        } catch (std::length_error const&) {
            threw = true;
        }
        CHECK(threw);
        CHECK(hipFree(dp) == hipSuccess);
        CHECK(hipFree(dq) == hipSuccess);
        CHECK(hipFree(dv) == hipSuccess);
        CHECK(hipFree(dprobe) == hipSuccess);
    }
    {  // StaticMultiset<int32> (tests/static_multiset/semantics_test.cpp:87-129)
        aclco::StaticMultiset<int32_t> ms(aclco::Extent<size_t>{5}, std::numeric_limits<int32_t>::lowest(), s);
        std::vector<int32_t> keys = {7, 2, 7, std::numeric_limits<int32_t>::max(), -4, 2, 9};
        auto* dk = to_device(keys);
        CHECK(ms.Insert(dk, keys.size(), s) == 2);
        CHECK(ms.Size(s) == 5);
        int32_t* dout = nullptr;
        CHECK(hipMalloc(&dout, 8 * sizeof(int32_t)) == hipSuccess);
        CHECK(ms.RetrieveAll(dout, 8, s) == 5);
        auto all = to_host(dout, 5);
        CHECK(std::equal(all.begin(), all.end(), keys.begin()));
        std::vector<int32_t> q = {7, 8, 7, -4};
        auto* dq = to_device(q);
        uint64_t* dc = nullptr;
        CHECK(hipMalloc(&dc, 4 * sizeof(uint64_t)) == hipSuccess);
        ms.CountEach(dq, dc, q.size(), s);
        auto c = to_host(dc, 4);
        CHECK(c[0] == 2 && c[1] == 0 && c[2] == 2 && c[3] == 1);
        CHECK(ms.Count(dq, q.size(), s) == 5);
        bool threw = false;
        try {
            (void)ms.RetrieveAll(dout, 4, s);
        } catch (std::length_error const&) {
            threw = true;
        }
        CHECK(threw);
        CHECK(hipFree(dk) == hipSuccess);
        CHECK(hipFree(dq) == hipSuccess);
        CHECK(hipFree(dc) == hipSuccess);
        CHECK(hipFree(dout) == hipSuccess);
    }
    {  // 1- and 2-byte keys: Pair<uint8_t, uint32_t> (3 padding bytes) and an
       // int16 set whose two slots share one 32-bit CAS word
        static_assert(sizeof(aclco::Pair<uint8_t, uint32_t>) == 8, "Pair<uint8_t, uint32_t> layout");
        static_assert(sizeof(aclco::Pair<int16_t, int64_t>) == 16, "Pair<int16_t, int64_t> layout");
        aclco::StaticMap<uint8_t, uint32_t, aclco::Extent<size_t>, aclco::EqualTo<uint8_t>,
                         aclco::LinearProbing<aclco::murmurhash3_32<uint8_t>>, aclco::Storage<1>>
            map(aclco::Extent<size_t>{128}, 255, 0xffffffffu);
        std::vector<aclco::Pair<uint8_t, uint32_t>> pairs(255);
        for (int i = 0; i < 255; ++i) pairs[i] = {static_cast<uint8_t>(i), static_cast<uint32_t>(i * 1000)};
        auto* dp = to_device(pairs);
        CHECK(map.Insert(dp, pairs.size(), s) == 255 - 128);  // 128 slots
        std::vector<uint8_t> keys(255);
        for (int i = 0; i < 255; ++i) keys[i] = static_cast<uint8_t>(i);
        auto* dk = to_device(keys);
        CHECK(map.Count(dk, keys.size(), s) == 128);
        uint32_t* dv;
        CHECK(hipMalloc(&dv, 255 * sizeof(uint32_t)) == hipSuccess);
        map.Find(dk, dv, keys.size(), s);
        auto v = to_host(dv, 255);
        int found = 0;
        for (int i = 0; i < 255; ++i) {
            CHECK(v[i] == 0xffffffffu || v[i] == static_cast<uint32_t>(i * 1000));
            found += v[i] != 0xffffffffu;
        }
        CHECK(found == 128);

        aclco::StaticSet<int16_t, aclco::Extent<size_t>, aclco::EqualTo<int16_t>,
                         aclco::LinearProbing<aclco::murmurhash3_32<int16_t>>, aclco::Storage<1>>
            set(aclco::Extent<size_t>{2}, -1);
        CHECK(set.Capacity() == 2);
        std::vector<int16_t> both = {2, -2};
        auto* db = to_device(both);
        CHECK(set.Insert(db, 1, s) == 0);
        CHECK(set.Insert(db + 1, 1, s) == 0);
        CHECK(set.Count(db, 2, s) == 2);
        CHECK(set.Erase(db + 1, 1, s) == 0);
        CHECK(set.Count(db, 2, s) == 1);
        CHECK(hipFree(dp) == hipSuccess && hipFree(dk) == hipSuccess && hipFree(dv) == hipSuccess);
        CHECK(hipFree(db) == hipSuccess);
    }
    {  // Bloom filter policies of tests/bloom_filter/generalization_test.cpp:30-55
        using DefaultP = aclco::DefaultFilterPolicy<aclco::xxhash_64<uint32_t>, uint32_t, 8>;
        using ExpectedDefaultP =
            aclco::BloomFilterPolicy<uint32_t, aclco::xxhash_64<uint32_t>, uint32_t, 8, 8, 8, 1, 1, 8, false, false>;
        using Default16 = aclco::DefaultFilterPolicy<aclco::xxhash_64<uint64_t>, uint32_t, 16>;
        using Expected16 = aclco::BloomFilterPolicy<uint64_t, aclco::xxhash_64<uint64_t>, uint32_t, 16>;
        using ArrowP = aclco::ArrowFilterPolicy<uint32_t>;
        static_assert(std::is_same_v<DefaultP, ExpectedDefaultP>);
        static_assert(std::is_same_v<Default16, Expected16>);
        static_assert(std::is_same_v<ArrowP, ExpectedDefaultP>);
        static_assert(aclco::detail::UsesMultiplyHighBlockIndex<DefaultP>::value);
        using Scalar16 = aclco::BloomFilterPolicy<uint64_t, aclco::xxhash_64<uint64_t>, uint32_t, 8, 16, 1, 8, 1, 8,
                                                  true, true>;
        static_assert(Scalar16::patternBits == 16 && Scalar16::bitsPerWord == 2 && Scalar16::conditionalAdd);
        using Filter = aclco::BloomFilter<uint64_t, aclco::Extent<size_t>, Scalar16>;
        const size_t nb = 257;
        Scalar16 policy(typename Scalar16::Hasher{42});
        Filter f(aclco::Extent<size_t>{nb}, policy, typename Filter::AllocatorType{}, s);
        CHECK(f.NumWords() == nb * 8 && f.SizeBytes() == nb * 32);
        std::vector<uint64_t> keys(2051);
        for (size_t i = 0; i < keys.size(); ++i) keys[i] = i * 17 + 3;
        auto* dk = to_device(keys);
        f.Add(dk, keys.size(), s);
        // the words the policy's own host functions give
        std::vector<uint32_t> golden(nb * 8, 0);
        for (uint64_t k : keys) {
            const uint64_t h = policy.HashKey(k);
            const uint32_t blk = policy.BlockIndex(static_cast<uint32_t>(h >> 32), nb);
            for (uint32_t w = 0; w < 8; ++w) golden[blk * 8 + w] |= policy.WordPatternForLane(w, static_cast<uint32_t>(h));
        }
        CHECK(to_host(std::as_const(f).Data(), nb * 8) == golden);
        CHECK(policy.WordPattern<3>(12345u) == policy.WordPatternForLane(3, 12345u));
        // WordsPerBlock 16 (64-byte blocks), 32 pattern bits, and a caller allocator
        using P16 = aclco::BloomFilterPolicy<uint64_t, aclco::xxhash_64<uint64_t>, uint32_t, 16, 32>;
        aclco::BloomFilter<uint64_t, aclco::Extent<size_t>, P16> g(aclco::Extent<size_t>{100}, P16{}, s);
        g.Add(dk, keys.size(), s);
        bool* dc;
        CHECK(hipMalloc(&dc, keys.size()) == hipSuccess);
        g.Contains(dk, keys.size(), dc, s);
        auto c = to_host(reinterpret_cast<uint8_t*>(dc), keys.size());
        for (uint8_t b : c) CHECK(b == 1);
        bool threw = false;
        try {
            aclco::BloomFilter<uint64_t, aclco::Extent<size_t>, P16> h2(aclco::Extent<size_t>{100}, P16{7}, s);
            g.Merge(h2, s);  // different seeds
        } catch (std::invalid_argument const&) {
            threw = true;
        }
        CHECK(threw);
        CHECK(hipFree(dk) == hipSuccess && hipFree(dc) == hipSuccess);
    }
    CHECK(hipStreamDestroy(s) == hipSuccess);
    std::printf("OK\n");
    return 0;
}
