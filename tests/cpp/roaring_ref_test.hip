// RoaringBitmap and RoaringBitmapRef through the C++ headers, compiled with
// hipcc so RoaringBitmapRef::Contains runs in a caller's kernel.  Bitmaps are
// written here byte by byte in the portable serialization; the expected
// answers are the key sets themselves.  Covers the reference's cases of
// tests/roaring_bitmap/contains_test.cpp:369-430 (64-bit buckets with array
// and run containers, metadata, moves, empty queries) and malformed input.
// Prints "OK" on success.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <utility>
#include <vector>

#include "mxco/aclco.h"

#define CHECK(c)                                                                \
    do {                                                                        \
        if (!(c)) {                                                             \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            std::exit(1);                                                       \
        }                                                                       \
    } while (0)

namespace {

struct Container {
    uint16_t key;
    bool run;
    std::vector<uint16_t> values;                    // array container
    std::vector<std::pair<uint16_t, uint16_t>> runs;  // (start, length - 1)
};

void put16(std::vector<uint8_t>& b, uint32_t v)
{
    b.push_back(v & 0xff);
    b.push_back((v >> 8) & 0xff);
}
void put32(std::vector<uint8_t>& b, uint32_t v)
{
    put16(b, v & 0xffff);
    put16(b, v >> 16);
}

uint32_t card(Container const& c)
{
    if (!c.run) return (uint32_t)c.values.size();
    uint32_t n = 0;
    for (auto const& r : c.runs) n += r.second + 1u;
    return n;
}

// Run-format cookie when any container is a run container; offsets with the
// no-run cookie or with 4 or more containers.
std::vector<uint8_t> make32(std::vector<Container> const& cs)
{
    bool runs = false;
    for (auto const& c : cs) runs |= c.run;
    std::vector<uint8_t> b;
    if (runs) {
        put32(b, 12347u | (uint32_t)(cs.size() - 1) << 16);
        std::vector<uint8_t> bits((cs.size() + 7) / 8, 0);
        for (size_t i = 0; i < cs.size(); ++i)
            if (cs[i].run) bits[i / 8] |= (uint8_t)(1u << (i % 8));
        b.insert(b.end(), bits.begin(), bits.end());
    } else {
        put32(b, 12346u);
        put32(b, (uint32_t)cs.size());
    }
    for (auto const& c : cs) {
        put16(b, c.key);
        put16(b, card(c) - 1);
    }
    const bool offsets = !runs || cs.size() >= 4;
    const size_t table = b.size();
    if (offsets) b.resize(b.size() + 4 * cs.size());
    for (size_t i = 0; i < cs.size(); ++i) {
        if (offsets) {
            const uint32_t at = (uint32_t)b.size();
            for (int k = 0; k < 4; ++k) b[table + 4 * i + k] = (uint8_t)(at >> (8 * k));
        }
        if (cs[i].run) {
            put16(b, (uint32_t)cs[i].runs.size());
            for (auto const& r : cs[i].runs) {
                put16(b, r.first);
                put16(b, r.second);
            }
        } else if (cs[i].values.size() > 4096) {  // bitset container
            std::vector<uint8_t> bits(8192, 0);
            for (uint16_t v : cs[i].values) bits[v / 8] |= (uint8_t)(1u << (v % 8));
            b.insert(b.end(), bits.begin(), bits.end());
        } else {
            for (uint16_t v : cs[i].values) put16(b, v);
        }
    }
    return b;
}

std::vector<uint8_t> make64(std::vector<std::pair<uint32_t, std::vector<uint8_t>>> const& buckets)
{
    std::vector<uint8_t> b;
    put32(b, (uint32_t)buckets.size());
    put32(b, 0);
    for (auto const& [hk, bm] : buckets) {
        put32(b, hk);
        b.insert(b.end(), bm.begin(), bm.end());
    }
    return b;
}

template <class Key>
__global__ void ref_contains(aclco::RoaringBitmapRef<Key> ref, Key const* keys, uint8_t* out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = ref.Contains(keys[i]) ? 1 : 0;
}

template <class Key>
std::vector<uint8_t> via_class(aclco::RoaringBitmap<Key> const& bm, std::vector<Key> const& keys, hipStream_t s)
{
    Key* dk = nullptr;
    uint8_t* dout = nullptr;
    CHECK(hipMalloc(&dk, keys.size() * sizeof(Key)) == hipSuccess);
    CHECK(hipMalloc(&dout, keys.size()) == hipSuccess);
    CHECK(hipMemcpy(dk, keys.data(), keys.size() * sizeof(Key), hipMemcpyHostToDevice) == hipSuccess);
    bm.Contains(dk, dout, aclco::Extent<size_t>{keys.size()}, s);
    std::vector<uint8_t> out(keys.size());
    CHECK(hipMemcpy(out.data(), dout, keys.size(), hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(hipMemset(dout, 7, keys.size()) == hipSuccess);
    ref_contains<Key><<<((int)keys.size() + 63) / 64, 64, 0, s>>>(bm.Ref(), dk, dout, (int)keys.size());
    CHECK(hipStreamSynchronize(s) == hipSuccess);
    std::vector<uint8_t> viaRef(keys.size());
    CHECK(hipMemcpy(viaRef.data(), dout, keys.size(), hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(viaRef == out);
    CHECK(hipFree(dk) == hipSuccess);
    CHECK(hipFree(dout) == hipSuccess);
    return out;
}

template <class F>
bool throws_invalid(F&& f)
{
    try {
        f();
    } catch (std::invalid_argument const&) {
        return true;
    }
    return false;
}

}  // namespace

int main()
{
    hipStream_t s;
    CHECK(hipStreamCreate(&s) == hipSuccess);

    // 64-bit buckets: array {1, 7, 42} under high key 0, runs {10..13} of
    // container 2 under high key 0x80000000.
    auto low = make32({{0, false, {1, 7, 42}, {}}});
    auto high = make32({{2, true, {}, {{10, 3}}}});
    auto bm64 = make64({{0u, low}, {0x80000000u, high}});
    {
        aclco::RoaringBitmap<uint64_t> bm(bm64.data(), bm64.size(), {}, s);
        CHECK(bm.Size() == 7 && !bm.Empty() && bm.SizeBytes() == bm64.size() && bm.Data() != nullptr);
        CHECK(bm.Ref().Size() == bm.Size() && bm.Ref().SizeBytes() == bm.SizeBytes());
        const uint64_t hk = 0x80000000ull << 32;
        std::vector<uint64_t> keys{1, 2, 42, (1ull << 32) | 1, hk | (2u << 16) | 10, hk | (2u << 16) | 13,
                                   hk | (2u << 16) | 14};
        CHECK((via_class(bm, keys, s) == std::vector<uint8_t>{1, 0, 1, 0, 1, 1, 0}));
    }

    // 32-bit: array, bitset (5000 values) and run containers, 4 containers
    // with the run cookie so the offsets are serialized.
    std::vector<uint16_t> dense;
    for (uint32_t v = 0; v < 10000; v += 2) dense.push_back((uint16_t)v);
    auto bm32 = make32({{1, false, {2, 4, 8}, {}},
                        {3, false, dense, {}},
                        {7, true, {}, {{0, 0}, {100, 49}, {65000, 535}}},
                        {9, false, {65535}, {}}});
    {
        aclco::RoaringBitmap<uint32_t> bm(bm32.data(), bm32.size(), {}, s);
        CHECK(bm.Size() == 3 + 5000 + 1 + 50 + 536 + 1);
        std::vector<uint32_t> keys{(1u << 16) | 4, (1u << 16) | 5, (3u << 16) | 9998, (3u << 16) | 9999,
                                   (3u << 16) | 10000, (7u << 16), (7u << 16) | 1, (7u << 16) | 149,
                                   (7u << 16) | 150, (7u << 16) | 65535, (9u << 16) | 65535, (8u << 16) | 100,
                                   0xffffffffu};
        CHECK((via_class(bm, keys, s) == std::vector<uint8_t>{1, 0, 1, 0, 0, 1, 0, 1, 0, 1, 1, 0, 0}));

        // moves: the moved-from object answers empty queries and rejects others
        aclco::RoaringBitmap<uint32_t> moved(std::move(bm));
        CHECK(bm.Data() == nullptr && moved.Size() == 3 + 5000 + 1 + 50 + 536 + 1);
        bm.Contains(nullptr, nullptr, aclco::Extent<size_t>{0}, s);
        uint32_t* dk = nullptr;
        uint8_t* dout = nullptr;
        CHECK(hipMalloc(&dk, 4) == hipSuccess && hipMalloc(&dout, 1) == hipSuccess);
        CHECK(throws_invalid([&] { bm.ContainsAsync(dk, dout, aclco::Extent<size_t>{1}, s); }));
        CHECK(throws_invalid([&] { moved.ContainsAsync(nullptr, dout, aclco::Extent<size_t>{1}, s); }));
        auto repl = make32({{2, false, {16}, {}}});
        aclco::RoaringBitmap<uint32_t> assigned(repl.data(), repl.size(), {}, s);
        assigned = std::move(moved);
        CHECK(assigned.Size() == 3 + 5000 + 1 + 50 + 536 + 1 && moved.Data() == nullptr);
        (void)assigned.GetAllocator();
        CHECK(hipFree(dk) == hipSuccess && hipFree(dout) == hipSuccess);
    }

    // unbounded constructor (length from the format), empty bitmap
    {
        aclco::RoaringBitmap<uint32_t> bm(bm32.data(), s);
        CHECK(bm.SizeBytes() == bm32.size());
        auto empty = make32({});
        aclco::RoaringBitmap<uint32_t> e(empty.data(), empty.size(), {}, s);
        CHECK(e.Empty() && e.Ref().Empty());
        CHECK((via_class(e, std::vector<uint32_t>{0, 1, 0xffffffffu}, s) == std::vector<uint8_t>{0, 0, 0}));
    }

    // malformed input
    {
        auto bad = bm32;
        bad[0] = 0x11;
        CHECK(throws_invalid([&] { aclco::RoaringBitmap<uint32_t> b(bad.data(), bad.size(), {}, s); }));
        CHECK(throws_invalid([&] { aclco::RoaringBitmap<uint32_t> b(bm32.data(), bm32.size() - 1, {}, s); }));
        auto unordered = make32({{5, false, {1}, {}}, {4, false, {1}, {}}});
        CHECK(throws_invalid([&] { aclco::RoaringBitmap<uint32_t> b(unordered.data(), unordered.size(), {}, s); }));
        auto buckets = make64({{3u, low}, {3u, low}});
        CHECK(throws_invalid([&] { aclco::RoaringBitmap<uint64_t> b(buckets.data(), buckets.size(), {}, s); }));
    }

    CHECK(hipStreamDestroy(s) == hipSuccess);
    std::printf("OK\n");
    return 0;
}
