// Caller-defined KeyEqual, hasher and predicate types through the C++
// headers, compiled with hipcc (include/mxco/detail/caller_probe.h).  Checks
// against results computed on the host:
//   - the reference's minimum-signed-hash case (tests/static_map/insert_test.cpp:280-320):
//     every key starts at (uint32_t)INT32_MIN % capacity and the three keys
//     fill the three slots from there;
//   - a caller hash places each key of a LinearProbing / Storage<1> map at
//     hash(key) % capacity, and a DoubleHashing<H1, H2> map starts a walk at
//     H1{}(key) and steps by H2{1}(key); Insert, InsertOrAssign, Erase,
//     InsertAndFind, Find, Contains, Count and ForEach give the host's map;
//   - a KeyEqual on a padded key type: keys equal under it are duplicates,
//     and a key equal to the empty sentinel under it is never inserted
//     (EqualWrapper::IsSentinelValue for types without a unique object
//     representation);
//   - caller predicates (stateless on StaticMap / StaticSet / StaticMultimap,
//     stateful on StaticMultiset) select exactly what the built-in IsNonZero
//     selects on a host-computed mask.
// Prints "OK" on success.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
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

template <typename T>
T* upload(std::vector<T> const& v)
{
    T* p = nullptr;
    CHECK(hipMalloc(&p, v.size() * sizeof(T) + 16) == hipSuccess);
    if (!v.empty()) CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess);
    return p;
}

template <typename T>
std::vector<T> download(const T* p, size_t n)
{
    std::vector<T> v(n);
    if (n) CHECK(hipMemcpy(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess);
    return v;
}

using Ext = aclco::Extent<std::size_t>;

// ---------------------------------------------------------------------------
// hashers
// ---------------------------------------------------------------------------
template <typename Key>
struct MinSignedHash {
    using ArgumentType = Key;
    using ResultType = int32_t;
    __host__ __device__ constexpr ResultType operator()(Key const&) const noexcept
    {
        return std::numeric_limits<int32_t>::min();
    }
};

// key * 3 (a bijection on small keys, so slots are predictable)
struct TimesThree {
    __host__ __device__ uint32_t operator()(uint32_t k) const noexcept { return k * 3u; }
};

// Knuth multiplicative hash plus an offset; DoubleHashing default-constructs
// the second one from {1}
struct OffsetHash {
    uint32_t offset = 0;
    __host__ __device__ uint64_t operator()(uint64_t k) const noexcept
    {
        return (k * 0x9E3779B97F4A7C15ull >> 20) + offset;
    }
};

// ---------------------------------------------------------------------------
// padded key with a KeyEqual that ignores all but the low bit of `a`
// ---------------------------------------------------------------------------
struct PaddedKey {
    uint8_t a;
    uint16_t b;
};
static_assert(sizeof(PaddedKey) == 4 && !std::has_unique_object_representations_v<PaddedKey>);

PaddedKey pk(uint8_t a, uint16_t b)
{
    PaddedKey k;
    std::memset(&k, 0, sizeof(k));
    k.a = a;
    k.b = b;
    return k;
}

struct LowBitEqual {
    __host__ __device__ bool operator()(PaddedKey l, PaddedKey r) const noexcept
    {
        return (l.a & 1) == (r.a & 1) && l.b == r.b;
    }
};
struct LowBitHash {
    __host__ __device__ uint32_t operator()(PaddedKey k) const noexcept
    {
        return (uint32_t)(k.a & 1) * 40503u + (uint32_t)k.b * 2654435761u;
    }
};

// ---------------------------------------------------------------------------
// predicates
// ---------------------------------------------------------------------------
struct AboveFive {
    __device__ bool operator()(int32_t s) const { return s > 5; }
};
struct AtLeast {  // stateful: StaticMultiset takes the predicate object
    int64_t threshold;
    __device__ bool operator()(int64_t s) const { return s >= threshold; }
};

struct SumValues {  // acc[0] += value, acc[1] += 1
    unsigned long long* acc;
    __device__ explicit SumValues(uint8_t* args) : acc(reinterpret_cast<unsigned long long*>(args)) {}
    __device__ void operator()(aclco::Pair<uint32_t, uint32_t> slot) const
    {
        atomicAdd(acc, static_cast<unsigned long long>(slot.second));
        atomicAdd(acc + 1, 1ull);
    }
};

void min_signed_hash_case(hipStream_t s)
{
    using Map = aclco::StaticMap<uint32_t, uint32_t, Ext, aclco::EqualTo<uint32_t>,
                                 aclco::LinearProbing<MinSignedHash<uint32_t>>, aclco::Storage<1>>;
    Map map(Ext{128}, 0xFFFFFFFFu, 0xFFFFFFFFu, {}, {}, {}, s);
    const std::vector<aclco::Pair<uint32_t, uint32_t>> in = {{7, 31}, {8, 32}, {9, 33}};
    auto* d = upload(in);
    CHECK(map.Insert(d, Ext{3}, s) == 0);
    std::vector<uint32_t> q = {7, 8, 9, 10};
    auto* dq = upload(q);
    auto* dout = upload(std::vector<uint32_t>(4, 0));
    CHECK(map.Count(dq, Ext{4}, s) == 3);
    map.Find(dq, dout, Ext{4}, s);
    CHECK(hipStreamSynchronize(s) == hipSuccess);
    auto out = download(dout, 4);
    CHECK(out[0] == 31 && out[1] == 32 && out[2] == 33 && out[3] == 0xFFFFFFFFu);
    // SanitizeHash: the int32 minimum as the 32-bit probe size type
    const size_t cap = map.Capacity();
    const size_t start = (size_t)(2147483648ull % cap);
    auto slots = download(map.Data(), cap);
    std::set<uint32_t> got;
    for (size_t j = 0; j < 3; ++j) got.insert(slots[(start + j) % cap].first);
    CHECK(got == std::set<uint32_t>({7, 8, 9}));
    hipFree(d);
    hipFree(dq);
    hipFree(dout);
}

void caller_hash_map_case(hipStream_t s)
{
    // linear, one slot per bucket: each key at hash % capacity
    using Lin = aclco::StaticMap<uint32_t, uint32_t, Ext, aclco::EqualTo<uint32_t>, aclco::LinearProbing<TimesThree>,
                                 aclco::Storage<1>>;
    Lin lin(Ext{1000}, 0xFFFFFFFFu, 0xFFFFFFFFu, {}, {}, {}, s);
    const size_t cap = lin.Capacity();
    std::vector<aclco::Pair<uint32_t, uint32_t>> in;
    std::set<size_t> pos;
    for (uint32_t k = 1; in.size() < 200; ++k) {
        if (pos.insert((size_t)(k * 3u) % cap).second) in.push_back({k, k * 7 + 1});
    }
    auto* d = upload(in);
    CHECK(lin.Insert(d, Ext{in.size()}, s) == 0);
    auto slots = download(lin.Data(), cap);
    for (auto const& p : in) {
        CHECK(slots[(p.first * 3u) % cap].first == p.first);
        CHECK(slots[(p.first * 3u) % cap].second == p.second);
    }
    // ForEach walks with the same hash
    unsigned long long* acc = upload(std::vector<unsigned long long>(2, 0));
    std::vector<uint32_t> fq;
    for (auto const& p : in) fq.push_back(p.first);
    fq.push_back(100000);
    auto* dfq = upload(fq);
    lin.ForEach<SumValues>(dfq, Ext{fq.size()}, acc, s);
    auto sums = download(acc, 2);
    unsigned long long want = 0;
    for (auto const& p : in) want += p.second;
    CHECK(sums[0] == want && sums[1] == in.size());
    hipFree(d);
    hipFree(acc);
    hipFree(dfq);

    // double hashing, buckets of 2: every operation against the host map
    using Dbl = aclco::StaticMap<uint64_t, uint64_t, Ext, aclco::EqualTo<uint64_t>,
                                 aclco::DoubleHashing<OffsetHash, OffsetHash>, aclco::Storage<2>>;
    const uint64_t E = ~0ull;
    Dbl dbl(Ext{40000}, E, E, {}, {}, {}, s);
    std::mt19937_64 rng(3);
    std::map<uint64_t, uint64_t> host;
    std::vector<aclco::Pair<uint64_t, uint64_t>> ins;
    while (ins.size() < 20000) {
        const uint64_t k = rng() >> 1;
        if (host.emplace(k, k ^ 0x55).second) ins.push_back({k, k ^ 0x55});
    }
    // one key inserted alone lands at its first slot: (H1{}(k) % buckets) * 2
    {
        Dbl one(Ext{4000}, E, E, {}, {}, {}, s);
        auto* d1 = upload(std::vector<aclco::Pair<uint64_t, uint64_t>>{ins[0]});
        CHECK(one.Insert(d1, Ext{1}, s) == 0);
        const size_t c1 = one.Capacity();
        const uint32_t h1 = (uint32_t)OffsetHash{}(ins[0].first);
        auto s1 = download(one.Data(), c1);
        CHECK(s1[(h1 % (c1 / 2)) * 2].first == ins[0].first);
        hipFree(d1);
    }
    auto* dins = upload(ins);
    CHECK(dbl.Insert(dins, Ext{ins.size()}, s) == 0);
    // InsertOrAssign: half new keys, half existing ones with new payloads
    std::vector<aclco::Pair<uint64_t, uint64_t>> ioa;
    for (size_t i = 0; i < 2000; ++i) ioa.push_back({ins[i].first, 7 * i});
    while (ioa.size() < 4000) {
        const uint64_t k = rng() >> 1;
        if (!host.count(k)) {
            host[k] = 0;
            ioa.push_back({k, k + 1});
        }
    }
    for (auto const& p : ioa) host[p.first] = p.second;
    auto* dioa = upload(ioa);
    Dbl::SizeType assigned = 0;
    CHECK(dbl.InsertOrAssign(dioa, Ext{ioa.size()}, s, assigned) == 0);
    CHECK(assigned == 2000);
    // Erase 1000 keys (a tombstone key so the walks go on)
    dbl.SetErasedKey(E - 1, s);
    std::vector<uint64_t> er;
    for (size_t i = 5000; i < 6000; ++i) {
        er.push_back(ins[i].first);
        host.erase(ins[i].first);
    }
    auto* der = upload(er);
    CHECK(dbl.Erase(der, Ext{er.size()}, s) == 0);
    // InsertAndFind: present keys report their payload, new keys their own
    std::vector<aclco::Pair<uint64_t, uint64_t>> iaf = {{ins[10].first, 1}, {ins[5000].first, 2}, {12345, 3}};
    host.erase(12345);
    auto* diaf = upload(iaf);
    auto* dfound = upload(std::vector<uint64_t>(3, 0));
    auto* dflag = upload(std::vector<uint8_t>(3, 9));
    dbl.InsertAndFind(diaf, dfound, dflag, Ext{3}, s);
    auto found = download(dfound, 3);
    auto flag = download(dflag, 3);
    CHECK(flag[0] == 0 && found[0] == host[ins[10].first]);
    CHECK(flag[1] == 1 && found[1] == 2 && flag[2] == 1 && found[2] == 3);
    host[ins[5000].first] = 2;
    host[12345] = 3;
    // Find / Contains / Count over present, erased and absent keys
    std::vector<uint64_t> q;
    for (size_t i = 0; i < 8000; ++i) q.push_back(ins[i].first);
    for (size_t i = 0; i < 1000; ++i) q.push_back(rng() >> 1);
    auto* dq = upload(q);
    auto* dv = upload(std::vector<uint64_t>(q.size(), 0));
    auto* dc = upload(std::vector<uint8_t>(q.size(), 9));
    dbl.Find(dq, dv, Ext{q.size()}, s);
    dbl.Contains(dq, dc, Ext{q.size()}, s);
    const auto cnt = dbl.Count(dq, Ext{q.size()}, s);
    auto v = download(dv, q.size());
    auto c = download(dc, q.size());
    size_t present = 0;
    for (size_t i = 0; i < q.size(); ++i) {
        auto it = host.find(q[i]);
        const bool has = it != host.end();
        present += has;
        CHECK(c[i] == (has ? 1 : 0));
        CHECK(v[i] == (has ? it->second : E));
    }
    CHECK(cnt == present);
    for (void* p : {(void*)dins, (void*)dioa, (void*)der, (void*)diaf, (void*)dfound, (void*)dflag, (void*)dq,
                    (void*)dv, (void*)dc})
        hipFree(p);
}

void key_equal_case(hipStream_t s)
{
    using Set = aclco::StaticSet<PaddedKey, Ext, LowBitEqual, aclco::LinearProbing<LowBitHash>, aclco::Storage<1>>;
    const PaddedKey empty = pk(0xFF, 0xFFFF);
    Set set(Ext{512}, empty, {}, {}, {}, s);
    // {1,5} and {3,5} are equal under LowBitEqual: one of them is a duplicate;
    // {1,0xFFFF} equals the empty sentinel under it: never inserted
    std::vector<PaddedKey> in = {pk(1, 5), pk(3, 5), pk(2, 5), pk(0, 9), pk(1, 0xFFFF)};
    auto* d = upload(in);
    CHECK(set.Insert(d, Ext{in.size()}, s) == 2);
    std::vector<PaddedKey> q = {pk(5, 5), pk(4, 5), pk(2, 9), pk(7, 5), pk(1, 0xFFFF), pk(6, 6)};
    auto* dq = upload(q);
    auto* dc = upload(std::vector<uint8_t>(q.size(), 9));
    set.Contains(dq, dc, Ext{q.size()}, s);
    auto c = download(dc, q.size());
    CHECK(c[0] == 1 && c[1] == 1 && c[2] == 1 && c[3] == 1);  // {5,5}~{1,5}, {4,5}~{2,5}, {2,9}~{0,9}, {7,5}~{1,5}
    CHECK(c[4] == 0 && c[5] == 0);
    CHECK(set.Count(dq, Ext{q.size()}, s) == 4);
    hipFree(d);
    hipFree(dq);
    hipFree(dc);
}

void predicate_case(hipStream_t s)
{
    const int64_t E = -1;
    std::mt19937_64 rng(9);
    const size_t n = 50000;
    std::vector<aclco::Pair<int64_t, int64_t>> in;
    std::vector<int64_t> keys;
    std::vector<int32_t> st(n);
    std::vector<uint8_t> mask(n);
    for (size_t i = 0; i < n; ++i) {
        const int64_t k = (int64_t)(i * 2654435761ull % 1000003ull) + 1;
        in.push_back({k, k * 3});
        keys.push_back(k);
        st[i] = (int32_t)(rng() % 10);
        mask[i] = st[i] > 5;
    }
    auto* din = upload(in);
    auto* dk = upload(keys);
    auto* dst = upload(st);
    auto* dm = upload(mask);
    // StaticMap: InsertIf / FindIf / ContainsIf / AssignIf with AboveFive vs IsNonZero on the mask
    aclco::StaticMap<int64_t, int64_t> a(Ext{2 * n}, E, E, {}, {}, {}, s), b(Ext{2 * n}, E, E, {}, {}, {}, s);
    const auto fa = a.InsertIf<int32_t, AboveFive>(din, dst, Ext{n}, s);
    const auto fb = b.InsertIf<uint8_t, aclco::IsNonZero>(din, dm, Ext{n}, s);
    CHECK(fa == fb);
    auto* va = upload(std::vector<int64_t>(n, 0));
    auto* vb = upload(std::vector<int64_t>(n, 0));
    a.Find(dk, va, Ext{n}, s);
    b.Find(dk, vb, Ext{n}, s);
    auto ha = download(va, n), hb = download(vb, n);
    size_t sel = 0;
    for (size_t i = 0; i < n; ++i) {
        CHECK(ha[i] == hb[i]);
        CHECK(ha[i] == (mask[i] ? keys[i] * 3 : E));
        sel += mask[i];
    }
    CHECK(sel > n / 4 && sel < n / 2);
    a.FindIf<int32_t, AboveFive>(dk, dst, va, Ext{n}, s);
    b.FindIf<uint8_t, aclco::IsNonZero>(dk, dm, vb, Ext{n}, s);
    CHECK(download(va, n) == download(vb, n));
    auto* ca = upload(std::vector<uint8_t>(n, 9));
    auto* cb = upload(std::vector<uint8_t>(n, 9));
    a.ContainsIf<int32_t, AboveFive>(dk, dst, ca, Ext{n}, s);
    b.ContainsIf<uint8_t, aclco::IsNonZero>(dk, dm, cb, Ext{n}, s);
    CHECK(download(ca, n) == download(cb, n));
    a.AssignIfAsync<int32_t, AboveFive>(din, dst, Ext{n}, s);  // same pairs: no change
    a.Find(dk, va, Ext{n}, s);
    CHECK(download(va, n) == hb);

    // StaticSet and StaticMultimap with the same predicate
    aclco::StaticSet<int64_t> sa(Ext{2 * n}, E, {}, {}, {}, s), sb(Ext{2 * n}, E, {}, {}, {}, s);
    CHECK((sa.InsertIf<int32_t, AboveFive>(dk, dst, Ext{n}, s) ==
           sb.InsertIf<uint8_t, aclco::IsNonZero>(dk, dm, Ext{n}, s)));
    sa.Contains(dk, ca, Ext{n}, s);
    sb.Contains(dk, cb, Ext{n}, s);
    CHECK(download(ca, n) == download(cb, n));
    aclco::StaticMultimap<int64_t, int64_t> ma(n, E, E, s), mb(n, E, E, s);
    CHECK((ma.InsertIf<int32_t, AboveFive>(din, dst, n, s) == mb.InsertIf<uint8_t, aclco::IsNonZero>(din, dm, n, s)));
    CHECK(ma.Count(dk, n, s) == mb.Count(dk, n, s));
    ma.ContainsIf<int32_t, AboveFive>(dk, dst, ca, n, s);
    mb.ContainsIf<uint8_t, aclco::IsNonZero>(dk, dm, cb, n, s);
    CHECK(download(ca, n) == download(cb, n));

    // StaticMultiset: a stateful predicate object
    std::vector<int64_t> st64(st.begin(), st.end());
    std::vector<uint8_t> mask3(n);
    for (size_t i = 0; i < n; ++i) mask3[i] = st64[i] >= 3;
    auto* dst64 = upload(st64);
    auto* dm3 = upload(mask3);
    aclco::StaticMultiset<int64_t> xa(Ext{n}, E, s), xb(Ext{n}, E, s);
    CHECK(xa.InsertIf(dk, Ext{n}, dst64, AtLeast{3}, s) == xb.InsertIf(dk, Ext{n}, dm3, aclco::IsNonZero{}, s));
    CHECK(xa.Size(s) == xb.Size(s));
    xa.ContainsIf(dk, Ext{n}, dst64, AtLeast{5}, ca, s);
    std::vector<uint8_t> mask5(n);
    for (size_t i = 0; i < n; ++i) mask5[i] = st64[i] >= 5;
    auto* dm5 = upload(mask5);
    xb.ContainsIf(dk, Ext{n}, dm5, aclco::IsNonZero{}, cb, s);
    CHECK(download(ca, n) == download(cb, n));
    for (void* p : {(void*)din, (void*)dk, (void*)dst, (void*)dm, (void*)va, (void*)vb, (void*)ca, (void*)cb,
                    (void*)dst64, (void*)dm3, (void*)dm5})
        hipFree(p);
}

struct ByteHash {
    __host__ __device__ uint32_t operator()(uint8_t k) const noexcept { return k * 37u + 11u; }
};

void narrow_key_case(hipStream_t s)
{
    // 1-byte keys, a caller hash: every key but the sentinel goes in and is found
    using Map = aclco::StaticMap<uint8_t, uint16_t, Ext, aclco::EqualTo<uint8_t>, aclco::LinearProbing<ByteHash>>;
    Map m(Ext{400}, 0xFF, 0xFFFF, {}, {}, {}, s);
    std::vector<aclco::Pair<uint8_t, uint16_t>> in;
    std::vector<uint8_t> q;
    for (int k = 0; k < 255; ++k) {
        in.push_back({(uint8_t)k, (uint16_t)(k * 100)});
        q.push_back((uint8_t)k);
    }
    auto* d = upload(in);
    CHECK(m.Insert(d, Ext{in.size()}, s) == 0);
    CHECK(m.Insert(d, Ext{in.size()}, s) == 255);  // all duplicates now
    auto* dq = upload(q);
    auto* dv = upload(std::vector<uint16_t>(q.size(), 0));
    m.Find(dq, dv, Ext{q.size()}, s);
    auto v = download(dv, q.size());
    for (int k = 0; k < 255; ++k) CHECK(v[k] == (uint16_t)(k * 100));
    hipFree(d);
    hipFree(dq);
    hipFree(dv);
}

int main()
{
    hipStream_t s;
    CHECK(hipStreamCreate(&s) == hipSuccess);
    min_signed_hash_case(s);
    caller_hash_map_case(s);
    key_equal_case(s);
    predicate_case(s);
    narrow_key_case(s);
    CHECK(hipStreamSynchronize(s) == hipSuccess);
    CHECK(hipStreamDestroy(s) == hipSuccess);
    std::printf("OK\n");
    return 0;
}
