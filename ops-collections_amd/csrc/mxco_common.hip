// Error state, device queries and capacity primes of libmxco.so.
#include <stdlib.h>
#include "mxco_common.h"

#include <mutex>
#include <vector>

namespace mxco {

static thread_local std::string t_last_error;

mxco_status fail(mxco_status st, const std::string& msg)
{
    t_last_error = msg;
    return st;
}

mxco_status ok()
{
    t_last_error.clear();
    return MXCO_OK;
}

mxco_status hip_fail(hipError_t e, const char* what)
{
    std::string msg = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory) return fail(MXCO_OUT_OF_MEMORY, msg);
    return fail(MXCO_RUNTIME_ERROR, msg);
}

int cu_count()
{
    static std::mutex mu;
    static std::vector<int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    std::lock_guard<std::mutex> lk(mu);
    if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
    if (cache[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cache[dev] = n;
    }
    return cache[dev];
}

uint64_t max_batch()
{
    const uint64_t stride = (uint64_t)cu_count() * 1024u;
    return (1ull << 32) - stride;
}

unsigned grid_for(uint64_t threads, unsigned blocks_per_cu)
{
    uint64_t blocks = (threads + 255) / 256;
    const uint64_t cap = (uint64_t)cu_count() * blocks_per_cu;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    return (unsigned)blocks;
}

// ---------------------------------------------------------------------------
// Capacity primes.  The reference's table (include/utility/prime.h:15,
// 140,741 entries) is the set {3, 5, 11, 151} together with the chain
//   p0 = 2,  p_{k+1} = nextprime(p_k + (p_k < 2^17 ? 5 : 2^17))
// up to 17,177,758,133.  The chain is generated on demand, once, up to the
// largest value asked for.
// ---------------------------------------------------------------------------

static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((unsigned __int128)a * b % m); }

static uint64_t powmod(uint64_t a, uint64_t e, uint64_t m)
{
    uint64_t r = 1;
    a %= m;
    for (; e; e >>= 1) {
        if (e & 1) r = mulmod(r, a, m);
        a = mulmod(a, a, m);
    }
    return r;
}

// Deterministic Miller-Rabin for 64-bit integers (first twelve prime bases).
static bool is_prime(uint64_t n)
{
    static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (uint64_t b : bases) {
        if (n % b == 0) return n == b;
    }
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        ++s;
    }
    for (uint64_t b : bases) {
        uint64_t x = powmod(b, d, n);
        if (x == 1 || x == n - 1) continue;
        bool composite = true;
        for (int r = 1; r < s && composite; ++r) {
            x = mulmod(x, x, n);
            if (x == n - 1) composite = false;
        }
        if (composite) return false;
    }
    return true;
}

static const uint64_t kPrimeMax = 17177758133ull;

uint64_t prime_lower_bound(uint64_t n)
{
    static std::mutex mu;
    static std::vector<uint64_t> chain{2};
    if (n > kPrimeMax) return 0;
    // The extra members sit below the chain's gaps they fall into.
    static const uint64_t extra[] = {3, 5, 11, 151};
    std::lock_guard<std::mutex> lk(mu);
    while (chain.back() < n) {
        uint64_t p = chain.back();
        uint64_t q = p + (p < (1ull << 17) ? 5ull : (1ull << 17));
        while (!is_prime(q)) ++q;
        chain.push_back(q);
    }
    // lower_bound over chain U extra
    uint64_t best = 0;
    for (uint64_t e : extra) {
        if (e >= n) {
            best = e;
            break;
        }
    }
    size_t lo = 0, hi = chain.size();
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (chain[mid] < n) lo = mid + 1;
        else hi = mid;
    }
    uint64_t c = chain[lo];
    if (c > kPrimeMax) return 0;
    return (best && best < c) ? best : c;
}

uint64_t valid_extent(uint64_t requested, uint32_t bucket, uint32_t probing)
{
    if (bucket == 0) return 0;
    const uint64_t c = requested ? requested : 1;
    const uint64_t size = c / bucket + (c % bucket != 0);
    if (probing == MXCO_PROBING_LINEAR) return (size + (requested == 0)) * bucket;
    const uint64_t p = prime_lower_bound(size);
    return p ? p * bucket : 0;
}

int env_int(const char* name, int dflt)
{
    const char* s = getenv(name);
    return s && *s ? atoi(s) : dflt;
}

}  // namespace mxco

using namespace mxco;

extern "C" {

MXCO_API const char* mxco_last_error(void) { return t_last_error.c_str(); }

MXCO_API const char* mxco_status_string(mxco_status s)
{
    switch (s) {
    case MXCO_OK: return "ok";
    case MXCO_INVALID_ARGUMENT: return "invalid argument";
    case MXCO_LENGTH_ERROR: return "length error";
    case MXCO_OUT_OF_MEMORY: return "out of memory";
    case MXCO_RUNTIME_ERROR: return "runtime error";
    case MXCO_LOGIC_ERROR: return "logic error";
    case MXCO_UNSUPPORTED: return "unsupported";
    case MXCO_OVERFLOW_ERROR: return "overflow error";
    }
    return "unknown";
}

MXCO_API const char* mxco_version(void) { return "0.1.0"; }

MXCO_API mxco_status mxco_valid_extent(uint64_t requested, uint32_t bucket_size, uint32_t probing,
                                       uint64_t* capacity)
{
    if (!capacity || bucket_size == 0 || probing > 1) return fail(MXCO_INVALID_ARGUMENT, "mxco_valid_extent: bad args");
    const uint64_t c = valid_extent(requested, bucket_size, probing);
    if (c == 0) return fail(MXCO_INVALID_ARGUMENT, "Invalid input Extent");
    *capacity = c;
    return ok();
}

MXCO_API mxco_status mxco_device_cu_count(int* cus)
{
    if (!cus) return fail(MXCO_INVALID_ARGUMENT, "mxco_device_cu_count: null");
    *cus = cu_count();
    return ok();
}

}  // extern "C"
