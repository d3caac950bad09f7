// The host path of the reference's tests, timed as a CPU baseline (TEST
// INFRASTRUCTURE ONLY: called by bench.py's cpu_baseline leg).  The tests'
// golden map is a std::unordered_map filled in input order, first insert of
// a key wins (tests/common/golden.h:23-45), then queried with find.
#include <chrono>
#include <cstdint>
#include <unordered_map>

extern "C" int golden_map_run(const int64_t* keys, const int64_t* vals, uint64_t n, const int64_t* queries,
                              uint64_t nq, int64_t miss, int64_t* out, double* insert_s, double* find_s)
{
    using clock = std::chrono::steady_clock;
    std::unordered_map<int64_t, int64_t> m;
    const auto t0 = clock::now();
    m.reserve(n);
    for (uint64_t i = 0; i < n; ++i) m.emplace(keys[i], vals[i]);
    const auto t1 = clock::now();
    for (uint64_t j = 0; j < nq; ++j) {
        const auto it = m.find(queries[j]);
        out[j] = it == m.end() ? miss : it->second;
    }
    const auto t2 = clock::now();
    *insert_s = std::chrono::duration<double>(t1 - t0).count();
    *find_s = std::chrono::duration<double>(t2 - t1).count();
    return 0;
}
