"""The device hash code (include/mxco/detail/hash.h) compiled for the host:
murmurhash3_32, xxhash_32 and xxhash_64 of 1-, 2-, 4- and 8-byte keys with
seeds 0 and 1 (the probe hash and the double-hashing step) equal the oracle's
byte-generic implementations, which the SMHasher verification value and
python-xxhash pin for every input length."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include <cstdint>
#include "mxco/detail/hash.h"
using namespace mxco;
template <int KS>
void run(uint64_t seed_state)
{
    uint64_t x = seed_state;
    for (int i = 0; i < 2000; ++i) {
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        const uint64_t mask = KS == 8 ? ~0ull : ((1ull << (8 * KS)) - 1);
        const uint64_t key = (x ^ (x >> 29)) & mask;
        for (uint32_t seed = 0; seed < 2; ++seed)
            std::printf("%d %llu %u %u %u %llu\n", KS, (unsigned long long)key, seed, murmur3_32<KS>(key, seed),
                        xxhash32<KS>(key, seed), (unsigned long long)xxhash64<KS>(key, seed));
    }
}
int main()
{
    run<1>(1);
    run<2>(2);
    run<4>(4);
    run<8>(8);
    return 0;
}
"""


@pytest.fixture(scope="module")
def header_hashes(tmp_path_factory):
    d = tmp_path_factory.mktemp("hash")
    src = d / "hash_header.cpp"
    src.write_text(PROGRAM)
    exe = d / "hash_header"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                           str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return [tuple(int(x) for x in line.split()) for line in out.splitlines()]


@pytest.mark.parametrize("ks", [1, 2, 4, 8])
def test_device_hashes_equal_oracle(header_hashes, ks):
    L = O.lib()
    rows = [r for r in header_hashes if r[0] == ks]
    assert len(rows) == 4000
    distinct = set()
    for _, key, seed, m3, x32, x64 in rows:
        data = int(key).to_bytes(ks, "little")
        assert m3 == L.orc_murmur3_32(data, ks, seed), (ks, key, seed)
        assert x32 == L.orc_xxh32(data, ks, seed), (ks, key, seed)
        assert x64 == L.orc_xxh64(data, ks, seed), (ks, key, seed)
        distinct.add(key)
    assert len(distinct) > (200 if ks == 1 else 1000)
