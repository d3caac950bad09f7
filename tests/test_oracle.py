"""Pins the CPU oracle before it is trusted as the checker.

Every function of oracle/oracle.c is checked here against something that is
not a retyping of its own formula: the reference's known-answer tests, the
published test vectors of the hash algorithms, an independent hash library
(python-xxhash), the SHA-256 of the reference's prime table, brute force on
tiny inputs, and closed forms (dict goldens for the map semantics, linear
counting for HyperLogLog).
"""
import hashlib
import os
import math
import struct

import numpy as np
import pytest
import xxhash

import oracle as O

L = O.lib()


# --------------------------------------------------------------------------
# Hash functions
# --------------------------------------------------------------------------

def test_reference_hash_kats():
    # tests/utility/hash_test.cpp:17-46
    assert L.orc_fmix32(0x41424344, 0xFFFFEEEE) == 0x570DF661
    assert L.orc_fmix64(0x41424344, 0xFFFFEEEE) == 0x5EE45A8EB612F708
    assert L.orc_murmur3_32(struct.pack("<I", 0x41424344), 4, 0xFFFFEEEE) == 0xD6BDBBB4
    assert L.orc_xxh32(struct.pack("<I", 0x75BCD15), 4, 0) == 0xB20A85EE
    assert L.orc_xxh32(struct.pack("<I", 0x2A), 4, 0) == 0x454235D1
    # tests/utility/xxhash64_test.cpp:90-103
    for key, seed, exp in [(0, 0, 4246796580750024372), (0, 42, 3614696996920510707),
                           (42, 0, 15516826743637085169), (123456789, 0, 9462334144942111946)]:
        assert L.orc_xxh64(struct.pack("<I", key), 4, seed) == exp
    for key, seed, exp in [(0, 0, 3803688792395291579), (0, 42, 13194218611613725804),
                           (42, 0, 13066772586158965587), (123456789, 0, 14662639848940634189)]:
        assert L.orc_xxh64(struct.pack("<Q", key), 8, seed) == exp


def test_murmur3_published_vectors():
    # SMHasher MurmurHash3_x86_32 verification strings.
    cases = [(b"", 0, 0), (b"", 1, 0x514E28B7), (b"", 0xFFFFFFFF, 0x81F16F39),
             (b"\0\0\0\0", 0, 0x2362F9DE), (b"aaaa", 0x9747B28C, 0x5A97808A),
             (b"Hello, world!", 0x9747B28C, 0x24884CBA),
             (b"The quick brown fox jumps over the lazy dog", 0x9747B28C, 0x2FA826CD)]
    for data, seed, exp in cases:
        assert L.orc_murmur3_32(data, len(data), seed) == exp, (data, seed)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 4, 7, 8, 15, 16, 31, 32, 33, 64, 100])
def test_xxhash_against_library(n):
    rng = np.random.default_rng(n)
    for _ in range(20):
        data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        seed = int(rng.integers(0, 2**32))
        assert L.orc_xxh32(data, n, seed) == xxhash.xxh32_intdigest(data, seed)
        assert L.orc_xxh64(data, n, seed) == xxhash.xxh64_intdigest(data, seed)


M64 = (1 << 64) - 1


def _rotl64(x, r):
    return ((x << r) | (x >> (64 - r))) & M64


def _murmur3_x64_128(data, seed, fmix):
    """MurmurHash3_x64_128 with `fmix` as its 64-bit finalizer."""
    n, nb = len(data), len(data) // 16
    h1 = h2 = seed
    c1, c2 = 0x87C37B91114253D5, 0x4CF5AD432745937F
    for i in range(nb):
        k1, k2 = struct.unpack_from("<QQ", data, 16 * i)
        h1 ^= (_rotl64((k1 * c1) & M64, 31) * c2) & M64
        h1 = (((_rotl64(h1, 27) + h2) & M64) * 5 + 0x52DCE729) & M64
        h2 ^= (_rotl64((k2 * c2) & M64, 33) * c1) & M64
        h2 = (((_rotl64(h2, 31) + h1) & M64) * 5 + 0x38495AB5) & M64
    tail = data[16 * nb:]
    k1 = sum(b << (8 * j) for j, b in enumerate(tail[:8]))
    k2 = sum(b << (8 * j) for j, b in enumerate(tail[8:]))
    if len(tail) > 8:
        h2 ^= (_rotl64((k2 * c2) & M64, 33) * c1) & M64
    if tail:
        h1 ^= (_rotl64((k1 * c1) & M64, 31) * c2) & M64
    h1 ^= n
    h2 ^= n
    h1 = (h1 + h2) & M64
    h2 = (h2 + h1) & M64
    h1, h2 = fmix(h1), fmix(h2)
    h1 = (h1 + h2) & M64
    h2 = (h2 + h1) & M64
    return struct.pack("<QQ", h1, h2)


def _smhasher_verification(hashfn):
    """SMHasher's VerificationTest: hash keys {0..i-1} with seed 256-i,
    hash the concatenation with seed 0, read the first 4 bytes."""
    key, hashes = bytearray(256), bytearray()
    for i in range(256):
        key[i] = i
        hashes += hashfn(bytes(key[:i]), 256 - i)
    return struct.unpack_from("<I", hashfn(bytes(hashes), 0))[0]


def test_fmix_and_murmur_smhasher_verification_values():
    # fmix64 on full 64-bit states (the reference's fmix64 KAT has a 32-bit
    # input, which leaves the first shift unpinned): MurmurHash3_x64_128
    # built on the oracle's fmix64 must give SMHasher's published 0x6384BA69;
    # MurmurHash3_x86_32 must give 0xB0F57EE3.
    assert _smhasher_verification(lambda d, s: _murmur3_x64_128(d, s, lambda h: L.orc_fmix64(h, 0))) == 0x6384BA69
    assert _smhasher_verification(lambda d, s: struct.pack("<I", L.orc_murmur3_32(d, len(d), s))) == 0xB0F57EE3


def test_fmix_matches_murmur_finalizer():
    # fmix32 of the length-xored state is the tail of murmur3 over zero blocks:
    # murmur3("", seed) = fmix32(seed ^ 0).
    for seed in [0, 1, 0xDEADBEEF, 0xFFFFFFFF]:
        assert L.orc_murmur3_32(b"", 0, seed) == L.orc_fmix32(seed, 0)


# --------------------------------------------------------------------------
# Capacity primes and MakeValidExtent
# --------------------------------------------------------------------------

REF_PRIME_SHA256 = "614c7e6da9ec95170031f33739827f56b33c69b2b9854fd7a5585467c9ac8b43"


def test_prime_table_matches_reference_sha256():
    p = O.primes()
    assert len(p) == 140741
    assert p[0] == 2 and p[-1] == 17177758133
    assert hashlib.sha256(p.astype("<i8").tobytes()).hexdigest() == REF_PRIME_SHA256


def test_primes_are_prime_and_sorted():
    p = O.primes()
    assert np.all(np.diff(p) > 0)
    # brute-force primality on the small members
    small = p[p < 200000]
    sieve = np.ones(200000, bool)
    sieve[:2] = False
    for i in range(2, 448):
        if sieve[i]:
            sieve[i * i::i] = False
    assert sieve[small].all()


def test_extent_kats():
    # tests/utility/extent_test.cpp
    assert [O.valid_extent(c, 5, O.PROBE_LINEAR) for c in (0, 1, 5, 6, 7, 20)] == [10, 5, 5, 10, 10, 20]
    assert O.valid_extent(20, 5, O.PROBE_DOUBLE) == 25
    assert O.valid_extent(8, 1, O.PROBE_DOUBLE) == 11
    assert O.valid_extent(21, 3, O.PROBE_DOUBLE) == 21
    # workloads: 2^20 / 2^28 requested slots, bucket 5
    assert O.valid_extent(1 << 20, 5, O.PROBE_LINEAR) == 1048580
    assert O.valid_extent(1 << 28, 5, O.PROBE_LINEAR) == 268435460
    # beyond the largest prime: the reference throws
    assert O.valid_extent(17177758134 * 5, 5, O.PROBE_DOUBLE) == 0


def test_pair_layout():
    # Pair<F,S> alignas(min(16, bit_ceil(sF+sS))) with natural member alignment.
    assert O.pair_layout(4, 4) == (8, 4)
    assert O.pair_layout(8, 8) == (16, 8)
    assert O.pair_layout(4, 8) == (16, 8)
    assert O.pair_layout(8, 4) == (16, 8)
    assert O.pair_layout(4, 0) == (4, 0)


# --------------------------------------------------------------------------
# Probing
# --------------------------------------------------------------------------

def _probe(t, key):
    s, st = O.ctypes.c_uint32(), O.ctypes.c_uint32()
    k = np.array([key], dtype=t.kd)
    L.orc_oa_probe_start(t.t, O._ptr(k), O.ctypes.byref(s), O.ctypes.byref(st))
    return s.value, st.value


def test_probe_start_closed_form():
    t = O.OpenAddressing(np.int32, np.int32, 1000, -1, -1, bucket=5)
    nb = t.capacity // 5
    for key in [0, 1, 42, 123456789, -7]:
        h = L.orc_murmur3_32(struct.pack("<i", key), 4, 0)
        assert _probe(t, key) == ((h % nb) * 5, 5)
    d = O.OpenAddressing(np.int32, None, 1000, -1, None, bucket=5, probing=O.PROBE_DOUBLE,
                         hash_family=O.HASH_XXHASH_32)
    nb = d.capacity // 5
    for key in [0, 1, 42, 123456789]:
        h1 = L.orc_xxh32(struct.pack("<i", key), 4, 0)
        h2 = L.orc_xxh32(struct.pack("<i", key), 4, 1)
        assert _probe(d, key) == ((h1 % nb) * 5, (h2 % (nb - 1) + 1) * 5)


@pytest.mark.parametrize("probing,hashf", [(O.PROBE_LINEAR, O.HASH_MURMUR3_32), (O.PROBE_DOUBLE, O.HASH_XXHASH_32)])
def test_full_cycle_fills_every_slot(probing, hashf):
    # tests/utility/probing_scheme_test.cpp: the walk visits every bucket, so
    # `capacity` distinct keys fill the table and one more fails.
    t = O.OpenAddressing(np.int32, np.int32, 57, -1, -1, bucket=3, probing=probing, hash_family=hashf)
    cap = t.capacity
    keys = np.arange(cap + 1, dtype=np.int32) * 7919
    assert t.insert(keys[:cap], keys[:cap]) == 0
    assert t.insert(keys[cap:], keys[cap:]) == 1
    slots = t.slots().view(np.int32).reshape(cap, 2)
    assert set(slots[:, 0].tolist()) == set(keys[:cap].tolist())


# --------------------------------------------------------------------------
# Map / set semantics against dict goldens (tests/common/golden.h:22-150)
# --------------------------------------------------------------------------

def _golden_insert(keys, values, empty_key):
    g, failed = {}, 0
    for k, v in zip(keys.tolist(), values.tolist()):
        if k == empty_key or k in g:
            failed += 1
        else:
            g[k] = v
    return g, failed


@pytest.mark.parametrize("kd,vd", [(np.int32, np.int32), (np.int64, np.int64), (np.int32, np.int64),
                                   (np.int64, np.int32)])
@pytest.mark.parametrize("probing", [O.PROBE_LINEAR, O.PROBE_DOUBLE])
def test_map_semantics_vs_dict(kd, vd, probing):
    rng = np.random.default_rng(1)
    keys = rng.integers(0, 3000, 4000).astype(kd)
    keys[::97] = -1  # reserved empty key
    vals = rng.integers(-10**6, 10**6, 4000).astype(vd)
    t = O.OpenAddressing(kd, vd, 5000, -1, -1, bucket=5, probing=probing)
    g, failed = _golden_insert(keys, vals, -1)
    assert t.insert(keys, vals) == failed
    q = np.concatenate([keys, rng.integers(3000, 6000, 1000).astype(kd)])
    out, hit = t.find(q)
    exp = np.array([g.get(k, -1) for k in q.tolist()], dtype=vd)
    np.testing.assert_array_equal(out, exp)
    np.testing.assert_array_equal(hit, np.array([k in g for k in q.tolist()]))


@pytest.mark.parametrize("kd,vd", [(np.int32, np.int32), (np.int64, np.int64), (np.int64, np.int32)])
@pytest.mark.parametrize("probing", [O.PROBE_LINEAR, O.PROBE_DOUBLE])
def test_assign_if_vs_dict(kd, vd, probing):
    """AssignIf overwrites the payload of selected keys already present and
    inserts nothing (open_addressing_ref_impl.h:384-410, kernels.h:183-210)."""
    rng = np.random.default_rng(3)
    keys = rng.choice(6000, 3000, replace=False).astype(kd)
    t = O.OpenAddressing(kd, vd, 4000, -1, -1, bucket=5, probing=probing)
    assert t.insert(keys, (keys * 3).astype(vd)) == 0
    g = {k: 3 * k for k in keys.tolist()}
    q = rng.choice(9000, 4000, replace=False).astype(kd)
    nv = rng.integers(-10**6, 10**6, len(q)).astype(vd)
    sel = rng.integers(0, 2, len(q)).astype(bool)
    exp_assigned = 0
    for k, v, f in zip(q.tolist(), nv.tolist(), sel.tolist()):
        if f and k in g:
            g[k] = v
            exp_assigned += 1
    assert t.assign_if(q, nv, sel) == exp_assigned > 500
    probe = np.arange(-1, 9000).astype(kd)
    out, hit = t.find(probe)
    np.testing.assert_array_equal(hit, np.array([k in g for k in probe.tolist()]))
    np.testing.assert_array_equal(out, np.array([g.get(k, -1) for k in probe.tolist()], dtype=vd))
    # no stencil: every pair selected; the empty key assigns nothing
    q2 = np.concatenate([keys[:100], np.array([-1], kd)])
    assert t.assign_if(q2, np.full(len(q2), 11, vd)) == 100
    np.testing.assert_array_equal(t.find(keys[:100])[0], np.full(100, 11, vd))


def test_assign_if_stops_at_the_first_empty_slot():
    """A key stored past an empty slot of its own walk is not reached: the
    walk ends at EMPTY (Find alike)."""
    t = O.OpenAddressing(np.int64, np.int64, 64, -1, -1, bucket=1, probing=O.PROBE_LINEAR)
    key = np.array([12345], np.int64)
    start, _ = t.probe_start(12345)
    slots = t.slots()
    nxt = (start + 1) % t.capacity
    slots[nxt, :8] = key.view(np.uint8)
    slots[nxt, 8:16] = np.array([77], np.int64).view(np.uint8)
    t.load_slots(slots)
    assert t.assign_if(key, np.array([5], np.int64)) == 0
    assert t.slots()[nxt, 8:16].view(np.int64)[0] == 77
    assert not t.find(key)[1][0]


def test_insert_or_assign_and_insert_and_find_vs_dict():
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 500, 2000).astype(np.int64)
    vals = rng.integers(0, 10**9, 2000).astype(np.int64)
    t = O.OpenAddressing(np.int64, np.int64, 1000, -1, -1)
    # InsertOrAssign: last writer wins sequentially, assigned counts existing keys
    g, assigned = {}, 0
    for k, v in zip(keys.tolist(), vals.tolist()):
        assigned += k in g
        g[k] = v
    f, a = t.insert_or_assign(keys, vals)
    assert (f, a) == (0, assigned)
    out, hit = t.find(np.arange(500, dtype=np.int64))
    np.testing.assert_array_equal(out, np.array([g.get(k, -1) for k in range(500)]))
    # InsertAndFind: first occurrence inserts and returns its value, later ones see it
    t2 = O.OpenAddressing(np.int64, np.int64, 1000, -1, -1)
    found, inserted = t2.insert_and_find(keys, vals)
    first = {}
    exp_f, exp_i = [], []
    for k, v in zip(keys.tolist(), vals.tolist()):
        if k not in first:
            first[k] = v
            exp_i.append(True)
        else:
            exp_i.append(False)
        exp_f.append(first[k])
    np.testing.assert_array_equal(found, exp_f)
    np.testing.assert_array_equal(inserted, exp_i)


def test_negative_counts():
    # n > capacity -> fail == n - capacity; reserved key -> failed
    t = O.OpenAddressing(np.int32, np.int32, 10, -1, -1)
    cap = t.capacity
    keys = np.arange(cap + 7, dtype=np.int32)
    assert t.insert(keys, keys) == 7
    t2 = O.OpenAddressing(np.int32, np.int32, 10, -1, -1)
    assert t2.insert(np.array([-1, -1, 3], np.int32), np.array([1, 2, 3], np.int32)) == 2


def test_erase_tombstone_and_plain():
    keys = np.arange(100, dtype=np.int32)
    t = O.OpenAddressing(np.int32, np.int32, 200, -1, -1)
    t.insert(keys, keys)
    assert t.erase(keys[:50], erased_key=-2) == 0
    out, hit = t.find(keys)
    assert not hit[:50].any() and hit[50:].all()
    assert t.erase(keys[:50], erased_key=-2) == 50


def test_set_semantics():
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 1000, 3000).astype(np.int32)
    t = O.OpenAddressing(np.int32, None, 2000, -1, None, probing=O.PROBE_DOUBLE, hash_family=O.HASH_XXHASH_32)
    failed = 3000 - len(set(keys.tolist()))
    assert t.insert(keys) == failed
    q = np.arange(-5, 1500, dtype=np.int32)
    out, hit = t.find(q)
    s = set(keys.tolist())
    np.testing.assert_array_equal(hit, [k in s for k in q.tolist()])
    np.testing.assert_array_equal(out, [k if k in s else -1 for k in q.tolist()])


def test_multithreaded_restatement_matches_sequential():
    rng = np.random.default_rng(4)
    keys = rng.choice(2**31 - 2, 50000, replace=False).astype(np.int64)
    vals = keys * 3 + 1
    a = O.OpenAddressing(np.int64, np.int64, 100000, -1, -1)
    b = O.OpenAddressing(np.int64, np.int64, 100000, -1, -1)
    assert a.insert(keys, vals) == b.mt_insert(keys, vals, 4) == 0
    q = np.concatenate([keys, keys + 2**40])
    np.testing.assert_array_equal(a.find(q)[0], b.mt_find(q, 4))


def test_find_line_count_bounds():
    # every walk reads at least one line; a hit at LF 0.5 rarely reads more than two
    rng = np.random.default_rng(5)
    keys = rng.choice(2**31 - 2, 20000, replace=False).astype(np.int64)
    t = O.OpenAddressing(np.int64, np.int64, 40000, -1, -1)
    t.insert(keys, keys)
    _, _, lines = t.find(keys, lines=True)
    assert 20000 <= lines < 20000 * 2


# --------------------------------------------------------------------------
# Bloom filter: bit-level restatement with an independent XXH64
# --------------------------------------------------------------------------

SALTS = [0x47B6137B, 0x44974D91, 0x8824AD5B, 0xA2B7289D, 0x705495C7, 0x2DF1424B, 0x9EFC4947, 0x5C6BFB31]


def _py_bloom_add(words, num_blocks, key_bytes):
    h = xxhash.xxh64_intdigest(key_bytes, 0)
    blk = ((h >> 32) * num_blocks) >> 32
    lo = h & 0xFFFFFFFF
    for w in range(8):
        bit = ((SALTS[w] * lo) & 0xFFFFFFFF) >> 27
        words[blk * 8 + w] |= 1 << bit


# SHA-256 of BloomFilterPolicy's 64 salts as little-endian uint32
# (include/bloom_filter_policy.h:170-178; tests/common/bloom_filter_golden.h:29-37
# holds the same table).
BLOOM_SALTS_SHA256 = "a67bd0b9b64efd063db9fa8a0728d0762d9edc3b8716e48b6d1359f301c91101"


def test_bloom_salt_table_is_the_references():
    salts = O.bloom_salts()
    assert hashlib.sha256(salts.astype("<u4").tobytes()).hexdigest() == BLOOM_SALTS_SHA256
    assert salts[:8].tolist() == SALTS


def _py_bloom_add_policy(words, num_blocks, key_bytes, wpb, pbits, seed, salts):
    """MakeBloomGolden<Key, W, P> (tests/common/bloom_filter_golden.h:88-146),
    bit by bit, with python-xxhash as the hash."""
    h = xxhash.xxh64_intdigest(key_bytes, seed)
    blk = ((h >> 32) * num_blocks) >> 32
    lo = h & 0xFFFFFFFF
    bpw = pbits // wpb
    for w in range(wpb):
        for b in range(bpw):
            bit = ((salts[w * bpw + b] * lo) & 0xFFFFFFFF) >> 27
            words[blk * wpb + w] |= 1 << bit


@pytest.mark.parametrize("wpb,pbits", [(1, 1), (1, 32), (2, 4), (4, 4), (8, 16), (16, 16), (16, 32), (32, 32),
                                       (32, 64)])
@pytest.mark.parametrize("dtype", [np.uint32, np.int64])
def test_bloom_policies_vs_bitwise_restatement(wpb, pbits, dtype):
    rng = np.random.default_rng(wpb * 100 + pbits)
    nb, seed = 61, 42
    salts = [int(x) for x in O.bloom_salts()]
    keys = rng.integers(0, 2**31, 2500).astype(dtype)
    words = np.zeros(nb * wpb, np.uint32)
    O.bloom_add(words, nb, keys, seed, wpb, pbits)
    ref = [0] * (nb * wpb)
    for k in keys:
        _py_bloom_add_policy(ref, nb, k.tobytes(), wpb, pbits, seed, salts)
    np.testing.assert_array_equal(words, np.array(ref, np.uint32))
    assert O.bloom_contains(words, nb, keys, seed, wpb, pbits).all()
    # a query is contained exactly when every word holds its pattern
    q = rng.integers(2**31, 2**32, 3000).astype(dtype)
    got = O.bloom_contains(words, nb, q, seed, wpb, pbits)
    for k, g in zip(q[:300], got[:300]):
        one = [0] * (nb * wpb)
        _py_bloom_add_policy(one, nb, k.tobytes(), wpb, pbits, seed, salts)
        assert g == all((r & m) == m for r, m in zip(words.tolist(), one))


def test_bloom_and_hll_multithreaded_equal_sequential():
    """The CPU baseline's threaded Add / Contains give the sequential words,
    outputs and registers."""
    rng = np.random.default_rng(31)
    keys = rng.integers(-2**63, 2**63 - 1, 200000, dtype=np.int64)
    for wpb, pb in ((8, 8), (16, 32)):
        a = np.zeros(4096 * wpb, np.uint32)
        b = np.zeros(4096 * wpb, np.uint32)
        O.bloom_add(a, 4096, keys, 3, wpb, pb)
        O.bloom_add_mt(b, 4096, keys, 7, 3, wpb, pb)
        np.testing.assert_array_equal(a, b)
        q = np.concatenate([keys[:1000], rng.integers(0, 2**62, 5000)])
        np.testing.assert_array_equal(O.bloom_contains(a, 4096, q, 3, wpb, pb),
                                      O.bloom_contains_mt(a, 4096, q, 5, 3, wpb, pb))
    r1 = np.zeros(1 << 14, np.uint8)
    r2 = np.zeros(1 << 14, np.uint8)
    O.hll_add(r1, 14, keys)
    O.hll_add_mt(r2, 14, keys, 6)
    np.testing.assert_array_equal(r1, r2)


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.uint32, np.float32])
def test_bloom_vs_bitwise_restatement(dtype):
    rng = np.random.default_rng(6)
    nb = 37
    keys = rng.integers(-10**6, 10**6, 3000).astype(dtype)
    words = np.zeros(nb * 8, np.uint32)
    O.bloom_add(words, nb, keys)
    ref = [0] * (nb * 8)
    for k in keys:
        _py_bloom_add(ref, nb, k.tobytes())
    np.testing.assert_array_equal(words, np.array(ref, np.uint32))
    # no false negatives; contains == all pattern bits present
    assert O.bloom_contains(words, nb, keys).all()


# --------------------------------------------------------------------------
# HyperLogLog
# --------------------------------------------------------------------------

@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_hll_registers_closed_form(dtype):
    # tests/hyperloglog/add_test.cpp:321-335 GoldenAdd
    p = 13
    rng = np.random.default_rng(7)
    keys = rng.integers(-2**31, 2**31 - 1, 5000).astype(dtype)
    regs = np.zeros(1 << p, np.uint8)
    O.hll_add(regs, p, keys)
    ref = np.zeros(1 << p, np.uint8)
    for k in keys:
        h = xxhash.xxh64_intdigest(k.tobytes(), 0)
        idx = h >> (64 - p)
        w = ((h << p) & (2**64 - 1)) | (1 << (p - 1))
        rank = 64 - w.bit_length() + 1
        ref[idx] = max(ref[idx], rank)
    np.testing.assert_array_equal(regs, ref)


def _hll_tables(p):
    t = np.load(O.os.path.join(O._HERE, "data", "hllpp_p%d.npz" % p))
    return t["raw"], t["bias"], float(t["threshold"])


def _finalize(p, z, v):
    raw, bias, thr = _hll_tables(p)
    return int(L.orc_hll_finalize(p, z, v, O._ptr(raw), O._ptr(bias), len(raw), thr))


def _alpha_mm(m):
    return (0.7213 / (1.0 + 1.079 / m)) * m * m


def _nearest_six(p, z):
    """Brute-force HLL++ bias correction: the six raw estimates nearest to E
    over the WHOLE table (sort by |raw - E|, stable), their biases averaged
    (tests/hyperloglog/hll_finalizer_neighbor_count.cpp:36-55).  It searches
    no window, so it shares no step with the oracle's anchor-and-slide."""
    raw, bias, _ = _hll_tables(p)
    e = _alpha_mm(1 << p) / z
    order = sorted(range(len(raw)), key=lambda i: (abs(float(raw[i]) - e), i))
    corr = 0.0
    for i in order[:6]:
        corr += float(bias[i])
    return int(math.floor(e - corr / 6 + 0.5))


def test_hll_estimate_linear_counting_and_raw_regime():
    p = 14
    m = 1 << p
    raw, bias, thr = _hll_tables(p)
    # Linear counting: few keys -> estimate = round(m ln(m/V))
    regs = np.zeros(m, np.uint8)
    O.hll_add(regs, p, np.arange(1000, dtype=np.int64))
    h = O.hll_histogram(regs, p)
    z = L.orc_hll_z(O._ptr(h))
    est = L.orc_hll_finalize(p, z, int(h[0]), O._ptr(raw), O._ptr(bias), len(raw), thr)
    assert est == round(m * math.log(m / h[0]))
    # Large cardinality (E >= 5m): raw HLL estimate alpha m^2 / z
    regs[:] = 20
    h = O.hll_histogram(regs, p)
    z = L.orc_hll_z(O._ptr(h))
    est = L.orc_hll_finalize(p, z, int(h[0]), O._ptr(raw), O._ptr(bias), len(raw), thr)
    alpha = 0.7213 / (1 + 1.079 / m)
    assert est == round(alpha * m * m / z)
    assert abs(z - m * 2.0 ** -20) < 1e-12
    # no registers set: z = 0 -> the u64 maximum; all zero: linear counting of 0
    assert _finalize(p, 0.0, 0) == 2**64 - 1
    assert _finalize(p, float(m), m) == 0


@pytest.mark.parametrize("p", [13, 14, 15, 16, 17, 18])
def test_hll_bias_regime_nearest_six(p):
    """2.5m < E < 5m: the finalizer subtracts the mean bias of the six
    nearest raw estimates.  Checked against the brute-force nearest six on
    (i) a sweep of E over the whole regime, V = 0 and V = 1 (h = m ln m, above
    every precision's threshold), (ii) the reference's own tail cases (z built
    from registers of rank 2 and 3 with E beyond the last raw estimate) and
    its control sums (hll_finalizer_neighbor_count.cpp:74-113)."""
    m = 1 << p
    raw, _, thr = _hll_tables(p)
    assert m * math.log(m) > thr
    sums = [_alpha_mm(m) / e for e in np.linspace(2.5 * m * 1.0001, 5.0 * m * 0.9999, 301)]
    tail = []
    for r2 in range(0, m + 1, max(1, m >> 9)):
        z = r2 / 4.0 + (m - r2) / 8.0
        e = _alpha_mm(m) / z
        if raw[-1] < e < 5.0 * m:
            tail.append(z)
    assert tail, "no tail case at p=%d" % p
    controls = [r2 / 4.0 + (m - r2) / 8.0 for r2 in (m // 2, 3 * m // 4, m)]
    moved = 0
    for z in sums + tail + controls:
        e = _alpha_mm(m) / z
        exp = _nearest_six(p, z)
        if 2.5 * m < e < 5.0 * m:
            moved += exp != int(math.floor(e + 0.5))
        else:
            assert e >= 5.0 * m and exp is not None
            exp = int(math.floor(e + 0.5))
        assert _finalize(p, z, 0) == exp, (p, z, e)
        if e > 2.5 * m:
            assert _finalize(p, z, 1) == exp, (p, z, e)
    assert moved > 250  # the correction is not a no-op in this regime


def test_hll_threshold_branch():
    """V > 0, E > 2.5m and h = m ln(m/V) <= threshold: the estimate is the
    rounded linear-counting value h (finalizer.h:60-62).  Registers can reach
    this only at p = 18 (threshold 350000 > m ln(2.5/0.7213) needs
    threshold/m > 1.243), with about a quarter of the registers at zero and
    the rest at high ranks.  A control with fewer zeros crosses the threshold
    and is bias-corrected (nearest-six)."""
    p = 18
    m = 1 << p
    _, _, thr = _hll_tables(p)
    for zeros, in_branch in ((int(0.27 * m), True), (int(0.28 * m), True), (int(0.265 * m), True), (int(0.25 * m), False)):
        regs = np.full(m, 40, np.uint8)
        regs[:zeros] = 0
        h = O.hll_histogram(regs, p)
        z = L.orc_hll_z(O._ptr(h))
        e = _alpha_mm(m) / z
        lin = m * math.log(m / zeros)
        assert 2.5 * m < e < 5.0 * m
        assert (lin <= thr) == in_branch
        exp = int(math.floor(lin + 0.5)) if in_branch else _nearest_six(p, z)
        assert _finalize(p, z, zeros) == exp
        assert exp != _nearest_six(p, z) or not in_branch


# --------------------------------------------------------------------------
# StaticMultimap restatement against a dict-of-lists brute force
# --------------------------------------------------------------------------

def test_multimap_oracle_vs_brute_force():
    from multimap import MultimapOracle
    rng = np.random.default_rng(8)
    cap = 300
    mm = MultimapOracle(cap, -1, -1)
    ref, size = {}, 0
    for batch in range(4):
        k = rng.integers(0, 50, 120)
        v = rng.integers(0, 1000, 120)
        sel = rng.integers(0, 2, 120).astype(bool)
        failed = mm.insert(k, v, sel)
        chosen = [(a, b) for a, b, s in zip(k.tolist(), v.tolist(), sel.tolist()) if s]
        room = cap - size
        for a, b in chosen[:room]:
            ref.setdefault(a, []).append(b)
        size += min(room, len(chosen))
        assert failed == len(chosen) - min(room, len(chosen))
    q = np.concatenate([np.arange(-1, 60), np.array([3, 3, 7])])
    assert mm.contains(q).tolist() == [k in ref for k in q.tolist()]
    assert mm.find(q).tolist() == [min(ref[k]) if k in ref else -1 for k in q.tolist()]
    assert mm.count(q) == sum(len(ref.get(k, [])) for k in q.tolist())
    probe, vals = mm.retrieve(q)
    exp = [(k, x) for k in q.tolist() for x in sorted(ref.get(k, []))]
    assert list(zip(probe.tolist(), vals.tolist())) == exp
    with pytest.raises(ValueError):
        mm.insert(np.array([1, -1]), np.array([2, 3]))


# --------------------------------------------------------------------------
# StaticMultiset restatement against the reference's own test vectors
# --------------------------------------------------------------------------

def test_multiset_oracle_reference_vectors():
    from multiset import MultisetOracle
    # tests/static_multiset/semantics_test.cpp:87-129 (int32 and int64 alike)
    for big in (2**31 - 1, 2**63 - 1):
        ms = MultisetOracle(5, -big - 1)
        keys = [7, 2, 7, big, -4, 2, 9]
        assert ms.insert(keys) == 2
        assert ms.retrieve_all().tolist() == keys[:5]
        q = [7, 8, 7, -4]
        assert ms.count(q) == 5
        probe, match = ms.retrieve(q)
        assert probe.tolist() == [7, 7, 7, 7, -4] and match.tolist() == probe.tolist()
    # tests/static_multiset/insert_test.cpp:60-74: capacity boundary
    ms = MultisetOracle(10, -2**31)
    assert ms.insert(list(range(25))) == 15 and ms.size() == 10
    # sentinel keys fail, unselected inputs do not
    ms = MultisetOracle(4, -1)
    assert ms.insert([1, -1, 2, -1, 3], selected=[1, 1, 0, 0, 1]) == 1
    assert ms.retrieve_all().tolist() == [1, 3]


def test_multiset_oracle_brute_force():
    from multiset import MultisetOracle
    rng = np.random.default_rng(12)
    ms = MultisetOracle(300, -1)
    kept = []
    for _ in range(5):
        k = rng.integers(-1, 40, 100)
        sel = rng.integers(0, 3, 100) > 0
        failed = ms.insert(k, sel)
        chosen = [int(a) for a, s in zip(k, sel) if s]
        valid = [a for a in chosen if a != -1]
        take = valid[:300 - len(kept)]
        kept += take
        assert failed == len(chosen) - len(take)
    assert ms.retrieve_all().tolist() == kept
    q = np.arange(-2, 45)
    assert ms.count_each(q).tolist() == [kept.count(int(x)) if x != -1 else 0 for x in q]
    assert ms.count(q) == sum(kept.count(int(x)) for x in q if x != -1)
    assert ms.count_each_outer(q).tolist() == [max(1, kept.count(int(x))) if x != -1 else 1 for x in q]


# --------------------------------------------------------------------------
# DynamicMap growth rules (include/dynamic_map.h), hand-computed values and
# the properties the reference's tests check (tests/dynamic_map/*.cpp)
# --------------------------------------------------------------------------

def test_dynamic_map_oracle_growth_rules():
    import dynamic_map as D
    # floor(0.60f * c) - 10000: 0.60f is 0.60000002384..., so 100000 -> 60000
    assert D.usable(100000) == 50000 and D.usable(16666) == 0 and D.usable(200000) == 110000
    m = D.DynamicMapOracle(100000, -1, -2)
    assert m.num_submaps() == 1 and m.total_cap == 100000        # create_test.cpp:41-49
    assert m.insert(range(50000), range(50000)) == 50000 and m.num_submaps() == 1
    # one more key: room 0 -> GrowTarget(1) = 2 * size = 100000 -> next capacity 200000 (room 110000)
    assert m.insert([10**9], [1]) == 1
    assert m.caps == [100000, 200000] and m.next_cap == 400000
    # re-inserting old keys after growth inserts nothing (insert_test.cpp:122-160)
    assert m.insert(range(1000), range(5000, 6000)) == 0
    assert m.find([0, 999], -1) == [0, 999] and m.size == 50001
    # zero initial capacity grows (reserve_test.cpp:82-106)
    z = D.DynamicMapOracle(0, -1, -2)
    assert z.insert([7], [8]) == 1 and z.size == 1 and z.num_submaps() == 2
    with pytest.raises(ValueError):
        D.DynamicMapOracle(10, -1, -1)
    # erase counts distinct present keys once
    assert m.erase([0, 0, 1, 10**9, 123456789]) == 3 and m.size == 49998


# --------------------------------------------------------------------------
# Roaring portable serialization writer (used to build GPU test bitmaps)
# --------------------------------------------------------------------------

def test_roaring_writer_layout_by_hand():
    import roaring as R
    # {1, 2, 3} without runs: cookie 12346, 1 container, header (key 0,
    # cardinality-1 = 2), offset 16 (4 + 4 + 4 + 4 bytes before it), values.
    assert R.serialize32([1, 2, 3]) == bytes.fromhex("3a300000" "01000000" "00000200" "10000000" "010002000300")
    # the same as one run container: cookie 12347 | (1 - 1) << 16, run bitmap
    # 0b1, header, no offsets (fewer than 4 containers), 1 run (start 1, length-1 2)
    assert R.serialize32([1, 2, 3], run_containers=True) == bytes.fromhex("3b300000" "01" "00000200" "0100" "01000200")
    # empty 32-bit bitmap and empty 64-bit bitmap (reference contains_test.cpp:211-224)
    assert R.serialize32([]) == bytes.fromhex("3a300000" "00000000")
    assert R.serialize64([]) == bytes(8)
    # more than 4096 values in one container: an 8192-byte bitset
    b = R.serialize32(range(5000))
    assert len(b) == 4 + 4 + 4 + 4 + 8192 and b[16] == 0xFF


def test_roaring_writer_reproduces_reference_spec_files():
    """The reference's RoaringFormatSpec test files (written by CRoaring
    v5.1.0 from tests/utility/generate_roaring_testdata.cpp:48-101) are pinned
    by size and SHA-256 in cmake/roaring_testdata.cmake:35-48; the writer
    rebuilds all three byte for byte from the same key sets."""
    import hashlib
    import roaring as R
    k = list(range(0, 100000, 1000)) + [3 * x for x in range(100000, 200000)] + list(range(700000, 800000))
    low = list(range(0, 0x9001)) + list(range(0xA000, 0x10001)) + [0x20000, 0x20005]
    low += [0x80000 + j for j in range(0, 0x10000, 2)]
    cases = [
        (R.serialize32(k), 72616, "d719ae2e0150a362ef7cf51c361527585891f01460b1a92bcfb6a7257282a442"),
        (R.serialize32(k, run_containers="optimize"), 48056,
         "1f1909bfdd354fa2f0694fe88b8076833ca5383ad9fc3f68f2709c84a2ab70e3"),
        (R.serialize64(low + [(1 << 32) | x for x in low], run_containers="optimize"), 16506,
         "b5a553a759167f5f9ccb3fa21552d943b4c73235635b753376f4faf62067d178"),
    ]
    for data, size, sha in cases:
        assert len(data) == size and hashlib.sha256(data).hexdigest() == sha


def test_bench_roaring_writer_matches_oracle_writer():
    """The vectorised writer that builds the bench's bitmaps produces the same
    bytes as the oracle writer (itself pinned by the reference's SHA-256s)."""
    import roaring as R
    from mxco import workloads as W
    k = list(range(0, 100000, 1000)) + [3 * x for x in range(100000, 200000)] + list(range(700000, 800000))
    for mode in (False, "optimize"):
        assert W.roaring_portable32(k, mode) == R.serialize32(k, run_containers=mode)
    rng = np.random.default_rng(3)
    for t in range(12):
        ks = np.concatenate([rng.integers(0, 1 << int(rng.integers(8, 33)), int(rng.integers(0, 20000))),
                             np.concatenate([np.arange(s, s + int(rng.integers(1, 3000)))
                                             for s in rng.integers(0, 1 << 20, 5)])]) & 0xFFFFFFFF
        for mode in (False, "optimize"):
            assert W.roaring_portable32(ks, mode) == R.serialize32(ks, run_containers=mode), (t, mode)
    low = list(range(0, 0x9001)) + list(range(0xA000, 0x10001)) + [0x20000, 0x20005]
    low += [0x80000 + j for j in range(0, 0x10000, 2)]
    kk = low + [(1 << 32) | x for x in low]
    assert W.roaring_portable64(kk, "optimize") == R.serialize64(kk, run_containers="optimize")


def test_golden_map_baseline_library():
    """The std::unordered_map golden path timed by the CPU baseline: first
    insert of a key wins, misses return the given value."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(os.path.dirname(O.__file__), "libgolden_map.so"))
    keys = np.array([5, 7, 5, 9], np.int64)
    vals = np.array([50, 70, 51, 90], np.int64)
    q = np.array([5, 6, 7, 9, 10], np.int64)
    out = np.empty_like(q)
    a, b = ctypes.c_double(), ctypes.c_double()
    P = ctypes.c_void_p
    lib.golden_map_run(P(keys.ctypes.data), P(vals.ctypes.data), ctypes.c_uint64(4), P(q.ctypes.data),
                       ctypes.c_uint64(5), ctypes.c_int64(-1), P(out.ctypes.data), ctypes.byref(a), ctypes.byref(b))
    assert out.tolist() == [50, -1, 70, 90, -1]
