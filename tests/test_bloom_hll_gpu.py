"""BloomFilter and HyperLogLog on the GPU through the C ABI, against the
oracle: filter words and HLL registers bit-exact, Contains outputs exact,
estimates equal to the FP64 finalizer applied to the same registers."""
import math

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

TD = {np.int32: torch.int32, np.int64: torch.int64, np.float32: torch.float32, np.uint32: torch.int32,
      np.uint64: torch.int64}


def mx():
    import mxco
    return mxco


def dev(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    if a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(a).cuda()


def words_of(f):
    return f.words().cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("dtype", [np.int32, np.uint32, np.float32, np.int64, np.uint64])
@pytest.mark.parametrize("num_blocks,n", [(1, 100), (37, 5000), (4096, 200000)])
@pytest.mark.parametrize("seed", [0, 42])
def test_bloom_words_and_contains_bit_exact(dtype, num_blocks, n, seed):
    rng = np.random.default_rng(num_blocks + n)
    keys = rng.integers(-2**31, 2**31 - 1, n).astype(dtype)
    keys[: n // 10] = keys[n // 10: 2 * (n // 10)]  # duplicates
    f = mx().BloomFilter(num_blocks, TD[dtype], seed=seed)
    f.add(dev(keys))
    ref = np.zeros(num_blocks * 8, np.uint32)
    O.bloom_add(ref, num_blocks, keys, seed)
    np.testing.assert_array_equal(words_of(f), ref)
    q = np.concatenate([keys[:1000], rng.integers(-2**31, 2**31 - 1, 3000).astype(dtype)])
    got = f.contains(dev(q)).cpu().numpy()
    np.testing.assert_array_equal(got, O.bloom_contains(ref, num_blocks, q, seed))
    assert got[:min(1000, n)].all()  # no false negatives


def test_bloom_add_if_contains_if_merge_intersect():
    rng = np.random.default_rng(3)
    nb = 1024
    a_keys = rng.integers(0, 2**62, 20000).astype(np.int64)
    b_keys = rng.integers(0, 2**62, 20000).astype(np.int64)
    st = (rng.integers(0, 3, 20000) == 0).astype(np.uint8)
    a = mx().BloomFilter(nb)
    a.add_if(dev(a_keys), dev(st))
    ra = np.zeros(nb * 8, np.uint32)
    O.bloom_add(ra, nb, a_keys[st != 0])
    np.testing.assert_array_equal(words_of(a), ra)
    c = a.contains_if(dev(a_keys), dev(st)).cpu().numpy()
    assert c[st != 0].all() and not c[st == 0].any()
    b = mx().BloomFilter(nb)
    b.add(dev(b_keys))
    rb = np.zeros(nb * 8, np.uint32)
    O.bloom_add(rb, nb, b_keys)
    m = mx().BloomFilter(nb)
    m.merge(a)
    m.merge(b)
    np.testing.assert_array_equal(words_of(m), ra | rb)
    m.intersect(b)
    np.testing.assert_array_equal(words_of(m), (ra | rb) & rb)
    m.clear()
    assert not words_of(m).any()


def test_bloom_errors():
    import mxco
    with pytest.raises(ValueError):
        mxco.BloomFilter(0)
    f = mxco.BloomFilter(8)
    with pytest.raises(ValueError):
        f.add(0, n=4)
    with pytest.raises(ValueError):
        f.merge(mxco.BloomFilter(16))
    with pytest.raises(ValueError):
        f.merge(mxco.BloomFilter(8, seed=1))
    f.add(0, n=0)  # zero keys: no-op


def test_bloom_partitioned_add_ragged_ranges_stencil_duplicates():
    """Batches of at least 2^16 keys and half a key per block take the
    partitioned Add: a partial last range (3 ranges of 2048 blocks + 77), a
    stencil, and 50 distinct keys repeated 2000 times each must still give the
    oracle's words."""
    nb = 3 * 2048 + 77
    rng = np.random.default_rng(8)
    keys = np.repeat(rng.integers(0, 2**62, 50).astype(np.int64), 2000)
    rng.shuffle(keys)
    st = (rng.integers(0, 3, len(keys)) != 0).astype(np.uint8)
    f = mx().BloomFilter(nb)
    f.add_if(dev(keys), dev(st))
    ref = np.zeros(nb * 8, np.uint32)
    O.bloom_add(ref, nb, keys[st != 0])
    np.testing.assert_array_equal(words_of(f), ref)
    wide = rng.integers(0, 2**62, 300000).astype(np.int64)
    f.add(dev(wide))
    O.bloom_add(ref, nb, wide)
    np.testing.assert_array_equal(words_of(f), ref)


def test_bloom_multi_pass_add_equals_split_adds():
    """2^27 + 2^20 keys take two partition passes; the words equal those of
    two separate Adds of the same halves (OR is order-free), and the first
    2^20 keys' blocks match the oracle."""
    nb = 1 << 22
    g = torch.Generator(device="cuda")
    g.manual_seed(9)
    n = (1 << 27) + (1 << 20)
    keys = torch.randint(0, 2**62, (n,), dtype=torch.int64, device="cuda", generator=g)
    a = mx().BloomFilter(nb)
    a.add(keys)
    b = mx().BloomFilter(nb)
    b.add(keys[: 1 << 27])
    b.add(keys[1 << 27:])
    assert torch.equal(a.words(), b.words())
    c = mx().BloomFilter(nb)
    head = keys[: 1 << 20]
    c.add(head)
    ref = np.zeros(nb * 8, np.uint32)
    O.bloom_add(ref, nb, head.cpu().numpy())
    np.testing.assert_array_equal(words_of(c), ref)
    assert a.contains(head).all().item()


def test_bloom_large_fpr():
    """2^22 blocks (128 MiB) with 2^24 keys: words bit-exact, FPR as expected."""
    nb = 1 << 22
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 2**62, 1 << 24).astype(np.int64)
    f = mx().BloomFilter(nb)
    f.add(dev(keys))
    ref = np.zeros(nb * 8, np.uint32)
    O.bloom_add(ref, nb, keys)
    np.testing.assert_array_equal(words_of(f), ref)
    absent = rng.integers(2**62, 2**63 - 1, 1 << 20).astype(np.int64)
    fpr = f.contains(dev(absent)).float().mean().item()
    assert fpr == O.bloom_contains(ref, nb, absent).mean()
    assert fpr < 0.01  # 4 keys per block on average


# --------------------------------------------------------------------------
# BloomFilterPolicy<Key, xxhash_64, uint32_t, W, P>
# --------------------------------------------------------------------------

POLICIES = [(1, 1), (1, 4), (2, 2), (4, 8), (8, 16), (16, 16), (16, 32), (32, 32), (32, 64)]


@pytest.mark.parametrize("wpb,pbits", POLICIES)
@pytest.mark.parametrize("dtype", [np.uint32, np.uint64, np.float32])
def test_bloom_policy_words_and_contains_bit_exact(wpb, pbits, dtype):
    """The seeded PatternBits=16 case of tests/bloom_filter/generalization_test.cpp:104-146
    (257 blocks, keys i*17+3, seed 42) and the other W/P combinations."""
    nb, seed = 257, 42
    keys = (np.arange(2051, dtype=np.uint64) * 17 + 3).astype(dtype)
    misses = (np.arange(257, dtype=np.uint64) * 17 + 3 + 1000000).astype(dtype)
    f = mx().BloomFilter(nb, TD[dtype], seed=seed, words_per_block=wpb, pattern_bits=pbits)
    f.add(dev(keys))
    ref = np.zeros(nb * wpb, np.uint32)
    O.bloom_add(ref, nb, keys, seed, wpb, pbits)
    np.testing.assert_array_equal(words_of(f), ref)
    q = np.concatenate([keys, misses])
    np.testing.assert_array_equal(f.contains(dev(q)).cpu().numpy().astype(bool),
                                  O.bloom_contains(ref, nb, q, seed, wpb, pbits))
    # a stencil, the no-alloc entry point, merge and intersect under the same policy
    st = (np.arange(len(q)) % 3 == 0).astype(np.uint8)
    g = mx().BloomFilter(nb, TD[dtype], seed=seed, words_per_block=wpb, pattern_bits=pbits)
    g.add_if(dev(q), dev(st))
    g2 = mx().BloomFilter(nb, TD[dtype], seed=seed, words_per_block=wpb, pattern_bits=pbits)
    g2.add_noalloc(dev(q), stencil=dev(st))
    ref_g = np.zeros(nb * wpb, np.uint32)
    O.bloom_add(ref_g, nb, q[st == 1], seed, wpb, pbits)
    np.testing.assert_array_equal(words_of(g), ref_g)
    np.testing.assert_array_equal(words_of(g2), ref_g)
    g.merge(f)
    np.testing.assert_array_equal(words_of(g), ref | ref_g)
    g2.intersect(f)
    np.testing.assert_array_equal(words_of(g2), ref & ref_g)


def test_bloom_policy_hotspot_and_repeated_keys():
    """tests/bloom_filter/generalization_test.cpp:148-173: one key 2048 times,
    then 512 keys of one block."""
    nb = 257
    f = mx().BloomFilter(nb, torch.int32, words_per_block=8)
    rep = np.full(2048, 1234567, np.uint32)
    f.add(dev(rep))
    ref = np.zeros(nb * 8, np.uint32)
    O.bloom_add(ref, nb, rep)
    np.testing.assert_array_equal(words_of(f), ref)
    f.clear()
    cand = np.arange(200000, dtype=np.uint32)
    import xxhash
    blk = np.array([((xxhash.xxh64_intdigest(int(c).to_bytes(4, "little"), 0) >> 32) * nb) >> 32 for c in cand])
    hot = cand[blk == 73][:512]
    assert len(hot) == 512
    f.add(dev(hot))
    ref = np.zeros(nb * 8, np.uint32)
    O.bloom_add(ref, nb, hot)
    np.testing.assert_array_equal(words_of(f), ref)
    assert (ref.reshape(nb, 8).any(axis=1).nonzero()[0] == [73]).all()


def test_bloom_caller_storage_and_policy_rules():
    mxco = mx()
    nb, wpb = 100, 16
    store = torch.full((nb * wpb + 8,), 7, dtype=torch.int32, device="cuda")
    f = mxco.BloomFilter(nb, torch.int64, words_per_block=wpb, pattern_bits=32, storage=store)
    assert f.data() == store.data_ptr()
    assert int(store[: nb * wpb].abs().sum().item()) == 0  # cleared
    assert int(store[nb * wpb:].sum().item()) == 7 * 8  # nothing beyond the filter
    keys = np.arange(5000, dtype=np.int64)
    f.add(dev(keys))
    ref = np.zeros(nb * wpb, np.uint32)
    O.bloom_add(ref, nb, keys, 0, wpb, 32)
    np.testing.assert_array_equal(store[: nb * wpb].cpu().numpy().view(np.uint32), ref)
    del f
    torch.cuda.synchronize()
    assert store.sum().item() != 0  # the caller's memory outlives the filter
    with pytest.raises(ValueError):
        mxco.BloomFilter(nb, torch.int64, words_per_block=16, storage=store[:10])
    for w, p in ((3, 3), (8, 12), (64, 64), (1, 33), (8, 72), (0, 0)):
        with pytest.raises(ValueError):
            mxco.BloomFilter(16, torch.int64, words_per_block=w, pattern_bits=p)
    a = mxco.BloomFilter(16, torch.int64, words_per_block=8, pattern_bits=8)
    b = mxco.BloomFilter(16, torch.int64, words_per_block=8, pattern_bits=16)
    with pytest.raises(ValueError):
        a.merge(b)


# --------------------------------------------------------------------------
# HyperLogLog
# --------------------------------------------------------------------------

def hll_tables(p):
    d = np.load(O.os.path.join(O._HERE, "data", "hllpp_p%d.npz" % p))
    return d["raw"], d["bias"], float(d["threshold"])


def oracle_estimate(regs, p):
    h = O.hll_histogram(regs, p)
    raw, bias, thr = hll_tables(p)
    z = O.lib().orc_hll_z(O._ptr(h))
    return int(O.lib().orc_hll_finalize(p, z, int(h[0]), O._ptr(raw), O._ptr(bias), len(raw), thr)), h


@pytest.mark.parametrize("p", [13, 14, 15, 16, 17, 18])
@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_hll_registers_and_estimate(p, dtype):
    rng = np.random.default_rng(p)
    s = mx().HyperLogLog(p, TD[dtype])
    ref = np.zeros(1 << p, np.uint8)
    # sizes on both sides of the direct / LDS-private dispatch boundary
    for n in (1000, 1 << (p + 3), 3 * (1 << 18) + 13):
        keys = rng.integers(-2**31, 2**31 - 1, n).astype(dtype)
        s.add(dev(keys))
        O.hll_add(ref, p, keys)
        regs = s.registers().cpu().numpy()
        np.testing.assert_array_equal(regs, ref)
        est, hist = oracle_estimate(ref, p)
        assert s.histogram() == hist.tolist()
        assert s.estimate() == est


def _alpha_mm(m):
    return (0.7213 / (1.0 + 1.079 / m)) * m * m


@pytest.mark.parametrize("p", [13, 14, 15, 16, 17, 18])
def test_hll_bias_regime_estimate(p):
    """About 3m and 4.5m distinct keys put the raw estimate E inside
    (2.5m, 5m), where HLL++ subtracts the interpolated bias: registers are
    bit-exact and the estimate equals the oracle's finalizer (itself checked
    against a brute-force nearest-six mean on the CPU)."""
    m = 1 << p
    for mult in (3.0, 4.5):
        n = int(mult * m)
        keys = (np.arange(n, dtype=np.int64) * np.int64(0x9E3779B97F4A7C15 & (2**62 - 1)) + p).astype(np.int64)
        s = mx().HyperLogLog(p)
        s.add(dev(keys))
        ref = np.zeros(m, np.uint8)
        O.hll_add(ref, p, keys)
        np.testing.assert_array_equal(s.registers().cpu().numpy(), ref)
        est, hist = oracle_estimate(ref, p)
        z = O.lib().orc_hll_z(O._ptr(hist))
        e = _alpha_mm(m) / z
        assert 2.5 * m < e < 5.0 * m, (p, n, e / m)
        assert est != int(math.floor(e + 0.5))  # the bias term moved it
        assert s.estimate() == est
        assert abs(est - n) / n < 3 * 1.04 / math.sqrt(m)


def test_hll_threshold_branch_estimate():
    """p = 18 registers with about a quarter at zero and the rest at rank 40:
    E is above 2.5m while h = m ln(m/V) is at or below the threshold, so the
    estimate is round(h).  Registers are written straight into the sketch."""
    import ctypes
    p, m = 18, 1 << 18
    s = mx().HyperLogLog(p)
    for zeros in (int(0.27 * m), int(0.28 * m), int(0.25 * m)):
        regs = np.full(m, 40, np.uint8)
        regs[:zeros] = 0
        np.random.default_rng(zeros).shuffle(regs)
        d = torch.from_numpy(regs).cuda()
        torch.cuda.synchronize()
        assert mx()._hip.hipMemcpy(ctypes.c_void_p(s.data()), ctypes.c_void_p(d.data_ptr()),
                                   ctypes.c_size_t(m), 3) == 0
        est, hist = oracle_estimate(regs, p)
        lin = m * math.log(m / zeros)
        if zeros != int(0.25 * m):
            assert est == int(math.floor(lin + 0.5))
        assert s.histogram() == hist.tolist()
        assert s.estimate() == est


def test_hll_merge_and_accuracy():
    p = 15
    rng = np.random.default_rng(9)
    parts = [rng.integers(0, 2**62, 200000).astype(np.int64) for _ in range(4)]
    sk = [mx().HyperLogLog(p) for _ in parts]
    for s, k in zip(sk, parts):
        s.add(dev(k))
    total = mx().HyperLogLog(p)
    for s in sk:
        total.merge(s)
    ref = np.zeros(1 << p, np.uint8)
    for k in parts:
        O.hll_add(ref, p, k)
    np.testing.assert_array_equal(total.registers().cpu().numpy(), ref)
    true = len(np.unique(np.concatenate(parts)))
    est = total.estimate()
    assert est == oracle_estimate(ref, p)[0]
    assert abs(est - true) / true < 3 * 1.04 / math.sqrt(1 << p)


def test_hll_factories_and_errors():
    import mxco
    assert mxco.HyperLogLog.create_with_sketch_size_kb(32).precision == 15
    assert mxco.HyperLogLog.create_with_standard_deviation(0.01).precision == 14
    with pytest.raises(ValueError):
        mxco.HyperLogLog(12)
    with pytest.raises(ValueError):
        mxco.HyperLogLog.create_with_sketch_size_kb(12)
    with pytest.raises(ValueError):
        mxco.HyperLogLog(14).merge(mxco.HyperLogLog(15))
    s = mxco.HyperLogLog(13)
    with pytest.raises(ValueError):
        s.add(0, n=5)
    s.add(0, n=0)
    assert s.estimate() == 0
