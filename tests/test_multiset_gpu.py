"""StaticMultiset on the GPU through the C ABI against the oracle and the
reference's own test vectors (tests/static_multiset/*.cpp): kept keys in
insertion order exactly, multiplicities exact, Retrieve in query order."""
import numpy as np
import pytest
import torch

from multiset import MultisetOracle

pytestmark = pytest.mark.gpu

TD = {np.int32: torch.int32, np.int64: torch.int64}


def mx():
    import mxco
    return mxco


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def lowest(kd):
    return int(np.iinfo(kd).min)


@pytest.mark.parametrize("kd", [np.int32, np.int64])
def test_reference_order_retrieve_and_short_output(kd):
    # tests/static_multiset/semantics_test.cpp:87-129
    big = int(np.iinfo(kd).max)
    keys = np.array([7, 2, 7, big, -4, 2, 9], kd)
    s = mx().StaticMultiset(5, lowest(kd), TD[kd])
    q = dev(np.array([7, 8, 7, -4], kd))
    for _ in range(3):
        s.clear()
        assert s.insert(dev(keys)) == 2
        out = torch.zeros(8, dtype=TD[kd], device="cuda")
        assert s.retrieve_all(out) == 5
        assert out[:5].cpu().numpy().tolist() == keys[:5].tolist()
        assert s.count(q) == 5
        probe = torch.zeros(8, dtype=TD[kd], device="cuda")
        match = torch.zeros(8, dtype=TD[kd], device="cuda")
        assert s.retrieve(q, probe, match) == 5
        assert probe[:5].cpu().tolist() == [7, 7, 7, 7, -4]
        assert match[:5].cpu().tolist() == [7, 7, 7, 7, -4]
        canary = torch.full((8,), 42, dtype=TD[kd], device="cuda")
        with pytest.raises(mx().LengthError):
            s.retrieve_all(canary, 4)
        assert canary.cpu().tolist() == [42] * 8
        p2, m2 = canary.clone(), canary.clone()
        with pytest.raises(mx().LengthError):
            s.retrieve(q, p2, m2, 4)
        assert p2.cpu().tolist() == [42] * 8 and m2.cpu().tolist() == [42] * 8


@pytest.mark.parametrize("kd", [np.int32, np.int64])
@pytest.mark.parametrize("capacity", [17, 4096, 100000])
def test_batches_vs_oracle(kd, capacity):
    rng = np.random.default_rng(capacity)
    e = lowest(kd)
    s = mx().StaticMultiset(capacity, e, TD[kd])
    ref = MultisetOracle(capacity, e)
    for b in range(4):
        n = int(rng.integers(1, 3 * capacity // 2 + 2))
        keys = (rng.integers(-1000, 1000, n) * 104729).astype(kd)
        keys[rng.random(n) < 0.1] = e
        if b % 2:
            st = rng.integers(0, 3, n).astype(np.int32)
            got = s.insert_if(dev(keys), dev(st), "nonzero")
            exp = ref.insert(keys, st != 0)
        else:
            got = s.insert(dev(keys))
            exp = ref.insert(keys)
        assert got == exp
        assert s.size() == ref.size()
    out = torch.zeros(capacity, dtype=TD[kd], device="cuda")
    assert s.retrieve_all(out) == ref.size()
    assert np.array_equal(out[:ref.size()].cpu().numpy(), ref.retrieve_all().astype(kd))
    q = np.concatenate([ref.retrieve_all()[:500].astype(kd), (rng.integers(-1000, 1000, 500) * 104729).astype(kd),
                        np.array([e], kd)])
    dq = dev(q)
    assert np.array_equal(s.count_each(dq).cpu().numpy().view(np.uint64), ref.count_each(q))
    assert np.array_equal(s.count_each_outer(dq).cpu().numpy().view(np.uint64), ref.count_each_outer(q))
    assert np.array_equal(s.contains(dq).cpu().numpy(), ref.contains(q))
    assert np.array_equal(s.find(dq).cpu().numpy(), ref.find(q))
    assert s.count(dq) == ref.count(q)
    total = ref.count(q)
    probe = torch.zeros(max(total, 1), dtype=TD[kd], device="cuda")
    match = torch.zeros(max(total, 1), dtype=TD[kd], device="cuda")
    assert s.retrieve(dq, probe, match) == total
    rp, rm = ref.retrieve(q)
    assert np.array_equal(probe[:total].cpu().numpy(), rp.astype(kd))
    assert np.array_equal(match[:total].cpu().numpy(), rm.astype(kd))
    st = (rng.integers(0, 2, len(q))).astype(np.int8)
    assert np.array_equal(s.contains(dq, stencil=dev(st), pred="nonzero").cpu().numpy(),
                          ref.contains(q, st != 0))
    assert np.array_equal(s.find(dq, stencil=dev(st), pred="nonzero").cpu().numpy(), ref.find(q, st != 0))


@pytest.mark.parametrize("kd", [np.int32, np.int64])
def test_count_beyond_uint32_and_full_table(kd):
    # tests/static_multiset/semantics_test.cpp:177-200
    n = 65536
    big = int(np.iinfo(kd).max)
    s = mx().StaticMultiset(n, lowest(kd), TD[kd])
    rep = dev(np.full(n, big, kd))
    assert s.insert(rep) == 0
    assert s.count(rep) == 1 << 32
    full = mx().StaticMultiset(16, lowest(kd), TD[kd])
    assert full.insert(dev(np.arange(-8, 8, dtype=kd))) == 0
    miss = dev(np.array([42], kd))
    assert full.count(miss) == 0
    assert full.insert(miss) == 1


def test_errors_and_clear():
    mxco = mx()
    with pytest.raises(ValueError):
        mxco.StaticMultiset(0, -1)
    s = mxco.StaticMultiset(8, -1)
    assert s.insert(None, n=3) == 3 and s.size() == 0
    keys = dev(np.array([5, 5, 6], np.int64))
    assert s.insert(keys) == 0
    with pytest.raises(OverflowError):
        s.count(keys, n=2**63)  # n * Size() may exceed 64 bits: rejected before any read
    assert s.count(None, n=0) == 0
    with pytest.raises(ValueError):
        s.count(None, n=1)
    s.clear()
    assert s.size() == 0 and s.count(keys) == 0
    assert s.retrieve_all(None, 0) == 0
    assert s.insert(keys) == 0 and s.count_each(keys).cpu().tolist() == [2, 2, 1]
