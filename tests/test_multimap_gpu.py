"""StaticMultimap on the GPU through the C ABI against the numpy contract
restatement (oracle/multimap.py): Find/Contains/Count exact, Retrieve exact
(grouped by query, values ascending), RetrieveAll as a sorted multiset."""
import numpy as np
import pytest
import torch

import oracle as O
from multimap import MultimapOracle

pytestmark = pytest.mark.gpu

TD = {np.int32: torch.int32, np.int64: torch.int64}


def mx():
    import mxco
    return mxco


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def pairs(k, v):
    return dev(O.pack_pairs(k, v))


@pytest.mark.parametrize("kd,vd", [(np.int32, np.int32), (np.int32, np.int64), (np.int64, np.int32),
                                   (np.int64, np.int64)])
def test_multimap_vs_oracle(kd, vd):
    rng = np.random.default_rng(1)
    cap = 50000
    mm = mx().StaticMultimap(cap, -1, -1, TD[kd], TD[vd])
    ref = MultimapOracle(cap, -1, -1, kd, vd)
    assert mm.capacity() == cap
    for _ in range(3):
        base = rng.integers(0, 4000, 8000)
        k = np.repeat(base, rng.integers(1, 9, 8000))[:20000].astype(kd)  # 1..8 copies
        v = rng.integers(0, 10**6, len(k)).astype(vd)
        assert mm.insert(pairs(k, v)) == ref.insert(k, v)
        assert mm.size() == ref.size()
    q = np.concatenate([rng.integers(-5, 5000, 6000), np.array([-1, 3, 3])]).astype(kd)
    np.testing.assert_array_equal(mm.contains(dev(q)).cpu().numpy(), ref.contains(q))
    np.testing.assert_array_equal(mm.find(dev(q)).cpu().numpy(), ref.find(q))
    assert mm.count(dev(q)) == ref.count(q)
    total = ref.count(q)
    po = torch.empty(total, dtype=TD[kd], device="cuda")
    vo = torch.empty(total, dtype=TD[vd], device="cuda")
    assert mm.retrieve(dev(q), po, vo) == total
    ep, ev = ref.retrieve(q)
    np.testing.assert_array_equal(po.cpu().numpy(), ep)
    np.testing.assert_array_equal(vo.cpu().numpy(), ev)
    ko = torch.empty(mm.size(), dtype=TD[kd], device="cuda")
    vo2 = torch.empty(mm.size(), dtype=TD[vd], device="cuda")
    assert mm.retrieve_all(ko, vo2) == mm.size()
    ak, av = ref.retrieve_all()
    got = np.lexsort((vo2.cpu().numpy(), ko.cpu().numpy()))
    np.testing.assert_array_equal(ko.cpu().numpy()[got], ak)
    np.testing.assert_array_equal(vo2.cpu().numpy()[got], av)


def test_multimap_prefix_admission_and_insert_if():
    rng = np.random.default_rng(2)
    mm = mx().StaticMultimap(1000, -1, -1)
    ref = MultimapOracle(1000, -1, -1)
    k = rng.integers(0, 300, 1500).astype(np.int64)
    v = np.arange(1500, dtype=np.int64)
    st = (rng.integers(0, 4, 1500) != 0).astype(np.int32)
    f = mm.insert_if(pairs(k, v), dev(st), "nonzero")
    assert f == ref.insert(k, v, st != 0)
    assert mm.size() == 1000
    q = np.arange(300, dtype=np.int64)
    np.testing.assert_array_equal(mm.find(dev(q)).cpu().numpy(), ref.find(q))
    assert mm.insert(pairs(k[:10], v[:10])) == 10  # full: nothing admitted
    mm.clear()
    assert mm.size() == 0 and mm.count(dev(q)) == 0


def test_multimap_errors():
    import mxco
    with pytest.raises(ValueError):
        mxco.StaticMultimap(0, -1, -1)
    mm = mxco.StaticMultimap(100, -1, -1)
    with pytest.raises(ValueError):  # reserved value in a selected pair: nothing inserted
        mm.insert(pairs(np.array([1, 2], np.int64), np.array([5, -1], np.int64)))
    assert mm.size() == 0
    mm.insert(pairs(np.array([1, 1, 2], np.int64), np.array([5, 6, 7], np.int64)))
    po = torch.empty(1, dtype=torch.int64, device="cuda")
    vo = torch.full((1,), 99, dtype=torch.int64, device="cuda")
    with pytest.raises(mxco.LengthError):
        mm.retrieve(dev(np.array([1], np.int64)), po, vo)
    assert vo.item() == 99  # nothing written
    with pytest.raises(ValueError):
        mm.insert(0, n=3)
    assert mm.find(dev(np.array([-1], np.int64))).item() == -1  # the empty key always misses


@pytest.mark.parametrize("kd,vd", [(np.int32, np.int32), (np.int64, np.int64)])
def test_retrieve_long_and_short_segments(kd, vd):
    """Retrieve sorts short segments in place and long ones (more than 64
    values) with a segmented sort: multiplicities 1, 63, 64, 65, 500 and 3000,
    negative values included, queries repeated and absent."""
    rng = np.random.default_rng(2)
    mult = {11: 1, 12: 63, 13: 64, 14: 65, 15: 500, 16: 3000}
    k = np.concatenate([np.full(c, key) for key, c in mult.items()]).astype(kd)
    v = rng.integers(-10**6, 10**6, len(k)).astype(vd)
    perm = rng.permutation(len(k))
    k, v = k[perm], v[perm]
    mm = mx().StaticMultimap(20000, -1, -1, TD[kd], TD[vd])
    ref = MultimapOracle(20000, -1, -1, kd, vd)
    assert mm.insert(pairs(k, v)) == ref.insert(k, v)
    q = np.array([16, 11, 99, 13, 15, 14, 12, 16, 12], kd)
    total = ref.count(q)
    po = torch.empty(total, dtype=TD[kd], device="cuda")
    vo = torch.empty(total, dtype=TD[vd], device="cuda")
    assert mm.retrieve(dev(q), po, vo) == total
    ep, ev = ref.retrieve(q)
    np.testing.assert_array_equal(po.cpu().numpy(), ep)
    np.testing.assert_array_equal(vo.cpu().numpy(), ev)
    # calls of other sizes reuse (and grow) the scratch: a smaller one, then
    # a larger one with an output larger than needed
    for q2 in (np.array([15, 11], kd), np.tile(np.array([16, 14, 99, 12], kd), 50)):
        total2 = ref.count(q2)
        po2 = torch.empty(total2 + 7, dtype=TD[kd], device="cuda")
        vo2 = torch.empty(total2 + 7, dtype=TD[vd], device="cuda")
        assert mm.retrieve(dev(q2), po2, vo2) == total2
        ep2, ev2 = ref.retrieve(q2)
        np.testing.assert_array_equal(po2[:total2].cpu().numpy(), ep2)
        np.testing.assert_array_equal(vo2[:total2].cpu().numpy(), ev2)
