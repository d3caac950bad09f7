"""DynamicMap on the GPU through the C ABI against the oracle's restatement
of include/dynamic_map.h: inserted / erased counts, Size(), Capacity() and
NumSubmaps() after every call, Find / Contains outputs bit-exact."""
import numpy as np
import pytest
import torch

from dynamic_map import DynamicMapOracle

pytestmark = pytest.mark.gpu

EK, EV, RK = -1, -1, -2


def mx():
    import mxco
    return mxco


def pairs_of(keys, vals):
    return mx().make_pairs(torch.from_numpy(keys).cuda(), torch.from_numpy(vals).cuda())


def same_state(d, o):
    assert (d.size(), d.capacity(), d.num_submaps()) == (o.size, o.total_cap, o.num_submaps())


def check_queries(d, o, q):
    got_v = d.find(torch.from_numpy(q).cuda()).cpu().numpy()
    got_c = d.contains(torch.from_numpy(q).cuda()).cpu().numpy()
    assert got_v.tolist() == o.find(q.tolist(), EV)
    assert got_c.tolist() == o.contains(q.tolist())


def test_growth_dedup_assign_erase_vs_oracle():
    rng = np.random.default_rng(21)
    d = mx().DynamicMap(100000, EK, EV, RK)
    o = DynamicMapOracle(100000, EK, RK)
    same_state(d, o)
    universe = np.arange(0, 3_000_000, dtype=np.int64) * 7919 + 3
    for step in range(6):
        n = int(rng.integers(20000, 400000))
        keys = rng.choice(universe, n)                   # duplicates and keys already present
        keys[rng.random(n) < 0.01] = EK                  # reserved key: never inserted
        vals = keys * 7 + 1
        assert d.insert(pairs_of(keys, vals)) == o.insert(keys.tolist(), vals.tolist())
        same_state(d, o)
    assert d.num_submaps() > 1
    q = np.concatenate([rng.choice(universe, 200000), rng.integers(-10**6, 0, 1000)])
    check_queries(d, o, q)
    # InsertOrAssign: existing keys take new values, new keys count as inserted
    keys = rng.choice(universe, 300000)
    vals = keys * 3 + 5
    assert d.insert_or_assign(pairs_of(keys, vals)) == o.insert_or_assign(keys.tolist(), vals.tolist())
    same_state(d, o)
    check_queries(d, o, q)
    # Erase, then re-insert some erased keys
    er = rng.choice(universe, 150000)
    assert d.erase(torch.from_numpy(er).cuda()) == o.erase(er.tolist())
    same_state(d, o)
    check_queries(d, o, np.concatenate([q, er[:50000]]))
    back = er[:40000]
    assert d.insert(pairs_of(back, back * 7 + 1)) == o.insert(back.tolist(), (back * 7 + 1).tolist())
    same_state(d, o)
    check_queries(d, o, np.concatenate([q, er[:50000]]))


def test_async_append_reserve_clear_vs_oracle():
    rng = np.random.default_rng(22)
    d = mx().DynamicMap(50000, EK, EV, RK)
    o = DynamicMapOracle(50000, EK, RK)
    d.reserve(1_000_000)
    o.reserve(1_000_000)
    same_state(d, o)
    k = np.arange(1, 300001, dtype=np.int64) * 13
    assert d.insert(pairs_of(k, k + 1), mode=mx().DM_APPEND_UNIQUE) == o.insert(k.tolist(), (k + 1).tolist(), "append")
    same_state(d, o)
    k2 = np.arange(300001, 700001, dtype=np.int64) * 13
    assert d.insert_async(pairs_of(k2, k2 + 1)) == o.insert(k2.tolist(), (k2 + 1).tolist(), sync=False)
    torch.cuda.synchronize()
    same_state(d, o)
    q = np.concatenate([k[::7], k2[::5], rng.integers(-10**6, 0, 1000)])
    check_queries(d, o, q)
    d.clear()
    o.clear()
    same_state(d, o)
    check_queries(d, o, q[:1000])
    assert d.insert(pairs_of(k[:1000], k[:1000])) == o.insert(k[:1000].tolist(), k[:1000].tolist())
    same_state(d, o)


def test_zero_capacity_and_errors():
    d = mx().DynamicMap(0, EK, EV, RK)
    o = DynamicMapOracle(0, EK, RK)
    same_state(d, o)
    one = np.array([7], np.int64)
    assert d.insert(pairs_of(one, one + 1)) == o.insert([7], [8]) == 1
    same_state(d, o)
    assert d.find(torch.from_numpy(one).cuda()).cpu().tolist() == [8]
    assert d.insert(None, n=5) == 0
    with pytest.raises(ValueError):
        mx().DynamicMap(10, -1, -1, -1)
