"""The slot array behind Data(), read as the reference's tests read it
(DumpTable, tests/common/dump_table.h:20-57, compared with EqualAsMap against
the std::unordered_map golden of tests/common/golden.h): after Insert,
InsertOrAssign and Erase the live slots equal the golden map, for every
Pair<Key, Value> layout the library compiles (4/8-byte keys and values) and
for sets, with the erased sentinel set as the reference's erase tests set it
(tests/static_map/erase_test.cpp:95-213)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TD = {np.int32: torch.int32, np.int64: torch.int64, np.float32: torch.float32, np.float64: torch.float64}


def mx():
    import mxco
    return mxco


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def golden_insert(keys, values):
    """First occurrence wins (values are a function of the key, so the
    winner among duplicates does not matter)."""
    g = {}
    for k, v in zip(keys.tolist(), values.tolist()):
        g.setdefault(k, v)
    return g


@pytest.mark.parametrize("kd,vd", [(np.int32, np.int32), (np.int64, np.int64), (np.int32, np.int64),
                                   (np.int64, np.int32), (np.int64, np.float32), (np.int32, np.float64)])
@pytest.mark.parametrize("probing", ["linear", "double"])
def test_map_dump_after_insert_assign_erase(kd, vd, probing):
    mxco = mx()
    rng = np.random.default_rng(7)
    n = 40000
    uniq = rng.choice(1 << 30, n, replace=False).astype(kd)
    keys = np.concatenate([uniq, uniq[rng.integers(0, n, n // 4)]])  # a quarter duplicated
    vals = (keys.astype(np.int64) * 7 + 1).astype(vd)
    m = mxco.StaticMap(2 * n, -1, -1, TD[kd], TD[vd], probing=probing)
    m.set_erased_key(-2)
    pairs = mxco.make_pairs(dev(keys), dev(vals), TD[kd], TD[vd])
    assert m.insert(pairs) == n // 4  # duplicates count as failed inserts
    gold = golden_insert(keys, vals)

    # InsertOrAssign of existing and new keys: last writer wins per key
    akeys = np.concatenate([uniq[: n // 8], rng.choice(np.arange(1 << 30, (1 << 30) + n), n // 8, replace=False)])
    akeys = akeys.astype(kd)
    avals = (akeys.astype(np.int64) * 3 + 5).astype(vd)
    m.insert_or_assign(mxco.make_pairs(dev(akeys), dev(avals), TD[kd], TD[vd]))
    for k, v in zip(akeys.tolist(), avals.tolist()):
        gold[k] = v

    # erase the even-indexed keys, then the odd-indexed ones of a prefix
    for sel in (uniq[0::2][: n // 4], uniq[1::2][: n // 8]):
        assert m.erase(dev(sel)) == 0
        for k in sel.tolist():
            gold.pop(k, None)

    dk, dv = m.dump()
    assert len(dk) == len(gold) == len(set(dk.tolist()))
    assert dict(zip(dk.tolist(), dv.tolist())) == gold


@pytest.mark.parametrize("kd", [np.int32, np.int64])
def test_set_dump_after_insert_erase(kd):
    mxco = mx()
    rng = np.random.default_rng(8)
    n = 30000
    uniq = rng.choice(1 << 30, n, replace=False).astype(kd)
    keys = np.concatenate([uniq, uniq[: n // 3]])
    s = mxco.StaticSet(2 * n, -1, TD[kd])
    s.set_erased_key(-2)
    assert s.insert(dev(keys)) == n // 3
    assert s.erase(dev(uniq[::3])) == 0
    gold = set(uniq.tolist()) - set(uniq[::3].tolist())
    d = s.dump()
    assert len(d) == len(gold) and set(d.tolist()) == gold


def test_dump_skips_only_sentinels():
    """A key that differs from the empty sentinel in one byte is live; after
    Clear the dump is empty."""
    mxco = mx()
    m = mxco.StaticMap(1000, -1, -1, torch.int64, torch.int64)
    k = np.array([-1 ^ (1 << 40), 0, 1], np.int64)
    m.insert(mxco.make_pairs(dev(k), dev(k), torch.int64, torch.int64))
    dk, dv = m.dump()
    assert sorted(dk.tolist()) == sorted(k.tolist()) and sorted(dv.tolist()) == sorted(k.tolist())
    m.clear()
    dk, dv = m.dump()
    assert len(dk) == 0 and len(dv) == 0
