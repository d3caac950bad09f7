"""StaticMap / StaticSet on the GPU through the C ABI, checked against the
oracle on the same seeded inputs (Find/Contains/InsertAndFind outputs and
failure counts bit-exact), plus the reference's negative tests."""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

GROUPS = [1, 2, 4, 8]
PROBE = {O.PROBE_LINEAR: "linear", O.PROBE_DOUBLE: "double"}
HASH = {O.HASH_MURMUR3_32: "murmurhash3_32", O.HASH_XXHASH_32: "xxhash_32", O.HASH_XXHASH_64: "xxhash_64"}
TD = {np.int32: torch.int32, np.int64: torch.int64, np.float32: torch.float32, np.float64: torch.float64}


def mx():
    import mxco
    return mxco


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def dev_pairs(keys, values):
    return dev(O.pack_pairs(keys, values))


def distinct_keys(rng, n, hi, dtype):
    return rng.choice(hi, n, replace=False).astype(dtype)


def absent_keys(rng, n, hi, present, dtype):
    out = []
    pset = present
    while sum(len(x) for x in out) < n:
        c = rng.integers(0, hi, 2 * n).astype(dtype)
        c = c[~np.isin(c, pset)]
        out.append(c)
    return np.concatenate(out)[:n]


def i32_map_2p20_inputs():
    """Plain StaticMap int32->int32: 2^19 distinct keys in [0, 2^31-2],
    value = key*7+1 in uint32 arithmetic, 2^19 queries half present, shuffled."""
    rng = np.random.default_rng(0)
    n = 1 << 19
    keys = distinct_keys(rng, n, 2**31 - 1, np.int32)
    vals = (keys.astype(np.uint32) * np.uint32(7) + np.uint32(1)).astype(np.int32)
    q = np.concatenate([rng.choice(keys, n // 2, replace=False), absent_keys(rng, n // 2, 2**31 - 1, keys, np.int32)])
    rng.shuffle(q)
    return keys, vals, q


@pytest.mark.parametrize("group", GROUPS)
def test_i32_map_2p20_find_contains_bit_exact(group):
    keys, vals, q = i32_map_2p20_inputs()
    ref = O.OpenAddressing(np.int32, np.int32, 1 << 20, -1, -1)
    assert ref.insert(keys, vals) == 0
    exp_v, exp_c = ref.find(q)
    m = mx().StaticMap(1 << 20, -1, -1, torch.int32, torch.int32, group_size=group)
    assert m.capacity() == 1048580
    assert m.insert(dev_pairs(keys, vals)) == 0
    out = m.find(dev(q)).cpu().numpy()
    con = m.contains(dev(q)).cpu().numpy()
    assert out.tobytes() == exp_v.tobytes()
    assert con.tobytes() == exp_c.astype(np.uint8).tobytes()
    assert m.count(dev(q)) == int(exp_c.sum())


LAYOUTS = [(np.int32, np.int32), (np.int64, np.int64), (np.int32, np.int64), (np.int64, np.int32),
           (np.int32, None), (np.int64, None)]


@pytest.mark.parametrize("kd,vd", LAYOUTS, ids=lambda x: getattr(x, "__name__", str(x)))
@pytest.mark.parametrize("probing", [O.PROBE_LINEAR, O.PROBE_DOUBLE])
@pytest.mark.parametrize("bucket", [1, 5, 8, 17])
def test_layouts_vs_oracle(kd, vd, probing, bucket):
    rng = np.random.default_rng(bucket * 10 + probing)
    n = 30000
    keys = rng.integers(0, 40000, n).astype(kd)  # duplicates on purpose
    keys[::101] = -1                              # reserved key
    vals = None if vd is None else (keys.astype(np.int64) * 3 + 11).astype(vd)  # value = f(key)
    hashf = O.HASH_MURMUR3_32 if vd is not None else O.HASH_XXHASH_32
    ref = O.OpenAddressing(kd, vd, 40000, -1, -1 if vd is not None else None, bucket=bucket, probing=probing,
                           hash_family=hashf)
    exp_failed = ref.insert(keys, vals)
    q = np.concatenate([keys[:5000], rng.integers(-100, 80000, 10000).astype(kd)])
    exp_v, exp_c = ref.find(q)
    for g in GROUPS:
        if vd is None:
            m = mx().StaticSet(40000, -1, TD[kd], bucket_size=bucket, probing=PROBE[probing], hash=HASH[hashf],
                               group_size=g)
            pairs = dev(keys)
        else:
            m = mx().StaticMap(40000, -1, -1, TD[kd], TD[vd], bucket_size=bucket, probing=PROBE[probing],
                               hash=HASH[hashf], group_size=g)
            pairs = dev_pairs(keys, vals)
        assert m.capacity() == ref.capacity
        assert m.insert(pairs) == exp_failed, g
        out = m.find(dev(q)).cpu().numpy()
        con = m.contains(dev(q)).cpu().numpy()
        assert out.tobytes() == exp_v.tobytes(), g
        assert con.tobytes() == exp_c.astype(np.uint8).tobytes(), g


@pytest.mark.parametrize("group", GROUPS)
def test_insert_or_assign_counts_and_values(group):
    rng = np.random.default_rng(11)
    keys = distinct_keys(rng, 20000, 2**40, np.int64)
    vals = keys * 2
    m = mx().StaticMap(40000, -1, -1, torch.int64, torch.int64, group_size=group)
    assert m.insert(dev_pairs(keys, vals)) == 0
    upd = keys[:8000]
    newv = upd * 5 + 3
    extra = distinct_keys(rng, 3000, 2**40, np.int64) + 2**41
    f, a = m.insert_or_assign(dev_pairs(np.concatenate([upd, extra]), np.concatenate([newv, extra])))
    assert (f, a) == (0, 8000)
    ref = O.OpenAddressing(np.int64, np.int64, 40000, -1, -1)
    ref.insert(keys, vals)
    assert ref.insert_or_assign(np.concatenate([upd, extra]), np.concatenate([newv, extra])) == (0, 8000)
    q = np.concatenate([keys, extra, extra + 1])
    assert m.find(dev(q)).cpu().numpy().tobytes() == ref.find(q)[0].tobytes()


@pytest.mark.parametrize("kd,vd", [(np.int32, np.int32), (np.int64, np.int64), (np.int32, None)],
                         ids=["i32i32", "i64i64", "set32"])
@pytest.mark.parametrize("group", GROUPS)
def test_insert_and_find(kd, vd, group):
    rng = np.random.default_rng(12)
    keys = rng.integers(0, 5000, 20000).astype(kd)
    vals = None if vd is None else rng.integers(0, 10**6, 20000).astype(vd)
    if vd is None:
        m = mx().StaticSet(30000, -1, TD[kd], group_size=group)
        found, ins = m.insert_and_find(dev(keys))
    else:
        m = mx().StaticMap(30000, -1, -1, TD[kd], TD[vd], group_size=group)
        found, ins = m.insert_and_find(dev_pairs(keys, vals))
    found, ins = found.cpu().numpy(), ins.cpu().numpy()
    uniq = np.unique(keys)
    # exactly one inserting input per distinct key; every duplicate sees the winner's value
    assert ins.sum() == len(uniq)
    win = {}
    for i in np.nonzero(ins)[0]:
        win[keys[i]] = found[i] if vd is None else vals[i]
        if vd is not None:
            assert found[i] == vals[i]
    for k, f in zip(keys.tolist(), found.tolist()):
        assert f == win[k]
    # distinct-key batch: exactly the oracle
    dk = distinct_keys(rng, 5000, 10**6, kd) + 10**7
    dv = None if vd is None else (dk * 3).astype(vd)
    ref = O.OpenAddressing(kd, vd, 30000, -1, None if vd is None else -1,
                           probing=O.PROBE_DOUBLE if vd is None else O.PROBE_LINEAR,
                           hash_family=O.HASH_XXHASH_32 if vd is None else O.HASH_MURMUR3_32)
    ef, ei = ref.insert_and_find(dk, dv)
    f2, i2 = m.insert_and_find(dev(dk) if vd is None else dev_pairs(dk, dv))
    assert f2.cpu().numpy().tobytes() == ef.tobytes()
    assert (i2.cpu().numpy() == ei).all()


def test_full_table_and_reserved_key():
    m = mx().StaticMap(10, -1, -1)
    cap = m.capacity()
    keys = np.arange(cap + 7, dtype=np.int32)
    assert m.insert(dev_pairs(keys, keys)) == 7
    found = m.find(dev(keys)).cpu().numpy()
    assert (found[found != -1] == keys[found != -1]).all() and (found != -1).sum() == cap
    m2 = mx().StaticMap(10, -1, -1)
    assert m2.insert(dev_pairs(np.array([-1, -1, 3], np.int32), np.array([1, 2, 3], np.int32))) == 2
    # miss after a full cycle of a full table
    assert not m.contains(dev(np.array([10**6], np.int32))).cpu().numpy()[0]


def test_null_and_empty_batches():
    m = mx().StaticMap(100, -1, -1)
    assert m.insert(0, n=5) == 5            # null pointer: every input counted as failed
    assert m.insert(torch.empty((0, 8), dtype=torch.uint8, device="cuda")) == 0
    f, a = m.insert_or_assign(0, n=4)
    assert (f, a) == (4, 0)
    assert m.count(0, n=3) == 0
    out = torch.full((4,), 7, dtype=torch.int32, device="cuda")
    m.find(0, out=out, n=4)                   # silent no-op
    assert (out.cpu() == 7).all()


def test_oversized_batch_raises_length_error():
    import mxco
    m = mx().StaticMap(100, -1, -1)
    buf = torch.zeros(16, dtype=torch.uint8, device="cuda")
    with pytest.raises(mxco.LengthError):
        m.insert(buf, n=1 << 32)
    with pytest.raises(mxco.LengthError):
        m.find(buf, out=buf, n=1 << 32)


@pytest.mark.parametrize("group", GROUPS)
def test_erase_with_and_without_tombstone(group):
    keys = np.arange(1, 3001, dtype=np.int32) * 13
    m = mx().StaticMap(6000, -1, -1, group_size=group)
    m.insert(dev_pairs(keys, keys))
    m.set_erased_key(-2)
    assert m.erase(dev(keys[:1000])) == 0
    assert m.erase(dev(keys[:1000])) == 1000
    con = m.contains(dev(keys)).cpu().numpy()
    assert not con[:1000].any() and con[1000:].all()
    ref = O.OpenAddressing(np.int32, np.int32, 6000, -1, -1)
    ref.insert(keys, keys)
    assert ref.erase(keys[:1000], erased_key=-2) == 0
    # Plain erase (no tombstone) empties the slot, which can cut later probe
    # chains exactly as in the reference; compare with the oracle on the GPU's
    # own slot layout.
    s = mx().StaticSet(6000, -1, group_size=group)
    s.insert(dev(keys))
    ref = O.OpenAddressing(np.int32, None, 6000, -1, None, probing=O.PROBE_DOUBLE, hash_family=O.HASH_XXHASH_32)
    ref.load_slots(s.slots().cpu().numpy())
    assert s.erase(dev(keys[:300])) == ref.erase(keys[:300]) == 0
    assert s.contains(dev(keys)).cpu().numpy().tobytes() == ref.find(keys)[1].astype(np.uint8).tobytes()
    np.testing.assert_array_equal(s.slots().cpu().numpy(), ref.slots())


@pytest.mark.parametrize("group", GROUPS)
@pytest.mark.parametrize("probing", [O.PROBE_LINEAR, O.PROBE_DOUBLE])
@pytest.mark.parametrize("kd", [np.int64, np.int32])
def test_probe_stats_match_oracle_line_count(group, probing, kd):
    """The roofline's algorithmic bytes use the lines Find's walk reads; the
    kernel's count equals the oracle's replay on the same slot array (16-byte
    slots read one per vector, 8-byte slots in pairs)."""
    rng = np.random.default_rng(15)
    m = mx().StaticMap(200000, -1, -1, TD[kd], TD[kd], probing=PROBE[probing], group_size=group)
    n = int(0.8 * m.capacity())
    hi = 2**62 if kd == np.int64 else 2**30
    keys = distinct_keys(rng, n, hi, kd)
    m.insert(dev_pairs(keys, keys))
    ref = O.OpenAddressing(kd, kd, 200000, -1, -1, probing=probing)
    ref.load_slots(m.slots().cpu().numpy())
    q = np.concatenate([keys[:50000], rng.integers(hi, 2 * hi - 1, 50000).astype(kd)])
    lines, slots = m.probe_stats(dev(q))
    _, _, exp_lines = ref.find(q, lines=True)
    if probing == O.PROBE_LINEAR:
        assert lines == exp_lines
    else:
        # A double-hashing walk can come back to a line it read two or more
        # buckets earlier; the kernel dedups consecutive windows only.
        assert exp_lines <= lines <= exp_lines * 1.001


@pytest.mark.parametrize("group", GROUPS)
def test_stencil_predicates(group):
    rng = np.random.default_rng(13)
    keys = distinct_keys(rng, 10000, 10**7, np.int32)
    st = rng.integers(0, 100, 10000).astype(np.int32)
    m = mx().StaticMap(20000, -1, -1, group_size=group)
    assert m.insert_if(dev_pairs(keys, keys), dev(st), "odd") == 0
    con = m.contains(dev(keys)).cpu().numpy()
    np.testing.assert_array_equal(con, st % 2 == 1)
    out = m.find(dev(keys), stencil=dev(st), pred="even").cpu().numpy()
    assert (out == -1).all()
    out = m.find(dev(keys), stencil=dev(st), pred="odd").cpu().numpy()
    np.testing.assert_array_equal(out, np.where(st % 2 == 1, keys, -1))
    c = m.contains(dev(keys), stencil=dev(st), pred="even").cpu().numpy()
    assert not c.any()


ASSIGN_PRED = {"always": lambda st: np.ones(len(st), bool), "nonzero": lambda st: st != 0,
               "odd": lambda st: st % 2 == 1, "even": lambda st: st % 2 == 0}


@pytest.mark.parametrize("kd,vd", LAYOUTS[:4], ids=lambda x: getattr(x, "__name__", str(x)))
@pytest.mark.parametrize("probing", [O.PROBE_LINEAR, O.PROBE_DOUBLE])
@pytest.mark.parametrize("group", GROUPS)
def test_assign_if_vs_oracle(kd, vd, probing, group):
    """AssignIf (static_map.h:189): present selected keys get the new payload,
    absent keys stay absent.  From the same slots, the table after each
    AssignIf equals the oracle's byte for byte."""
    rng = np.random.default_rng(17 + group)
    cap_req = 30000
    keys = distinct_keys(rng, 20000, 10**8, kd)
    vals = rng.integers(0, 2**30, len(keys)).astype(vd)
    m = mx().StaticMap(cap_req, -1, -1, TD[kd], TD[vd], probing=PROBE[probing], group_size=group)
    ref = O.OpenAddressing(kd, vd, cap_req, -1, -1, probing=probing)
    assert m.insert(dev_pairs(keys, vals)) == ref.insert(keys, vals) == 0
    # concurrent Insert places keys in its own order: start both from the GPU's slots
    ref.load_slots(m.slots().cpu().numpy())
    q = np.concatenate([keys[rng.permutation(len(keys))[:12000]], absent_keys(rng, 6000, 10**8, keys, kd),
                        np.full(3, -1, kd)])
    rng.shuffle(q)
    nv = rng.integers(0, 2**30, len(q)).astype(vd)
    st = rng.integers(-3, 50, len(q)).astype(np.int32)
    for pred, sel in ASSIGN_PRED.items():
        nv = nv + np.asarray(1, vd)
        m.assign_if_async(dev_pairs(q, nv), dev(st), pred)
        torch.cuda.synchronize()
        assert ref.assign_if(q, nv, sel(st)) > 0
        assert m.slots().cpu().numpy().tobytes() == ref.slots().tobytes(), pred
    probe = np.concatenate([keys, absent_keys(rng, 5000, 10**8, keys, kd)])
    rv, rf = ref.find(probe)
    assert m.find(dev(probe)).cpu().numpy().tobytes() == rv.tobytes()
    assert m.contains(dev(probe)).cpu().numpy().astype(bool).tolist() == rf.tolist()


def test_assign_if_null_and_empty():
    m = mx().StaticMap(100, -1, -1)
    m.assign_if_async(dev_pairs(np.array([1], np.int32), np.array([2], np.int32)), None, "always", n=0)
    k = np.array([5, 6], np.int32)
    assert m.insert(dev_pairs(k, k)) == 0
    m.assign_if_async(dev_pairs(k, k * 10), None, "always")  # no stencil: every pair
    torch.cuda.synchronize()
    assert m.find(dev(k)).cpu().numpy().tolist() == [50, 60]


@pytest.mark.parametrize("group", GROUPS)
def test_insert_stats_counts_compare_and_swaps(group):
    """The retry-counting Insert: every distinct key lands with exactly one
    successful compare-and-swap, so made - lost = keys inserted; the table
    equals Insert's."""
    rng = np.random.default_rng(23)
    keys = distinct_keys(rng, 300000, 2**31 - 1, np.int32)
    dup = np.concatenate([keys, keys[:50000]])
    rng.shuffle(dup)
    m = mx().StaticMap(400000, -1, -1, group_size=group)
    lost, made = m.insert_stats(dev_pairs(dup, dup))
    assert 0 <= lost < made
    assert made - lost >= len(keys)  # plus duplicates caught by a lost compare-and-swap
    assert m.count(dev(keys)) == len(keys)
    assert np.array_equal(m.find(dev(keys)).cpu().numpy(), keys)
    m2 = mx().StaticMap(400000, -1, -1, group_size=group)
    lost, made = m2.insert_stats(dev_pairs(keys, keys))
    assert made - lost == len(keys)


def test_float_keys_use_operator_equal():
    # KeyEqual is == on float (so -0.0 equals +0.0 once the walk reaches it);
    # the sentinel test stays bitwise and the probe hash is over the bytes.
    keys = np.array([0.0, 1.5, -2.25, 7.0, -0.0], np.float32)
    vals = np.array([1, 2, 3, 4, 5], np.int32)
    q = np.array([-0.0, 0.0, 1.5, 3.0, -2.25], np.float32)
    for g in GROUPS:
        m = mx().StaticMap(100, float("nan"), -1, torch.float32, torch.int32, group_size=g)
        ref = O.OpenAddressing(np.float32, np.int32, 100, float("nan"), -1)
        assert m.insert(dev_pairs(keys, vals)) == ref.insert(keys, vals)
        assert m.find(dev(q)).cpu().numpy().tobytes() == ref.find(q)[0].tobytes()


@pytest.mark.parametrize("group", [1, 2, 8])
def test_large_table_properties(group):
    """A 2^26-slot int64 map at load factor 0.8: every inserted key is found
    with value f(key), every absent key misses, Count agrees."""
    g = torch.Generator(device="cuda")
    g.manual_seed(14)
    m = mx().StaticMap(1 << 26, -1, -1, torch.int64, torch.int64, group_size=group)
    n = int(0.8 * m.capacity())
    keys = torch.unique(torch.randint(0, 2**62, (n,), device="cuda", dtype=torch.int64, generator=g))
    n = keys.numel()
    vals = keys * 7 + 1
    assert m.insert(mx().make_pairs(keys, vals)) == 0
    out = m.find(keys)
    assert torch.equal(out, vals)
    absent = keys + 2**62  # disjoint from the inserted range
    assert m.count(absent) == 0
    assert m.count(keys) == n


@pytest.mark.parametrize("group", GROUPS)
def test_zipf_duplicate_heavy_insert(group):
    """Duplicate-heavy stream (2^20 draws of Zipf(s=1.1) over 2^14 values):
    the not-inserted count and every Find equal the oracle's sequential
    insert of the same stream (value = f(key), so racing duplicates agree)."""
    from mxco.workloads import Zipf
    z = Zipf(1 << 14, 1.1)
    keys_t = Zipf.keys(z.draw(1 << 20, seed=3))
    keys = keys_t.cpu().numpy()
    vals = keys * 7 + 1
    ref = O.OpenAddressing(np.int64, np.int64, 1 << 16, -1, -1)
    ref_failed = ref.insert(keys, vals)
    assert ref_failed == len(keys) - len(np.unique(keys))
    m = mx().StaticMap(1 << 16, -1, -1, torch.int64, torch.int64, group_size=group)
    assert m.insert(dev_pairs(keys, vals)) == ref_failed
    q = np.concatenate([np.unique(keys), Zipf.keys(torch.arange(1 << 14, 1 << 15)).numpy()])
    ref_out, ref_con = ref.find(q)
    assert np.array_equal(m.find(dev(q)).cpu().numpy(), ref_out)
    assert np.array_equal(m.contains(dev(q)).cpu().numpy(), ref_con)
