"""Insert of a large batch (n >= 2^20 and n >= capacity / 2, LinearProbing)
on a table of more than 256 MiB of slots is built region by region in LDS
(MXCO_OA_REGION_INSERT=0 turns it off; these tests lower the size threshold).
Its failures and Find outputs equal the oracle's sequential CAS walk on the
same inputs, and the table holds every inserted key exactly once: on empty tables, tables of 2^24 slots (a two-level
partition), an int32 set of 2^29 + 2^26 slots (256 regions per first-level bin, the most), tables already holding keys and tombstones, streams with one
very hot key (regions left to the CAS walk), full tables, InsertAsync and
graph replays."""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def region_insert_on(monkeypatch):
    # these tables are smaller than the default 256 MiB threshold
    monkeypatch.setenv("MXCO_OA_REGION_INSERT", "1")
    monkeypatch.setenv("MXCO_OA_REGION_INSERT_MIN_BYTES", "0")

TD = {np.int32: torch.int32, np.int64: torch.int64}


def mx():
    import mxco
    return mxco


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def dev_pairs(keys, values):
    return dev(O.pack_pairs(keys, values))


def hi_of(kd):
    return 2**30 if kd == np.int32 else 2**62


def distinct(rng, n, kd, exclude=None):
    k = np.unique(rng.integers(0, hi_of(kd), int(n * 1.05) + 16))
    if exclude is not None:
        k = k[~np.isin(k, exclude)]
    rng.shuffle(k)
    assert len(k) >= n
    return k[:n].astype(kd)


def value_of(keys, vd):
    return ((keys.astype(np.int64) * 7 + 1) & 0x3fffffff).astype(vd)


def misses(rng, n, kd):
    return rng.integers(hi_of(kd), 2 * hi_of(kd) - 1, n).astype(kd)


def table_keys(m, kd):
    s = m.slots().cpu().numpy()
    k = np.ascontiguousarray(s[:, :np.dtype(kd).itemsize]).view(kd).reshape(-1)
    return k[(k != kd(-1)) & (k != kd(-2))]


def check_against(m, ref, kd, vd, q):
    out = m.find(dev(q)).cpu().numpy()
    rv, rf = ref.find(q)
    assert out.tobytes() == rv.tobytes()
    assert m.contains(dev(q)).cpu().numpy().astype(bool).tolist() == rf.tolist()


@pytest.mark.parametrize("kd,vd", [(np.int64, np.int64), (np.int32, np.int32), (np.int64, np.int32),
                                   (np.int32, np.int64)])
def test_region_insert_matches_oracle(kd, vd):
    rng = np.random.default_rng(1)
    cap_req = 1 << 21
    m = mx().StaticMap(cap_req, -1, -1, TD[kd], TD[vd])
    cap = m.capacity()
    k = distinct(rng, int(0.8 * cap), kd)
    allk = np.concatenate([k, k[rng.integers(0, len(k), len(k) // 10)], np.full(5, -1, kd)])
    rng.shuffle(allk)
    assert m.insert_uses_region_build(len(allk))
    failed = m.insert(dev_pairs(allk, value_of(allk, vd)))
    ref = O.OpenAddressing(kd, vd, cap_req, -1, -1)
    assert failed == ref.insert(allk, value_of(allk, vd)) == len(allk) - len(k)
    got = table_keys(m, kd)
    assert len(got) == len(k) and np.array_equal(np.sort(got), np.sort(k))
    check_against(m, ref, kd, vd, np.concatenate([k[:200000], misses(rng, 200000, kd)]))


@pytest.mark.parametrize("kd", [np.int32, np.int64])
def test_region_insert_sets_match_oracle(kd):
    """StaticSet under LinearProbing: 4- and 8-byte slots, whose regions walk
    (their sort scratch does not fit the region's LDS)."""
    rng = np.random.default_rng(8)
    cap_req = 1 << 21
    s = mx().StaticSet(cap_req, -1, TD[kd], probing="linear")
    cap = s.capacity()
    k = distinct(rng, int(0.8 * cap), kd)
    allk = np.concatenate([k, k[rng.integers(0, len(k), len(k) // 10)], np.full(3, -1, kd)])
    rng.shuffle(allk)
    assert s.insert_uses_region_build(len(allk))
    failed = s.insert(dev(allk))
    ref = O.OpenAddressing(kd, None, cap_req, -1, None, probing=O.PROBE_LINEAR, hash_family=O.HASH_XXHASH_32)
    assert failed == ref.insert(allk) == len(allk) - len(k)
    got = table_keys(s, kd)
    assert len(got) == len(k) and np.array_equal(np.sort(got), np.sort(k))
    q = np.concatenate([k[:200000], misses(rng, 200000, kd)])
    rv, rf = ref.find(q)
    assert s.find(dev(q)).cpu().numpy().tobytes() == rv.tobytes()
    assert s.contains(dev(q)).cpu().numpy().astype(bool).tolist() == rf.tolist()


def test_region_insert_two_level_partition():
    """2^24 + 2^20 slots of 16 bytes: 4353 regions, grouped two per bin."""
    kd = vd = np.int64
    rng = np.random.default_rng(2)
    cap_req = (1 << 24) + (1 << 20)
    m = mx().StaticMap(cap_req, -1, -1, TD[kd], TD[vd])
    cap = m.capacity()
    k = distinct(rng, int(0.85 * cap), kd)
    assert m.insert_uses_region_build(len(k))
    assert m.insert(dev_pairs(k, value_of(k, vd))) == 0
    ref = O.OpenAddressing(kd, vd, cap_req, -1, -1)
    assert ref.insert(k, value_of(k, vd)) == 0
    assert len(table_keys(m, kd)) == len(k)
    check_against(m, ref, kd, vd, np.concatenate([k[::7], misses(rng, 500000, kd)]))


def test_region_insert_most_regions_per_bin():
    """2^29 + 2^26 slots of an int32 set: 147,457 regions of 4096 slots, so
    the first level groups the most regions a bin may hold (256; 577 bins).
    At this size the checks are properties: no failures, every key found,
    no miss found, and the table holds exactly the inserted keys."""
    s = mx().StaticSet((1 << 29) + (1 << 26), -1, torch.int32, probing="linear")
    cap = s.capacity()
    n = int(0.6 * cap)
    # i * odd mod 2^30 is a bijection of [0, 2^30): n distinct keys, none -1
    k = ((torch.arange(n, dtype=torch.int64, device="cuda") * 2654435761) % (1 << 30)).to(torch.int32)
    assert s.insert_uses_region_build(n)
    assert s.insert(k) == 0
    assert bool(s.contains(k).bool().all())
    miss = torch.randint(1 << 30, (1 << 31) - 1, (1 << 22,), dtype=torch.int64, device="cuda").to(torch.int32)
    assert not bool(s.contains(miss).bool().any())
    t = s.slots().view(torch.int32).reshape(-1)
    t = t[t != -1]
    assert t.numel() == n
    assert torch.equal(torch.sort(t).values, torch.sort(k).values)
    del t


def test_region_insert_into_table_with_keys_and_tombstones():
    kd = vd = np.int64
    rng = np.random.default_rng(3)
    cap_req = 1 << 21
    m = mx().StaticMap(cap_req, -1, -1, TD[kd], TD[vd])
    ref = O.OpenAddressing(kd, vd, cap_req, -1, -1)
    cap = m.capacity()
    first = distinct(rng, int(0.3 * cap), kd)
    # small batches: the CAS walk
    for c in np.array_split(first, 4):
        assert not m.insert_uses_region_build(len(c))
        assert m.insert(dev_pairs(c, value_of(c, vd))) == 0
        assert ref.insert(c, value_of(c, vd)) == 0
    m.set_erased_key(-2)
    gone = first[:len(first) // 5]
    assert m.erase(dev(gone)) == 0
    assert ref.erase(gone, erased_key=-2) == 0
    # one large batch: new keys, present keys (duplicates) and erased keys
    new = distinct(rng, int(0.5 * cap), kd, exclude=first)
    big = np.concatenate([new, first[-50000:], gone[:50000]])
    rng.shuffle(big)
    assert m.insert_uses_region_build(len(big))
    f = m.insert(dev_pairs(big, value_of(big, vd)))
    assert f == ref.insert(big, value_of(big, vd)) == 50000
    q = np.concatenate([first, new[:300000], misses(rng, 100000, kd)])
    check_against(m, ref, kd, vd, q)


def test_region_insert_hot_key():
    """One key 3M times among 1M distinct: its region goes to the CAS walk."""
    kd = vd = np.int64
    rng = np.random.default_rng(4)
    cap_req = 1 << 22
    m = mx().StaticMap(cap_req, -1, -1, TD[kd], TD[vd])
    k = distinct(rng, 1 << 20, kd)
    allk = np.concatenate([k, np.full(3 << 20, k[0], kd)])
    rng.shuffle(allk)
    assert m.insert_uses_region_build(len(allk))
    assert m.insert(dev_pairs(allk, value_of(allk, vd))) == 3 << 20
    ref = O.OpenAddressing(kd, vd, cap_req, -1, -1)
    ref.insert(allk, value_of(allk, vd))
    assert len(table_keys(m, kd)) == len(k)
    check_against(m, ref, kd, vd, np.concatenate([k, misses(rng, 100000, kd)]))


@pytest.mark.parametrize("kd,vd", [(np.int32, np.int32), (np.int64, np.int64)])
def test_region_insert_full_table(kd, vd):
    rng = np.random.default_rng(5)
    m = mx().StaticMap(1 << 20, -1, -1, TD[kd], TD[vd])
    cap = m.capacity()
    # every failing key walks the whole table (as in the CAS walk): a few thousand over
    k = distinct(rng, cap + 3000, kd)
    assert m.insert_uses_region_build(len(k))
    assert m.insert(dev_pairs(k, value_of(k, vd))) == len(k) - cap
    got = table_keys(m, kd)
    assert len(got) == cap and len(np.unique(got)) == cap and np.isin(got, k).all()
    found = m.contains(dev(k)).cpu().numpy().astype(bool)
    assert np.array_equal(np.sort(k[found]), np.sort(got))


def test_region_insert_after_clear_and_captured_insert_on_filled_table():
    """An Insert right after Clear builds its regions from the empty slot; an
    Insert captured on an empty table and replayed after other keys went in
    still reads the regions, so those keys stay."""
    import mxco
    kd = vd = np.int64
    rng = np.random.default_rng(7)
    m = mxco.StaticMap(1 << 21, -1, -1, TD[kd], TD[vd])
    cap = m.capacity()
    a = distinct(rng, int(0.6 * cap), kd)
    b = distinct(rng, 200000, kd, exclude=a)
    pa, pb = dev_pairs(a, value_of(a, vd)), dev_pairs(b, value_of(b, vd))
    assert m.insert(pb) == 0
    m.clear()
    assert m.insert_uses_region_build(len(a))
    assert m.insert(pa) == 0  # after Clear: regions start empty
    assert not m.contains(dev(b)).cpu().numpy().any()
    assert np.array_equal(m.find(dev(a)).cpu().numpy(), value_of(a, vd))
    m.clear()
    g = mxco.Graph(lambda s: m.insert_async(pa, stream=s))  # captured while the table is empty
    for rep in range(2):
        m.clear()
        assert m.insert(pb) == 0  # small batch: the CAS walk
        g.launch()
        g.stream.synchronize()
        assert m.contains(dev(b)).cpu().numpy().all(), rep
        assert np.array_equal(m.find(dev(a)).cpu().numpy(), value_of(a, vd)), rep


def test_region_insert_thresholds(monkeypatch):
    m = mx().StaticMap(1 << 22, -1, -1, torch.int64, torch.int64)
    cap = m.capacity()
    assert m.insert_uses_region_build(cap // 2 + 1)
    monkeypatch.setenv("MXCO_OA_REGION_INSERT", "0")
    assert not m.insert_uses_region_build(cap // 2 + 1)
    monkeypatch.setenv("MXCO_OA_REGION_INSERT", "1")
    assert not m.insert_uses_region_build(cap // 2 - 5)
    small = mx().StaticMap(1 << 20, -1, -1, torch.int64, torch.int64)
    assert not small.insert_uses_region_build((1 << 20) - 1)
    dh = mx().StaticMap(1 << 22, -1, -1, torch.int64, torch.int64, probing="double")
    assert not dh.insert_uses_region_build(1 << 22)
    monkeypatch.delenv("MXCO_OA_REGION_INSERT_MIN_BYTES")
    assert not m.insert_uses_region_build(cap // 2 + 1)  # 64 MiB of slots: below the default threshold


def test_region_insert_async_and_graph_replays():
    import mxco
    kd = vd = np.int64
    rng = np.random.default_rng(6)
    m = mxco.StaticMap(1 << 21, -1, -1, TD[kd], TD[vd])
    cap = m.capacity()
    k = distinct(rng, int(0.75 * cap), kd)
    pairs = dev_pairs(k, value_of(k, vd))
    q = np.concatenate([k[:100000], misses(rng, 100000, kd)])
    qd = dev(q)
    expect = np.concatenate([value_of(k[:100000], vd), np.full(100000, -1, vd)])
    assert m.insert_uses_region_build(len(k))
    m.insert_async(pairs)
    assert np.array_equal(m.find(qd).cpu().numpy(), expect)
    out = torch.empty(len(q), dtype=TD[vd], device="cuda")

    def record(s):
        m.clear_async(stream=s)
        assert m.insert_uses_region_build(len(k), stream=s)  # the workspace exists: taken under capture
        m.insert_async(pairs, stream=s)
        m.find_async(qd, out=out, stream=s)

    g = mxco.Graph(record)
    for rep in range(3):
        out.fill_(7)
        torch.cuda.synchronize()
        g.launch()
        g.stream.synchronize()
        assert np.array_equal(out.cpu().numpy(), expect), rep
        m.clear()
        torch.cuda.synchronize()


def test_captured_clear_does_not_mark_table_empty():
    """A Clear captured into a graph has not run when the capture returns: an
    eager region-built Insert that follows must still read the regions, so
    the keys already in the table stay."""
    import mxco
    kd = vd = np.int64
    rng = np.random.default_rng(11)
    m = mxco.StaticMap(1 << 21, -1, -1, TD[kd], TD[vd])
    cap = m.capacity()
    a = distinct(rng, int(0.3 * cap), kd)
    b = distinct(rng, int(0.55 * cap), kd, exclude=a)
    assert m.insert(dev_pairs(a, value_of(a, vd))) == 0
    g = mxco.Graph(lambda s: m.clear_async(stream=s))  # captured, not launched
    assert m.insert_uses_region_build(len(b))
    assert m.insert(dev_pairs(b, value_of(b, vd))) == 0
    assert m.contains(dev(a)).cpu().numpy().all()
    assert np.array_equal(m.find(dev(a)).cpu().numpy(), value_of(a, vd))
    assert np.array_equal(m.find(dev(b)).cpu().numpy(), value_of(b, vd))
    g.launch()
    g.stream.synchronize()
    assert not m.contains(dev(a)).cpu().numpy().any()
    assert not m.contains(dev(b)).cpu().numpy().any()


def test_captured_region_insert_keeps_its_workspace():
    """Once a region-built Insert is captured, its workspace is never
    reallocated: a later, larger eager Insert takes the CAS walk (same
    results), and replays of the graph still read valid buffers."""
    import mxco
    kd = vd = np.int64
    rng = np.random.default_rng(12)
    cap_req = 1 << 21
    m = mxco.StaticMap(cap_req, -1, -1, TD[kd], TD[vd])
    cap = m.capacity()
    k1 = distinct(rng, int(0.55 * cap), kd)
    k2 = distinct(rng, int(0.75 * cap), kd, exclude=k1)
    p1, p2 = dev_pairs(k1, value_of(k1, vd)), dev_pairs(k2, value_of(k2, vd))
    assert m.insert_uses_region_build(len(k1)) and m.insert_uses_region_build(len(k2))
    assert m.insert(p1) == 0  # workspace sized for len(k1)

    def record(s):
        m.clear_async(stream=s)
        assert m.insert_uses_region_build(len(k1), stream=s)
        m.insert_async(p1, stream=s)

    g = mxco.Graph(record)
    assert m.insert_uses_region_build(len(k1))
    assert not m.insert_uses_region_build(len(k2))  # would grow the captured workspace
    m.clear()
    ref = O.OpenAddressing(kd, vd, cap_req, -1, -1)
    assert m.insert(p2) == ref.insert(k2, value_of(k2, vd)) == 0
    q = np.concatenate([k1[:100000], k2[:100000], misses(rng, 50000, kd)])
    check_against(m, ref, kd, vd, q)
    for rep in range(2):
        g.launch()
        g.stream.synchronize()
        ref = O.OpenAddressing(kd, vd, cap_req, -1, -1)
        ref.insert(k1, value_of(k1, vd))
        check_against(m, ref, kd, vd, q)


def test_region_inserts_on_two_streams_are_ordered():
    """Two region-built InsertAsync calls on different streams share the
    table's workspace; the second waits for the first."""
    import mxco
    kd = vd = np.int64
    rng = np.random.default_rng(13)
    cap_req = 1 << 21
    m = mxco.StaticMap(cap_req, -1, -1, TD[kd], TD[vd])
    cap = m.capacity()
    u = distinct(rng, int(0.8 * cap), kd)
    a = u[: int(0.55 * cap)]
    b = np.concatenate([u[int(0.55 * cap):], a[: int(0.3 * cap)]])
    rng.shuffle(b)
    assert len(a) * 2 >= cap and len(b) * 2 >= cap
    pa, pb = dev_pairs(a, value_of(a, vd)), dev_pairs(b, value_of(b, vd))
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    for rep in range(2):
        m.clear()
        m.insert_async(pa, stream=s1)
        m.insert_async(pb, stream=s2)
        s1.synchronize()
        s2.synchronize()
        ref = O.OpenAddressing(kd, vd, cap_req, -1, -1)
        ref.insert(a, value_of(a, vd))
        ref.insert(b, value_of(b, vd))
        got = table_keys(m, kd)
        assert len(got) == len(u) and np.array_equal(np.sort(got), np.sort(u)), rep
        check_against(m, ref, kd, vd, np.concatenate([u[:200000], misses(rng, 50000, kd)]))
