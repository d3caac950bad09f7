"""Batched BUILD (Clear of every table, then Insert, built region by region in
LDS): per-item failures and Find / Contains outputs equal Clear + an insert
plan's and the oracle's; each table holds every distinct key exactly once and
nothing else; tables without an item are left empty; graph replays repeat it;
plans the region build does not take (double hashing, a table named twice)
give the same results through Clear + the insert kernel."""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

LAYOUTS = [(np.int32, np.int32), (np.int64, np.int64), (np.int32, np.int64), (np.int64, np.int32),
           (np.int32, None), (np.int64, None)]
TD = {np.int32: torch.int32, np.int64: torch.int64}


def slot_size(kd, vd):
    ks = np.dtype(kd).itemsize
    if vd is None:
        return ks
    return 8 if ks == 4 and np.dtype(vd).itemsize == 4 else 16


def region_slots(kd, vd):
    """Slots of one region: 64 KiB of slots, at most 4096."""
    return min(65536 // slot_size(kd, vd), 4096)


def value_of(keys, vd):
    """Payload of a key (so duplicates of a key carry equal payloads)."""
    v = (keys.astype(np.int64) * 7 + 1) & 0x3fffffff
    return v.astype(vd)


def workload(kd, vd, caps, lfs, seed, dup=0.0, reserved=0):
    """Per table: round(lf * cap) distinct keys, plus `dup` of them again and
    `reserved` copies of the empty key -1, shuffled."""
    rng = np.random.default_rng(seed)
    hi = 2**30 if kd == np.int32 else 2**62
    keys, offs, counts, distinct = [], [], [], []
    off = 0
    for cap, lf in zip(caps, lfs):
        n = int(round(lf * cap))
        k = np.unique(rng.integers(0, hi, int(n * 1.1) + 8))[:n]
        rng.shuffle(k)
        extra = [k[rng.integers(0, max(1, len(k)), int(dup * len(k)))] if len(k) else k[:0],
                 np.full(reserved, -1, dtype=np.int64)]
        allk = np.concatenate([k.astype(np.int64)] + [e.astype(np.int64) for e in extra])
        rng.shuffle(allk)
        keys.append(allk.astype(kd))
        distinct.append(k.astype(kd))
        offs.append(off)
        counts.append(len(allk))
        off += len(allk)
    allkeys = np.concatenate(keys)
    vals = value_of(allkeys, vd) if vd is not None else None
    pairs = torch.from_numpy(O.pack_pairs(allkeys, vals)).cuda()
    return {"keys": keys, "distinct": distinct, "offs": offs, "counts": counts, "pairs": pairs}


def queries_for(w, kd, seed):
    """Every distinct key of each table plus as many absent keys."""
    rng = np.random.default_rng(seed + 1)
    hi = 2**30 if kd == np.int32 else 2**62
    qs, offs, counts = [], [], []
    off = 0
    for d in w["distinct"]:
        miss = rng.integers(hi, 2 * hi - 1, len(d) + 3).astype(kd)  # above every inserted key, never -1
        q = np.concatenate([d, miss])
        qs.append(q)
        offs.append(off)
        counts.append(len(q))
        off += len(q)
    return np.concatenate(qs), offs, counts


def run_both(kd, vd, caps, w, q, qoffs, qcounts, probing="linear", items=None):
    """Build plan on one batch, Clear + insert plan on another; returns both
    batches, failures and Find / Contains outputs."""
    import mxco
    T = len(caps)
    vt = TD[vd] if vd is not None else None
    res = []
    items = items or [(t, w["offs"][t], w["counts"][t]) for t in range(T)]
    for op in (mxco.BATCH_BUILD, mxco.BATCH_INSERT):
        b = mxco.BatchedStaticMap(caps, -1, -1 if vd is not None else None, key_dtype=TD[kd], value_dtype=vt,
                                  probing=probing)
        ins = b.plan(op, items)
        if op == mxco.BATCH_INSERT:
            b.clear()
        ins.run(w["pairs"])
        fails = ins.failures()
        qd = torch.from_numpy(q).cuda()
        out = torch.empty(len(q), dtype=vt if vt is not None else TD[kd], device="cuda")
        flags = torch.empty(len(q), dtype=torch.uint8, device="cuda")
        fnd = b.plan(mxco.BATCH_FIND, [(t, qoffs[t], qcounts[t]) for t in range(T)])
        con = b.plan(mxco.BATCH_CONTAINS, [(t, qoffs[t], qcounts[t]) for t in range(T)])
        fnd.run(qd, out)
        con.run(qd, flags)
        torch.cuda.synchronize()
        res.append((b, fails, out.cpu().numpy(), flags.cpu().numpy()))
    return res


def table_keys(b, t, kd):
    s = b.slots(t).cpu().numpy()
    ks = np.dtype(kd).itemsize
    k = np.ascontiguousarray(s[:, :ks]).view(kd).reshape(-1)
    return k[k != kd(-1)]


@pytest.mark.parametrize("kd,vd", LAYOUTS)
def test_build_equals_clear_insert(kd, vd):
    R = region_slots(kd, vd)
    caps_req = [5, 100, 3000, R - 1, R, R + 1, 3 * R + 7, 9 * R]
    import mxco
    caps = [mxco.valid_extent(c, 5, "linear") for c in caps_req]
    lfs = [0.6, 0.9, 0.95, 0.85, 0.97, 0.9, 0.95, 0.92]
    w = workload(kd, vd, caps, lfs, seed=11, dup=0.2, reserved=3)
    q, qo, qc = queries_for(w, kd, 11)
    (bb, fb, ob, cb), (bi, fi, oi, ci) = run_both(kd, vd, caps, w, q, qo, qc)
    expect_fail = [w["counts"][t] - len(w["distinct"][t]) for t in range(len(caps))]
    assert fi == expect_fail
    assert fb == expect_fail
    assert np.array_equal(ob, oi)
    assert np.array_equal(cb, ci)
    for t in range(len(caps)):
        # every distinct key exactly once, nothing else
        got = np.sort(table_keys(bb, t, kd))
        assert np.array_equal(got, np.sort(w["distinct"][t])), t
        qs = q[qo[t]:qo[t] + qc[t]]
        hit = cb[qo[t]:qo[t] + qc[t]].astype(bool)
        assert hit[:len(w["distinct"][t])].all() and not hit[len(w["distinct"][t]):].any(), t
        if vd is not None:
            o = ob[qo[t]:qo[t] + qc[t]]
            assert np.array_equal(o[hit], value_of(qs[hit], vd)), t


def test_build_matches_oracle():
    """int32 map: per-table results of the build equal the oracle's sequential inserts."""
    import mxco
    kd, vd = np.int32, np.int32
    R = region_slots(kd, vd)
    caps = [mxco.valid_extent(c, 5, "linear") for c in (77, R + 3, 4 * R + 1)]
    w = workload(kd, vd, caps, [0.9, 0.93, 0.96], seed=5, dup=0.3, reserved=2)
    q, qo, qc = queries_for(w, kd, 5)
    (bb, fb, ob, cb), _ = run_both(kd, vd, caps, w, q, qo, qc)
    for t, cap in enumerate(caps):
        ref = O.OpenAddressing(np.int32, np.int32, cap, -1, -1)
        failed = ref.insert(w["keys"][t], value_of(w["keys"][t], vd))
        assert failed == fb[t]
        rv = ref.find(q[qo[t]:qo[t] + qc[t]])[0]
        assert rv.tobytes() == ob[qo[t]:qo[t] + qc[t]].tobytes()


@pytest.mark.parametrize("kd,vd", [(np.int32, np.int32), (np.int64, np.int64), (np.int32, None)])
def test_build_full_tables(kd, vd):
    """More distinct keys than slots: one-region and many-region tables fill
    up, the rest fail, and the two paths agree on the counts."""
    import mxco
    R = region_slots(kd, vd)
    caps = [mxco.valid_extent(c, 5, "linear") for c in (50, R - 3, 3 * R + 2)]
    w = workload(kd, vd, caps, [1.3, 1.1, 1.05], seed=3)
    q, qo, qc = queries_for(w, kd, 3)
    (bb, fb, ob, cb), (bi, fi, oi, ci) = run_both(kd, vd, caps, w, q, qo, qc)
    for t, cap in enumerate(caps):
        n = len(w["distinct"][t])
        assert fb[t] == fi[t] == n - cap
        got = table_keys(bb, t, kd)
        assert len(got) == cap and len(np.unique(got)) == cap
        assert np.isin(got, w["distinct"][t]).all()
        # the keys Find reports present are exactly the stored ones
        qs = q[qo[t]:qo[t] + qc[t]]
        hit = cb[qo[t]:qo[t] + qc[t]].astype(bool)
        assert np.array_equal(np.sort(qs[hit]), np.sort(got))


def test_build_clears_tables_without_items():
    import mxco
    kd, vd = np.int32, np.int32
    R = region_slots(kd, vd)
    caps = [mxco.valid_extent(c, 5, "linear") for c in (300, 2 * R + 9, 1000, R + 1)]
    w = workload(kd, vd, caps, [0.8] * 4, seed=9)
    b = mxco.BatchedStaticMap(caps, -1, -1)
    ins = b.plan(mxco.BATCH_INSERT, [(t, w["offs"][t], w["counts"][t]) for t in range(4)])
    ins.run(w["pairs"])
    assert ins.failures() == [0] * 4
    # build with items for tables 1 and 3 only (in reverse order)
    bld = b.plan(mxco.BATCH_BUILD, [(3, w["offs"][3], w["counts"][3]), (1, w["offs"][1], w["counts"][1])])
    bld.run(w["pairs"])
    assert bld.failures() == [0, 0]
    for t in (0, 2):
        assert len(table_keys(b, t, kd)) == 0
    for t in (1, 3):
        assert np.array_equal(np.sort(table_keys(b, t, kd)), np.sort(w["distinct"][t]))


def test_build_graph_replays_match_eager():
    import mxco
    from mxco import workloads
    w = workloads.batched_tables(64, (8, 16), (0.5, 0.9), (2, 12), seed=21,
                                 valid_extent=lambda c: mxco.valid_extent(c, 5, "linear"))
    T = len(w["caps"])
    b = mxco.BatchedStaticMap(w["caps"], -1, -1)
    bld = b.plan(mxco.BATCH_BUILD, [(t, w["ins_off"][t], w["n_ins"][t]) for t in range(T)])
    fnd = b.plan(mxco.BATCH_FIND, [(t, w["find_off"][t], w["n_find"][t]) for t in range(T)])
    out = torch.empty_like(w["queries"])
    bld.run(w["pairs"])
    fnd.run(w["queries"], out)
    assert torch.equal(out, w["expected"])
    assert bld.failures() == [0] * T

    def record(s):
        bld.run(w["pairs"], stream=s)
        fnd.run(w["queries"], out, stream=s)

    g = mxco.Graph(record)
    for rep in range(4):
        out.fill_(123)
        torch.cuda.synchronize()
        g.launch()
        g.stream.synchronize()
        assert torch.equal(out, w["expected"]), rep
        assert bld.failures(g.stream) == [0] * T
        # eager work between replays
        b.clear()
        torch.cuda.synchronize()


@pytest.mark.parametrize("case", ["double_hashing", "table_twice"])
def test_build_fallback_paths(case):
    import mxco
    kd, vd = np.int32, np.int32
    R = region_slots(kd, vd)
    probing = "double" if case == "double_hashing" else "linear"
    caps = [mxco.valid_extent(c, 5, probing) for c in (200, 2 * R + 1)]
    w = workload(kd, vd, caps, [0.8, 0.85], seed=4, dup=0.1)
    q, qo, qc = queries_for(w, kd, 4)
    items = None
    if case == "table_twice":
        n1 = w["counts"][1]
        items = [(0, w["offs"][0], w["counts"][0]), (1, w["offs"][1], n1 // 2),
                 (1, w["offs"][1] + n1 // 2, n1 - n1 // 2)]
    (bb, fb, ob, cb), (bi, fi, oi, ci) = run_both(kd, vd, caps, w, q, qo, qc, probing=probing, items=items)
    assert sum(fb) == sum(fi) == sum(w["counts"]) - sum(len(d) for d in w["distinct"])
    assert np.array_equal(ob, oi)
    assert np.array_equal(cb, ci)
