"""Every measured configuration at its stated size, through the C ABI.

- StaticMap int64 on 2^28 requested slots (268,435,460 x 16 B) at load
  factors 0.5 / 0.7 / 0.8 / 0.9, uniform 64-bit keys: Insert by the default
  path (the region build), and once by the CAS walk; a Zipf(1.1) stream of
  2^27 draws over 2^24 keys; InsertOrAssign of 2^26 existing keys; Find of
  2^28 queries at 50 % hit.  Checked by closed form (hits return key*7+1,
  misses -1, no failure, Count), by the stored key set, and for the Zipf
  stream by the oracle's failure count.
- BloomFilter of 2^22 blocks: Add 2^27 keys, Contains 2^27 queries; words and
  outputs bit-exact against the oracle.
- HyperLogLog p = 15 fed 2^28 Zipf(1.0) keys over 1e8 values: registers
  bit-exact against the oracle; 64 sketches of disjoint seeds merged: the
  register-wise maximum, estimate equal to the oracle's finalizer.
- 1024 batched int32 tables of 2^10..2^22 slots: BUILD and INSERT plans, Find
  and Contains, against the closed form and per-table plain calls.
- Serving: 48 StaticMap + 16 StaticMultimap int64 tables of 2^20..2^26 slots,
  2000 requests with batches of 1..2^16 keys, served continuously and one by
  one: identical outputs, map outputs equal to the closed form."""
import math

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

MAP_2P28_REQ = 1 << 28


def mx():
    import mxco
    return mxco


def stored_keys(m):
    """Sorted keys held by an int64 map (every slot not holding -1)."""
    k = m.slots().view(torch.int64)[:, 0]
    k = k[k != -1]
    return torch.sort(k).values


def check_map(m, keys, q, hit):
    assert torch.equal(stored_keys(m), torch.sort(keys).values)
    out = m.find(q)
    assert torch.equal(out, torch.where(hit, q * 7 + 1, torch.full_like(q, -1)))
    assert torch.equal(m.contains(q).bool(), hit)
    assert m.count(q) == int(hit.sum().item())
    assert m.count(keys) == keys.numel()


@pytest.mark.parametrize("lf", [0.5, 0.7, 0.8, 0.9])
def test_map_2p28_slots(lf):
    from mxco.workloads import uniform_map
    mxco = mx()
    m = mxco.StaticMap(MAP_2P28_REQ, -1, -1, torch.int64, torch.int64)
    cap = m.capacity()
    assert cap == 268435460
    n = int(round(lf * cap))
    print("2^28-slot map, LF %.1f: %d keys into %d slots" % (lf, n, cap), flush=True)
    keys, q, hit = uniform_map(n, 1 << 28, 0.5, seed=int(lf * 10))
    pairs = mxco.make_pairs(keys, keys * 7 + 1)
    assert m.insert_uses_region_build(n)  # the default path at this size
    assert m.insert(pairs) == 0
    check_map(m, keys, q, hit)
    if lf == 0.8:
        # InsertOrAssign of 2^26 distinct existing keys, new value key*3+5
        k = keys[: 1 << 26]
        assert m.insert_or_assign(mxco.make_pairs(k, k * 3 + 5)) == (0, 1 << 26)
        assert torch.equal(m.find(k), k * 3 + 5)
        rest = keys[1 << 26:]
        assert torch.equal(m.find(rest), rest * 7 + 1)
        assert torch.equal(stored_keys(m), torch.sort(keys).values)


def test_map_2p28_cas_walk_lf08(monkeypatch):
    from mxco.workloads import uniform_map
    mxco = mx()
    monkeypatch.setenv("MXCO_OA_REGION_INSERT", "0")
    m = mxco.StaticMap(MAP_2P28_REQ, -1, -1, torch.int64, torch.int64)
    n = int(round(0.8 * m.capacity()))
    keys, q, hit = uniform_map(n, 1 << 28, 0.5, seed=81)
    assert not m.insert_uses_region_build(n)
    dup = torch.cat([keys, keys[: 1 << 20]])  # 2^20 duplicates: each fails once
    assert m.insert(mxco.make_pairs(dup, dup * 7 + 1)) == 1 << 20
    check_map(m, keys, q, hit)


def test_map_zipf_stream_2p27_draws():
    """2^27 draws of Zipf(s=1.1) over 2^24 keys into the 2^28-slot table:
    failures equal the oracle's sequential count (every repeat fails), the
    table holds each drawn key once with value key*7+1."""
    from mxco.workloads import Zipf
    mxco = mx()
    z = Zipf(1 << 24, 1.1)
    ranks = z.draw(1 << 27, seed=11)
    keys = Zipf.keys(ranks)
    del z, ranks
    m = mxco.StaticMap(MAP_2P28_REQ, -1, -1, torch.int64, torch.int64)
    failed = m.insert(mxco.make_pairs(keys, keys * 7 + 1))
    kh = keys.cpu().numpy()
    ref = O.OpenAddressing(np.int64, np.int64, 1 << 26, -1, -1)
    assert failed == ref.insert(kh, kh * 7 + 1)
    uk = torch.unique(keys)
    assert failed == keys.numel() - uk.numel()
    assert torch.equal(stored_keys(m), uk)
    assert torch.equal(m.find(uk), uk * 7 + 1)


def test_bloom_2p22_blocks_2p27_keys():
    nb, n = 1 << 22, 1 << 27
    g = torch.Generator(device="cuda")
    g.manual_seed(21)
    keys = torch.randint(-(1 << 63), (1 << 63) - 1, (n,), dtype=torch.int64, device="cuda", generator=g)
    absent = torch.randint(-(1 << 63), (1 << 63) - 1, (n - n // 2,), dtype=torch.int64, device="cuda", generator=g)
    q = torch.cat([keys[: n // 2], absent])
    bf = mx().BloomFilter(nb, torch.int64)
    bf.add(keys)
    out = bf.contains(q)
    words = np.zeros(nb * 8, np.uint32)
    O.bloom_add(words, nb, keys.cpu().numpy())
    assert np.array_equal(bf.words().cpu().numpy().view(np.uint32), words)
    ref = O.bloom_contains(words, nb, q.cpu().numpy())
    assert np.array_equal(out.cpu().numpy().astype(bool), ref)
    assert ref[: n // 2].all()
    fpr = ref[n // 2:].mean()
    print("bloom 2^22 blocks: FPR %.4f" % fpr)
    assert 0.02 < fpr < 0.05  # Poisson(32) keys per block, 8 bits each: about 3.3 %


def _hll_oracle_estimate(regs, p):
    d = np.load(O.os.path.join(O._HERE, "data", "hllpp_p%d.npz" % p))
    h = O.hll_histogram(regs, p)
    z = O.lib().orc_hll_z(O._ptr(h))
    return int(O.lib().orc_hll_finalize(p, z, int(h[0]), O._ptr(d["raw"]), O._ptr(d["bias"]), len(d["raw"]),
                                        float(d["threshold"])))


def test_hll_2p28_zipf_keys_and_64_sketch_merge():
    from mxco.workloads import Zipf
    mxco = mx()
    universe, n = 10 ** 8, 1 << 28
    z = Zipf(universe, 1.0)
    total = mxco.HyperLogLog.create_with_sketch_size_kb(32, torch.int64)
    p = total.precision
    assert p == 15
    h = mxco.HyperLogLog(p, torch.int64)
    seen = torch.zeros(universe, dtype=torch.bool, device="cuda")
    regmax = torch.zeros(1 << p, dtype=torch.uint8, device="cuda")
    for k in range(64):
        r = z.draw(n, seed=1000 + k)
        seen[r] = True
        keys = Zipf.keys(r)
        del r
        h.clear()
        h.add(keys)
        total.merge(h)
        regs = h.registers().view(torch.uint8)
        if k == 0:
            ref = np.zeros(1 << p, np.uint8)
            O.hll_add(ref, p, keys.cpu().numpy())
            assert np.array_equal(regs.cpu().numpy(), ref)
            assert h.estimate() == _hll_oracle_estimate(ref, p)
        regmax = torch.maximum(regmax, regs)
        del keys
    merged = total.registers().view(torch.uint8)
    assert torch.equal(merged, regmax)
    est = total.estimate()
    assert est == _hll_oracle_estimate(merged.cpu().numpy(), p)
    exact = int(seen.sum().item())
    assert abs(est - exact) / exact < 3 * 1.04 / math.sqrt(1 << p)


def test_batched_1024_tables_2p10_to_2p22_slots():
    mxco = mx()
    from mxco import workloads
    w = workloads.batched_tables(1024, (10, 22), (0.5, 0.85), (6, 20), seed=4,
                                 valid_extent=lambda c: mxco.valid_extent(c, 5, "linear"))
    T = len(w["caps"])
    assert T == 1024 and min(w["caps"]) >= 1 << 10 and max(w["caps"]) <= (1 << 22) + 5
    b = mxco.BatchedStaticMap(w["caps"], -1, -1)
    items = [(t, w["ins_off"][t], w["n_ins"][t]) for t in range(T)]
    qitems = [(t, w["find_off"][t], w["n_find"][t]) for t in range(T)]
    fnd = b.plan(mxco.BATCH_FIND, qitems)
    con = b.plan(mxco.BATCH_CONTAINS, qitems)
    out = torch.empty_like(w["queries"])
    flags = torch.empty(w["queries"].numel(), dtype=torch.uint8, device="cuda")
    for op in (mxco.BATCH_BUILD, mxco.BATCH_INSERT):
        b.clear()
        ins = b.plan(op, items)
        ins.run(w["pairs"])
        assert ins.failures() == [0] * T, op
        out.fill_(7)
        fnd.run(w["queries"], out)
        con.run(w["queries"], flags)
        assert torch.equal(out, w["expected"]), op
        assert torch.equal(flags.bool(), w["expected"] != -1), op
    # plain per-table calls on every 32nd table give the same outputs
    for t in range(0, T, 32):
        a, n = w["ins_off"][t], w["n_ins"][t]
        qa, qn = w["find_off"][t], w["n_find"][t]
        m = mxco.StaticMap(w["caps"][t], -1, -1)
        assert m.insert(w["pairs"][a:a + n]) == 0
        assert torch.equal(m.find(w["queries"][qa:qa + qn]), out[qa:qa + qn])


OPS = {"insert": 0, "find": 1, "contains": 2, "retrieve": 6}


def test_serving_64_tables_2p20_to_2p26_slots():
    mxco = mx()
    from mxco import workloads
    w = workloads.serving(num_maps=48, num_multimaps=16, slots_log2=(20, 26), num_requests=2000, qps=200000.0,
                          batch_log2=(0, 16), seed=5)
    assert min(w["caps"]) >= 1 << 20 and max(w["caps"]) <= 1 << 26

    def build():
        torch.cuda.synchronize()  # the default stream's inputs are written before submit
        s = mxco.Server()
        tables = [s.add_map(c) if t < w["num_maps"] else s.add_multimap(c) for t, c in enumerate(w["caps"])]
        for t, pairs in enumerate(w["init"]):
            op = mxco.REQ_INSERT if t < w["num_maps"] else mxco.REQ_MM_INSERT
            assert s.wait(s.submit(tables[t], op, pairs))[0] == 0
        return s, tables

    def outputs():
        o = []
        for r in w["requests"]:
            if r["op"] == "find":
                o.append((torch.full((r["n"],), 7, dtype=torch.int64, device="cuda"), None))
            elif r["op"] == "contains":
                o.append((torch.full((r["n"],), 7, dtype=torch.uint8, device="cuda"), None))
            elif r["op"] == "retrieve":
                c = 8 * r["n"] + 8
                o.append((torch.zeros(c, dtype=torch.int64, device="cuda"),
                          torch.zeros(c, dtype=torch.int64, device="cuda")))
            else:
                o.append((None, None))
        return o

    def plan(tables, outs):
        return [(tables[r["table"]], OPS[r["op"]], r["input"].data_ptr(), r["n"],
                 o1.data_ptr() if o1 is not None else None, o2.data_ptr() if o2 is not None else None,
                 o1.numel() if r["op"] == "retrieve" else 0) for r, (o1, o2) in zip(w["requests"], outs)]

    s1, t1 = build()
    o1 = outputs()
    torch.cuda.synchronize()
    r1, lat = s1.replay(plan(t1, o1), w["arrivals_us"])
    s1.close()
    s2, t2 = build()
    o2 = outputs()
    torch.cuda.synchronize()
    r2, _ = s2.replay(plan(t2, o2), w["arrivals_us"], plain=True)
    s2.close()
    assert r1 == r2
    for j, (req, a, b) in enumerate(zip(w["requests"], o1, o2)):
        if a[0] is not None:
            assert torch.equal(a[0], b[0]), (j, req["op"])
        if a[1] is not None:
            assert torch.equal(a[1], b[1]), (j, req["op"])
        q, hit = req["input"], req["hit"]
        if req["op"] == "find":
            assert torch.equal(a[0], torch.where(hit, q * 7 + 1, torch.full_like(q, -1))), j
        elif req["op"] == "contains":
            assert torch.equal(a[0].bool(), hit), j
        elif req["op"] == "retrieve":
            cnt = r1[j]
            probe, vals = a[0][:cnt], a[1][:cnt]
            assert torch.equal(vals >> 4, probe), j  # value = key * 16 + copy
            if cnt:
                assert torch.equal(torch.unique_consecutive(probe), torch.unique_consecutive(q[hit])), j
    assert sorted(lat)[int(0.99 * len(lat))] >= 0
