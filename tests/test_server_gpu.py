"""Serving engine: a continuously batched open-loop request stream over
resident tables gives exactly the results of the same requests served one
at a time with the plain synchronous calls; admission refuses tables beyond
the HBM budget."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def mx():
    import mxco
    return mxco


OPS = {"insert": 0, "find": 1, "contains": 2, "retrieve": 6}


def ready():
    """Requests carry device pointers: their inputs and outputs must be
    written before submit.  The tensors here come from the default stream,
    which does not order itself with the server's streams, so wait for it."""
    torch.cuda.synchronize()


def build(w, **cfg):
    import mxco
    ready()
    s = mxco.Server(**cfg)
    tables = []
    for t, cap in enumerate(w["caps"]):
        tables.append(s.add_map(cap) if t < w["num_maps"] else s.add_multimap(cap))
    for t, pairs in enumerate(w["init"]):
        op = mxco.REQ_INSERT if t < w["num_maps"] else mxco.REQ_MM_INSERT
        res, _ = s.wait(s.submit(tables[t], op, pairs))
        assert res == 0
    return s, tables


def outputs(w):
    outs = []
    for r in w["requests"]:
        if r["op"] == "find":
            outs.append((torch.full((r["n"],), 7, dtype=torch.int64, device="cuda"), None))
        elif r["op"] == "contains":
            outs.append((torch.full((r["n"],), 7, dtype=torch.uint8, device="cuda"), None))
        elif r["op"] == "retrieve":
            cap = 8 * r["n"] + 8
            outs.append((torch.zeros(cap, dtype=torch.int64, device="cuda"),
                         torch.zeros(cap, dtype=torch.int64, device="cuda")))
        else:
            outs.append((None, None))
    return outs


def plan(w, tables, outs):
    reqs = []
    for r, (o1, o2) in zip(w["requests"], outs):
        cap = o1.numel() if r["op"] == "retrieve" else 0
        reqs.append((tables[r["table"]], OPS[r["op"]] + (0 if r["table"] < w["num_maps"] else 0),
                     r["input"].data_ptr(), r["n"], o1.data_ptr() if o1 is not None else None,
                     o2.data_ptr() if o2 is not None else None, cap))
    return reqs


@pytest.mark.parametrize("resident", [True, False], ids=["resident", "launched"])
@pytest.mark.parametrize("qps", [2000.0, 200000.0])
def test_batched_serving_equals_one_by_one(qps, resident):
    from mxco import workloads
    w = workloads.serving(num_maps=6, num_multimaps=2, slots_log2=(12, 16), num_requests=300, qps=qps,
                          batch_log2=(0, 10), seed=11)
    s1, t1 = build(w, max_tick_keys=1 << 14, resident=resident)
    assert s1.resident_stats()[0] == resident
    s2, t2 = build(w)
    o1, o2 = outputs(w), outputs(w)
    ready()
    r1, lat1 = s1.replay(plan(w, t1, o1), w["arrivals_us"])
    r2, lat2 = s2.replay(plan(w, t2, o2), w["arrivals_us"], plain=True)
    assert r1 == r2
    for req, a, b in zip(w["requests"], o1, o2):
        if a[0] is not None:
            assert torch.equal(a[0], b[0]), req["op"]
        if a[1] is not None:
            assert torch.equal(a[1], b[1])
    ticks, batched = s1.stats()
    assert ticks >= 1 and batched >= 1
    if qps > 100000:
        assert batched > ticks  # requests were folded together
    assert all(x >= 0 for x in lat1)
    s1.close()
    s2.close()


def test_find_results_against_closed_form():
    """Map finds return key*7+1 for present keys and -1 for absent ones, the
    workload's closed form (a query's hit flag is fixed when it is drawn)."""
    from mxco import workloads
    w = workloads.serving(num_maps=3, num_multimaps=1, slots_log2=(12, 13), num_requests=60, qps=100000.0,
                          batch_log2=(0, 9), seed=3)
    w["requests"] = [r for r in w["requests"] if r["op"] in ("find", "contains")]
    w["arrivals_us"] = w["arrivals_us"][:len(w["requests"])]
    s, t = build(w)
    o = outputs(w)
    ready()
    s.replay(plan(w, t, o), w["arrivals_us"])
    for req, (a, _) in zip(w["requests"], o):
        hit = req["hit"]
        if req["op"] == "find":
            assert torch.equal(a, torch.where(hit, req["input"] * 7 + 1, torch.full_like(a, -1)))
        else:
            assert torch.equal(a.bool(), hit)
    s.close()


def test_same_table_inserts_keep_arrival_order():
    """Insert requests on one table that arrive together, with keys shared
    between requests and different values: the earliest request's value
    stays (as one by one), and each request's result equals the plain
    path's; a find of every key reads the first writer's value."""
    import mxco
    rng = np.random.default_rng(17)
    space, nreq, per = 1500, 24, 300
    reqs, inputs, expect = [], [], {}
    for r in range(nreq):
        k = rng.choice(space, per, replace=False).astype(np.int64)  # distinct within a request
        v = k * 1000 + r
        for kk, vv in zip(k.tolist(), v.tolist()):
            expect.setdefault(kk, vv)
        inputs.append(torch.from_numpy(np.stack([k, v], 1)).cuda().contiguous())
    q = torch.arange(space, dtype=torch.int64, device="cuda")
    results = []
    for plain in (False, True):
        s = mxco.Server()
        t = s.add_map(4096)
        outs = torch.full((space,), 7, dtype=torch.int64, device="cuda")
        plan_ = [(t, mxco.REQ_INSERT, x.data_ptr(), per, None, None, 0) for x in inputs]
        plan_.append((t, mxco.REQ_FIND, q.data_ptr(), space, outs.data_ptr(), None, 0))
        ready()
        res, _ = s.replay(plan_, [0.0] * len(plan_), plain=plain)  # all at once
        results.append((res, outs.cpu().numpy()))
        s.close()
    assert results[0][0] == results[1][0]
    exp = np.array([expect.get(i, -1) for i in range(space)], np.int64)
    np.testing.assert_array_equal(results[0][1], exp)
    np.testing.assert_array_equal(results[1][1], exp)


def _mm_pairs(rng, table_salt, nkeys, mult):
    """Pairs over keys 0..nkeys-1 (+ salt << 32), key k repeated mult(k)
    times with distinct values, shuffled."""
    ks, vs = [], []
    for k in range(nkeys):
        for c in range(mult(k)):
            ks.append(k + (table_salt << 32))
            vs.append(int(rng.integers(-10**9, 10**9)))
    perm = rng.permutation(len(ks))
    k = np.array(ks, np.int64)[perm]
    v = np.array(vs, np.int64)[perm]
    return torch.from_numpy(np.stack([k, v], 1)).cuda().contiguous(), k, v


def test_multimap_ticks_equal_plain_and_counts():
    """Multimap requests arriving together (inserts, Find, Count, Retrieve
    with segments of up to 300 values, on three tables): the batched ticks
    give the plain one-by-one results; counts equal the host's multiplicities
    at each point of the arrival order."""
    import mxco
    rng = np.random.default_rng(29)
    mult_a = lambda k: 1 + (k % 5) + (300 if k % 97 == 0 else 0)
    mult_b = lambda k: 1 + (k % 3)
    batches = []
    for t in range(3):
        a = _mm_pairs(rng, t, 400, mult_a)
        b = _mm_pairs(rng, t, 500, mult_b)  # overlaps a's keys 0..399 and adds 400..499
        batches.append((a, b))
    q = [torch.from_numpy(np.arange(-3, 520, dtype=np.int64) + (t << 32)).cuda() for t in range(3)]
    nq = q[0].numel()
    results = []
    for plain in (False, True):
        s = mxco.Server()
        tabs = [s.add_multimap(1 << 14) for _ in range(3)]
        outs, plan_ = [], []
        for phase in range(2):
            for t in range(3):
                pairs = batches[t][phase][0]
                plan_.append((tabs[t], mxco.REQ_MM_INSERT, pairs.data_ptr(), pairs.shape[0], None, None, 0))
            for t in range(3):
                f = torch.full((nq,), 7, dtype=torch.int64, device="cuda")
                cap = 9 * nq + 400 * 3
                po = torch.zeros(cap, dtype=torch.int64, device="cuda")
                vo = torch.zeros(cap, dtype=torch.int64, device="cuda")
                outs.append((f, po, vo))
                plan_.append((tabs[t], mxco.REQ_MM_FIND, q[t].data_ptr(), nq, f.data_ptr(), None, 0))
                plan_.append((tabs[t], mxco.REQ_MM_COUNT, q[t].data_ptr(), nq, None, None, 0))
                plan_.append((tabs[t], mxco.REQ_MM_RETRIEVE, q[t].data_ptr(), nq, po.data_ptr(), vo.data_ptr(), cap))
        ready()
        res, _ = s.replay(plan_, [0.0] * len(plan_), plain=plain)
        results.append((res, [tuple(x.cpu().numpy() for x in o) for o in outs]))
        s.close()
    assert results[0][0] == results[1][0]
    for a, b in zip(results[0][1], results[1][1]):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    # counts: phase 0 after batch a, phase 1 after a + b
    res = results[0][0]
    per_phase = 3 + 3 * 3
    for phase in range(2):
        for t in range(3):
            count = res[per_phase * phase + 3 + 3 * t + 1]
            expect = sum(mult_a(k) for k in range(400))
            if phase == 1:
                expect += sum(mult_b(k) for k in range(500))
            assert count == expect, (phase, t)


def test_multimap_retrieve_too_small_fails_alone():
    """A served Retrieve whose matches exceed its capacity raises
    std::length_error's counterpart and writes nothing; the other requests
    of its tick complete."""
    import mxco
    s = mxco.Server()
    t = s.add_multimap(1024)
    pairs = torch.tensor([[5, 1], [5, 2], [5, 3], [6, 4]], dtype=torch.int64, device="cuda")
    ready()
    s.wait(s.submit(t, mxco.REQ_MM_INSERT, pairs))
    k5 = torch.tensor([5], dtype=torch.int64, device="cuda")
    k6 = torch.tensor([6], dtype=torch.int64, device="cuda")
    po = torch.full((2,), 9, dtype=torch.int64, device="cuda")
    vo = torch.full((2,), 9, dtype=torch.int64, device="cuda")
    po6 = torch.zeros(2, dtype=torch.int64, device="cuda")
    vo6 = torch.zeros(2, dtype=torch.int64, device="cuda")
    ready()
    bad = s.submit(t, mxco.REQ_MM_RETRIEVE, k5, out=po, out2=vo, out_capacity=2)
    good = s.submit(t, mxco.REQ_MM_RETRIEVE, k6, out=po6, out2=vo6, out_capacity=2)
    with pytest.raises(mxco.LengthError):
        s.wait(bad)
    assert s.wait(good)[0] == 1
    assert po.cpu().tolist() == [9, 9] and vo.cpu().tolist() == [9, 9]
    assert po6.cpu().tolist()[0] == 6 and vo6.cpu().tolist()[0] == 4
    s.close()


def test_admission_by_budget():
    import mxco
    s = mxco.Server(hbm_budget_bytes=64 << 20)
    s.add_map(1 << 20)  # 16 MiB of slots
    with pytest.raises(MemoryError):
        s.add_map(1 << 23)  # 128 MiB: over the budget
    t = s.add_map(1 << 20)
    s.drop_table(t)
    s.close()


def test_drop_table_waits_for_its_requests_and_refuses_new_ones():
    """drop_table lets every request already submitted to the table finish,
    then frees it; later submits to it, and a second drop, fail."""
    import mxco
    s = mxco.Server()
    t = s.add_map(1 << 20)
    keys = torch.arange(1, 1 << 18, dtype=torch.int64, device="cuda")
    pairs = mxco.make_pairs(keys, keys * 7 + 1)
    out = torch.full((keys.numel(),), 5, dtype=torch.int64, device="cuda")
    ready()
    tickets = [s.submit(t, mxco.REQ_INSERT, pairs), s.submit(t, mxco.REQ_FIND, keys, out=out)]
    s.drop_table(t)
    assert s.wait(tickets[0])[0] == 0
    s.wait(tickets[1])
    torch.cuda.synchronize()
    assert torch.equal(out, keys * 7 + 1)
    with pytest.raises(Exception, match="unknown table"):
        s.submit(t, mxco.REQ_FIND, keys, out=out)
    with pytest.raises(Exception, match="unknown table"):
        s.drop_table(t)
    s.close()


def test_resident_kernel_serves_many_ticks_per_launch():
    """Closed loop: every StaticMap tick goes through the resident kernel's
    mailbox, one launch serves many ticks, and the results equal the
    launched ticks' and the workload's closed form (a read's hit flag holds
    whatever the inserts: present keys come from the initial fill, absent
    ones from indices no insert reaches)."""
    from mxco import workloads
    w = workloads.serving(num_maps=4, num_multimaps=1, slots_log2=(14, 16), num_requests=400, qps=1e6,
                          batch_log2=(0, 12), seed=23)
    runs = []
    for resident in (True, False):
        s, t = build(w, resident=resident)
        o = outputs(w)
        ready()
        res, _ = s.replay(plan(w, t, o), [0.0] * len(w["requests"]))
        ticks, _ = s.stats()
        on, launches, grid = s.resident_stats()
        assert on == resident
        if resident:
            assert 1 <= launches < ticks, (launches, ticks)
            assert grid > 1
        runs.append((res, [(a.cpu() if a is not None else None, b.cpu() if b is not None else None) for a, b in o]))
        s.close()
    assert runs[0][0] == runs[1][0]
    for (a, a2), (b, b2) in zip(runs[0][1], runs[1][1]):
        if a is not None:
            assert torch.equal(a, b)
        if a2 is not None:
            assert torch.equal(a2, b2)
    for req, (a, _) in zip(w["requests"], runs[0][1]):
        if req["op"] not in ("find", "contains"):
            continue
        hit, inp = req["hit"].cpu(), req["input"].cpu()
        if req["op"] == "find":
            assert torch.equal(a, torch.where(hit, inp * 7 + 1, torch.full_like(a, -1)))
        elif req["op"] == "contains":
            assert torch.equal(a.bool(), hit)


def test_resident_kernel_relaunches_after_idle_exit():
    """The resident kernel ends after 2 ms without a tick; the next request
    launches it again and is answered like the first."""
    import time
    import mxco
    s = mxco.Server()
    t = s.add_map(1 << 12)
    keys = torch.arange(1, 1001, dtype=torch.int64, device="cuda")
    pairs = torch.stack([keys, keys * 7 + 1], 1).contiguous()
    ready()
    assert s.wait(s.submit(t, mxco.REQ_INSERT, pairs))[0] == 0
    launches0 = s.resident_stats()[1]
    assert launches0 >= 1
    q = torch.arange(500, 1501, dtype=torch.int64, device="cuda")
    exp = torch.where(q <= 1000, q * 7 + 1, torch.full_like(q, -1))
    for rep in range(3):
        time.sleep(0.05)  # well past the idle limit
        out = torch.full_like(q, 3)
        ready()
        s.wait(s.submit(t, mxco.REQ_FIND, q, out=out))
        assert torch.equal(out, exp), rep
    assert s.resident_stats()[1] >= launches0 + 3
    s.close()


def test_resident_kernel_deadline_fails_the_tick():
    """A tick open past the resident kernel's deadline is given up: the
    kernel exits normally, the request fails with an error instead of
    waiting, and the server still closes.  The deadline is set to zero from
    the host, so the first tick trips it."""
    import os
    import mxco
    os.environ["MXCO_SERVE_HARD_US"] = "0"
    try:
        s = mxco.Server()
    finally:
        del os.environ["MXCO_SERVE_HARD_US"]
    # 2^22 keys make 2048 chunks: far more than the workers claim before the
    # deadline's exit flag reaches them (each runs at most the chunk it holds
    # and the one claimed ahead)
    t = s.add_map(1 << 23)
    keys = torch.arange(1, 1 << 22, dtype=torch.int64, device="cuda")
    pairs = torch.stack([keys, keys * 7 + 1], 1).contiguous()
    ready()
    failed = 0
    for _ in range(3):
        try:
            s.wait(s.submit(t, mxco.REQ_INSERT, pairs))
        except RuntimeError as e:
            assert "deadline" in str(e), e
            failed += 1
    assert failed >= 1
    s.close()


@pytest.mark.parametrize("dt,vt,group", [(torch.int64, torch.int64, 1), (torch.int64, torch.int64, 2),
                                         (torch.int64, torch.int64, 4), (torch.int64, torch.int64, 8),
                                         (torch.int32, torch.int32, 1), (torch.int32, torch.int64, 1),
                                         (torch.int64, torch.int32, 1)],
                         ids=["i64-g1", "i64-g2", "i64-g4", "i64-g8", "i32-g1", "i32i64-g1", "i64i32-g1"])
def test_resident_large_reads_match_closed_form(dt, vt, group):
    """Ticks with millions of read keys: their read chunks (2048 keys) run the
    wave-staged loop in the resident kernel, ragged last chunks of 1971 and
    1000 keys run the staged and the per-lane loop.  One tick holds Find and
    Contains of present and absent keys; the next holds an insert of new
    keys and a Find of them, which waits for that insert.  Finds return
    key*7+1 or -1, Contains the hit flag, as the launched ticks do."""
    import mxco
    g = torch.Generator(device="cuda").manual_seed(11)
    nk = 1 << 21
    perm = torch.randperm(1 << 24, device="cuda", generator=g)
    keys = (perm[:nk] * 2 + 1).to(dt)          # odd: present after the first insert
    new = (perm[nk:nk + (1 << 20)] * 2 + 1).to(dt)  # odd, inserted by the second tick
    absent = (perm[:nk] * 2 + 2).to(dt)         # even: never inserted
    nq = 1 << 20
    hit = torch.rand(nq + 1000, device="cuda", generator=g) < 0.5
    q = torch.where(hit, keys[:nq + 1000], absent[:nq + 1000])
    q_c = q[:nq - 77]
    runs = []
    for resident in (True, False):
        ready()
        s = mxco.Server(key_dtype=dt, value_dtype=vt, group_size=group, resident=resident)
        t = s.add_map(1 << 23)
        res, _ = s.wait(s.submit(t, mxco.REQ_INSERT, mxco.make_pairs(keys, (keys * 7 + 1).to(vt))))
        assert res == 0
        f_out = torch.full((q.numel(),), 5, dtype=vt, device="cuda")
        c_out = torch.full((q_c.numel(),), 5, dtype=torch.uint8, device="cuda")
        n_out = torch.full((new.numel(),), 5, dtype=vt, device="cuda")
        new_pairs = mxco.make_pairs(new, (new * 7 + 1).to(vt))
        ready()
        reqs = [(t, mxco.REQ_FIND, q.data_ptr(), q.numel(), f_out.data_ptr(), None, 0),
                (t, mxco.REQ_CONTAINS, q_c.data_ptr(), q_c.numel(), c_out.data_ptr(), None, 0)]
        res, _ = s.replay(reqs, [0.0, 0.0])
        reqs = [(t, mxco.REQ_INSERT, new_pairs.data_ptr(), new.numel(), None, None, 0),
                (t, mxco.REQ_FIND, new.data_ptr(), new.numel(), n_out.data_ptr(), None, 0)]
        res2, _ = s.replay(reqs, [0.0, 0.0])
        assert res2[0] == 0
        assert s.resident_stats()[0] == resident
        s.close()
        torch.cuda.synchronize()
        assert torch.equal(f_out, torch.where(hit, q * 7 + 1, torch.full_like(q, -1)).to(vt))
        assert torch.equal(c_out.bool(), hit[:nq - 77])
        assert torch.equal(n_out, (new * 7 + 1).to(vt))
        runs.append((f_out.cpu(), c_out.cpu(), n_out.cpu()))
    for a, b in zip(*runs):
        assert torch.equal(a, b)



def test_resident_tick_of_many_requests():
    """A tick of more requests than a resident worker keeps in LDS (256)
    reads its requests from device memory: 600 small Finds and Contains on
    two tables give the closed-form results.  They are submitted while an
    insert of 2^24 keys into a third table runs, so that most of them wait
    and are served in one tick."""
    import mxco
    ready()
    s = mxco.Server()
    tabs = [s.add_map(1 << 16), s.add_map(1 << 16)]
    big = s.add_map(1 << 25)
    keys = torch.arange(1, 1 << 14, dtype=torch.int64, device="cuda") * 3
    for t in tabs:
        assert s.wait(s.submit(t, mxco.REQ_INSERT, mxco.make_pairs(keys, keys * 7 + 1)))[0] == 0
    bkeys = torch.arange(1, (1 << 24) + 1, dtype=torch.int64, device="cuda")
    bpairs = mxco.make_pairs(bkeys, bkeys * 7 + 1)
    g = torch.Generator(device="cuda").manual_seed(2)
    reqs, checks = [], []
    for i in range(600):
        n = 1 + (i * 37) % 50
        q = keys[torch.randint(0, keys.numel(), (n,), device="cuda", generator=g)] + (i % 2)  # odd i: absent (not a multiple of 3)
        contains = i % 3 == 0
        out = torch.full((n,), 5, dtype=torch.uint8 if contains else torch.int64, device="cuda")
        reqs.append((tabs[i % 2], mxco.REQ_CONTAINS if contains else mxco.REQ_FIND, q.data_ptr(), n, out.data_ptr(), None, 0))
        checks.append((q, out, contains, i % 2 == 0))
    ready()
    ticket = s.submit(big, mxco.REQ_INSERT, bpairs)
    s.replay(reqs, [0.0] * len(reqs))
    assert s.wait(ticket)[0] == 0
    # each request of at most 50 keys is one chunk: a tick's read chunks count its requests
    largest = int(s.resident_trace()["rd_chunks"].max())
    s.close()
    torch.cuda.synchronize()
    assert largest > 256, largest
    for q, out, contains, present in checks:
        if contains:
            assert bool(out.bool().all()) == present and bool(out.bool().any()) == present
        else:
            exp = q * 7 + 1 if present else torch.full_like(q, -1)
            assert torch.equal(out, exp)
