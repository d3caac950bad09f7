"""Benchmark: StaticMap int64->int64 Insert + Find at load factor 0.8 on
2^28 requested slots (268,435,460 slots x 16 B = 4.3 GB) on one MI355X.

One step = Clear + Insert of round(0.8 * capacity) distinct uniform 64-bit
keys (value = key*7+1) + Find of 2^28 queries at a 50 % hit rate, all
device-resident before the timed region.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ops-collections_amd"))

PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--slots-log2", type=int, default=28)
    ap.add_argument("--lf", type=float, default=0.8)
    ap.add_argument("--group", type=int, default=0, help="lanes per key (0 = library default)")
    ap.add_argument("--blocks-per-cu", type=int, default=0, help="persistent grid blocks per CU (0 = default)")
    ap.add_argument("--inflight", type=int, default=0,
                    help="probe kernel variant: 1 plain, 2 two keys per group, 3 block-staged chunks, "
                         "4 wave-staged Find/Contains outputs (0 = library default, 4)")
    ap.add_argument("--probe-flags", type=int, default=-1, help="bit0 sector-sized first window, bit1 nt loads")
    ap.add_argument("--store-mode", type=int, default=-1, help="payload store after a key CAS (0 plain, 1 atomic, 2 nt)")
    ap.add_argument("--sweep", action="store_true", help="also sweep LF 0.5/0.7/0.8/0.9 and group widths")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain comparison")
    ap.add_argument("--no-streams", action="store_true", help="skip the Zipf duplicate-heavy insert stream")
    ap.add_argument("--no-filters", action="store_true", help="skip the BloomFilter and HyperLogLog legs")
    ap.add_argument("--hll-sketches", type=int, default=64)
    ap.add_argument("--no-roaring", action="store_true", help="skip the RoaringBitmap query leg")
    ap.add_argument("--no-containers", action="store_true",
                    help="skip the StaticMultimap / StaticMultiset / DynamicMap / Erase+Count leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving comparison")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (PMC byte-counter calibration)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def copy_calibration():
    import torch
    a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda").fill_(1)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copied 3 x 1 GiB")


def make_workload(torch, cap, lf, seed=1234):
    """round(lf * cap) distinct keys uniform over the signed 64-bit range
    (sentinel -1 excluded), value key*7+1; 2^28 queries (cap if smaller) at a
    50 % hit rate, misses uniform 64-bit keys guaranteed absent."""
    from mxco.workloads import uniform_map
    n_ins = int(round(lf * cap))
    n_find = 1 << 28 if cap >= (1 << 28) else cap
    keys, q, is_hit = uniform_map(n_ins, n_find, 0.5, seed)
    return keys, keys * 7 + 1, q, is_hit


def run_gpu(args, torch, mxco, lf, group, steps, warmup, verify=True):
    cap_req = 1 << args.slots_log2
    m = mxco.StaticMap(cap_req, -1, -1, torch.int64, torch.int64, group_size=group)
    cap = m.capacity()
    keys, vals, q, is_hit = make_workload(torch, cap, lf)
    pairs = mxco.make_pairs(keys, vals)
    out = torch.empty(q.numel(), dtype=torch.int64, device="cuda")
    s = torch.cuda.current_stream()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]

    def step(e=None):
        if e:
            e[0].record(s)
        m.clear_async()
        if e:
            e[1].record(s)
        m.insert_async(pairs)
        if e:
            e[2].record(s)
        m.find_async(q, out=out)
        if e:
            e[3].record(s)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step(ev[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_clear = sum(e[0].elapsed_time(e[1]) for e in ev) / steps
    t_ins = sum(e[1].elapsed_time(e[2]) for e in ev) / steps
    t_find = sum(e[2].elapsed_time(e[3]) for e in ev) / steps

    res = {"capacity": cap, "n_insert": keys.numel(), "n_find": q.numel(), "wall_ms_per_step": wall * 1e3 / steps,
           "clear_ms": t_clear, "insert_ms": t_ins, "find_ms": t_find, "group": group, "lf": lf,
           "insert_path": "region build" if m.insert_uses_region_build(keys.numel()) else "CAS walk"}
    if verify and res["insert_path"] == "region build":
        # the CAS walk on the same batch, for comparison (MXCO_OA_REGION_INSERT=0 is read at each call)
        os.environ["MXCO_OA_REGION_INSERT"] = "0"
        try:
            def clear_insert(e):
                m.clear_async()
                e[0].record(s)
                m.insert_async(pairs)
                e[1].record(s)
            evc = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps)]
            clear_insert(evc[0])
            for e in evc:
                clear_insert(e)
            torch.cuda.synchronize()
            res["insert_cas_walk_ms"] = sum(e[0].elapsed_time(e[1]) for e in evc) / steps
        finally:
            del os.environ["MXCO_OA_REGION_INSERT"]
    if verify:
        # Inserted keys hit with f(key); absent keys miss; no insert failed.
        exp = torch.where(is_hit, q * 7 + 1, torch.full_like(q, -1))
        res["verified"] = bool(torch.equal(out, exp))
        m.clear()
        lost, made = m.insert_stats(pairs)  # the CAS walk, counting its compare-and-swaps
        res["cas_walk_atomics"] = {"cas_made": made, "cas_lost": lost, "retry_rate": round(lost / max(made, 1), 6),
                                   "made_minus_lost_equals_keys": made - lost == keys.numel()}
        m.clear()
        res["insert_failed"] = m.insert(pairs)
        ln_f, sl_f = m.probe_stats(q)
        ln_i, sl_i = m.probe_stats(keys)
        res.update({"find_lines": ln_f, "find_slots": sl_f, "insert_lines": ln_i, "insert_slots": sl_i})
        res["insert_or_assign"] = run_insert_or_assign(torch, mxco, m, keys, steps)
    del m
    torch.cuda.empty_cache()
    return res


def _timed(torch, fn, steps):
    """Mean milliseconds of fn() over `steps` calls, bracketed by events on the current stream."""
    s = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in ev:
        a.record(s)
        fn()
        b.record(s)
    torch.cuda.synchronize()
    return sum(a.elapsed_time(b) for a, b in ev) / steps


def run_insert_or_assign(torch, mxco, m, keys, steps):
    """InsertOrAssign of 2^26 distinct keys already in the table, new value key*3+5."""
    n = min(1 << 26, keys.numel())
    k = keys[:n]
    pairs = mxco.make_pairs(k, k * 3 + 5)
    m.insert_or_assign_async(pairs)
    ms = _timed(torch, lambda: m.insert_or_assign_async(pairs), steps)
    failed, assigned = m.insert_or_assign(pairs)
    ok = failed == 0 and assigned == n and bool(torch.equal(m.find(k), k * 3 + 5))
    return {"n": n, "ms": round(ms, 4), "gkeys_s": round(n / ms / 1e6, 4), "failed": failed, "assigned": assigned,
            "verified": ok}


def run_zipf_insert(args, torch, mxco, steps=3):
    """Duplicate-heavy insert stream: 2^27 draws of Zipf(s=1.1) over 2^24
    distinct keys into an empty table of the same 2^28 slots (value = key*7+1,
    so first-writer-wins races leave the same table)."""
    from mxco.workloads import Zipf
    n, distinct_max = 1 << 27, 1 << 24
    z = Zipf(distinct_max, 1.1)
    ranks = z.draw(n, seed=11)
    seen = torch.zeros(distinct_max, dtype=torch.bool, device="cuda")
    seen[ranks] = True
    distinct = int(seen.sum().item())
    keys = Zipf.keys(ranks)
    del ranks, z
    pairs = mxco.make_pairs(keys, keys * 7 + 1)
    m = mxco.StaticMap(1 << args.slots_log2, -1, -1, torch.int64, torch.int64)

    def step():
        m.clear_async()
        m.insert_async(pairs)

    step()
    s = torch.cuda.current_stream()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps)]
    for e in ev:
        m.clear_async()
        e[0].record(s)
        m.insert_async(pairs)
        e[1].record(s)
    torch.cuda.synchronize()
    ms = sum(a.elapsed_time(b) for a, b in ev) / steps
    m.clear()
    lost, made = m.insert_stats(pairs)
    m.clear()
    not_inserted = m.insert(pairs)
    uk = Zipf.keys(torch.nonzero(seen).flatten())
    ok = not_inserted == n - distinct and bool(torch.equal(m.find(uk), uk * 7 + 1))
    del m, pairs, keys
    torch.cuda.empty_cache()
    return {"draws": n, "distinct": distinct, "s": 1.1, "ms": round(ms, 4), "gkeys_s": round(n / ms / 1e6, 4),
            "duplicates_reported": not_inserted, "verified": ok,
            "cas_walk_atomics": {"cas_made": made, "cas_lost": lost, "retry_rate": round(lost / max(made, 1), 6)}}


def run_bloom(torch, mxco, steps=5):
    """BloomFilter of 2^22 blocks x 32 B (2^30 bits): Add 2^27 uniform uint64
    keys, Contains 2^27 queries (first half added, second half absent)."""
    nb, n = 1 << 22, 1 << 27
    g = torch.Generator(device="cuda")
    g.manual_seed(21)
    keys = torch.randint(-(1 << 63), (1 << 63) - 1, (n,), dtype=torch.int64, device="cuda", generator=g)
    absent = torch.randint(-(1 << 63), (1 << 63) - 1, (n - n // 2,), dtype=torch.int64, device="cuda", generator=g)
    q = torch.cat([keys[: n // 2], absent])
    out = torch.empty(n, dtype=torch.bool, device="cuda")
    bf = mxco.BloomFilter(nb, torch.int64)
    bf.add(keys)
    s = torch.cuda.current_stream()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    for e in ev:
        bf.clear(sync=False)
        e[0].record(s)
        bf.add_async(keys)
        e[1].record(s)
        bf.contains(q, out=out, sync=False)
        e[2].record(s)
    torch.cuda.synchronize()
    add_ms = sum(e[0].elapsed_time(e[1]) for e in ev) / steps
    con_ms = sum(e[1].elapsed_time(e[2]) for e in ev) / steps
    no_false_neg = bool(out[: n // 2].all().item())
    fpr = out[n // 2:].float().mean().item()
    add_gbs = n * (32 + 32 + 8) / (add_ms * 1e-3) / 1e9
    con_gbs = n * (32 + 8 + 1) / (con_ms * 1e-3) / 1e9
    del bf
    return {"blocks": nb, "keys": n, "add_ms": round(add_ms, 4), "contains_ms": round(con_ms, 4),
            "add_gkeys_s": round(n / add_ms / 1e6, 3), "contains_gkeys_s": round(n / con_ms / 1e6, 3),
            "fpr": round(fpr, 5), "no_false_negatives": no_false_neg,
            "roofline_add": {"bound": "hbm", "achieved": round(add_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": round(add_gbs / PEAK_HBM_GBS, 4), "bytes_per_key": 72},
            "roofline_contains": {"bound": "hbm", "achieved": round(con_gbs, 1), "peak": PEAK_HBM_GBS,
                                  "unit": "GB/s", "frac": round(con_gbs / PEAK_HBM_GBS, 4), "bytes_per_key": 41}}


def run_hll(torch, mxco, sketches=64, n=1 << 28):
    """HyperLogLog with a 32 KiB sketch (p = 15): `sketches` sketches of n
    int64 keys each, Zipf(s=1.0) over 1e8 distinct values with disjoint seeds,
    merged into one; the estimate is compared with the exact distinct count."""
    from mxco.workloads import Zipf
    universe = 10 ** 8
    z = Zipf(universe, 1.0)
    seen = torch.zeros(universe, dtype=torch.bool, device="cuda")
    total = mxco.HyperLogLog.create_with_sketch_size_kb(32, torch.int64)
    h = mxco.HyperLogLog(total.precision, torch.int64)
    regmax = torch.zeros(1 << total.precision, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream()
    add_ms, merge_ms = [], []
    for k in range(sketches):
        r = z.draw(n, seed=1000 + k)
        seen[r] = True
        keys = Zipf.keys(r)
        del r
        h.clear()
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record(s)
        h.add(keys)
        e[1].record(s)
        total.merge(h)
        e[2].record(s)
        torch.cuda.synchronize()
        regmax = torch.maximum(regmax, h.registers().view(torch.uint8))
        add_ms.append(e[0].elapsed_time(e[1]))
        merge_ms.append(e[1].elapsed_time(e[2]))
        del keys
    est = total.estimate()
    exact = int(seen.sum().item())
    merged_ok = bool(torch.equal(total.registers().view(torch.uint8), regmax))
    ms = sum(add_ms[1:]) / max(1, len(add_ms) - 1)  # first add includes one-time set-up
    gbs = n * 8 / (ms * 1e-3) / 1e9
    rel = abs(est - exact) / exact
    tol = 3 * 1.04 / math.sqrt(1 << total.precision)
    del z, seen
    torch.cuda.empty_cache()
    return {"precision": total.precision, "sketches": sketches, "keys_per_sketch": n, "add_ms": round(ms, 4),
            "add_gkeys_s": round(n / ms / 1e6, 3), "merge_ms": round(sum(merge_ms) / len(merge_ms), 4),
            "estimate": est, "exact_distinct": exact, "rel_error": round(rel, 5), "tolerance": round(tol, 5),
            "merge_is_register_max": merged_ok, "within_tolerance": rel <= tol,
            "roofline_add": {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": round(gbs / PEAK_HBM_GBS, 4), "bytes_per_key": 8}}


def run_ceilings(torch, mxco, reps=5):
    """Measured ceilings of the containers' access patterns: random 128-,
    64- and 32-byte reads over an 8 GiB buffer (HBM; the probe walk reads
    one 128-byte line per window), random 256-byte reads (two adjacent
    lines: whether a walk's next line is cheaper than a random one), random
    32- and 128-byte reads over a
    128 MiB buffer (the Bloom filter's size: served from the 256 MB MALL),
    and a 4 GiB streaming copy.  Median of `reps` timed runs each."""
    import statistics
    s = torch.cuda.current_stream()
    sink = torch.empty(mxco.cu_count() * 8 * 256, dtype=torch.int64, device="cuda")
    big = torch.ones(8 << 30, dtype=torch.uint8, device="cuda")
    out = {}

    def timed(fn):
        ms = []
        fn()
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(s)
            fn()
            b.record(s)
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        return statistics.median(ms)

    for name, nbytes, access, n in (("hbm_random_128B", 8 << 30, 128, 1 << 28),
                                    ("hbm_random_64B", 8 << 30, 64, 1 << 28),
                                    ("hbm_random_32B", 8 << 30, 32, 1 << 28),
                                    ("hbm_random_256B", 8 << 30, 256, 1 << 27),
                                    ("mall_random_32B_128MiB", 128 << 20, 32, 1 << 27),
                                    ("mall_random_128B_128MiB", 128 << 20, 128, 1 << 27)):
        ms = timed(lambda: mxco.random_read(big, nbytes, access, n, sink))
        out[name] = {"ms": round(ms, 4), "gb_s": round(n * access / (ms * 1e-3) / 1e9, 1),
                     "accesses": n, "buffer_bytes": nbytes}
    half = 2 << 30
    ms = timed(lambda: mxco.stream_copy(big, big[half:], half))
    out["hbm_stream_copy"] = {"ms": round(ms, 4), "gb_s": round(2 * half / (ms * 1e-3) / 1e9, 1),
                              "bytes_moved": 2 * half}
    del big, sink
    torch.cuda.empty_cache()
    return out


def run_roaring(torch, mxco, steps=5):
    """RoaringBitmap Contains of 2^28 uint32 queries on two bitmaps:
    "spec", the RoaringFormatSpec key set of the reference's tests (200,100
    keys in 11 array/bitset/run containers, 48 KB) queried uniformly over
    [0, 2^20); "wide", 2^26 uniform keys over the whole 2^32 range (65,536
    array containers of ~1,024 values, 128 MiB) queried half from the keys,
    half uniformly.  Each query reads a 4-byte key and writes a 1-byte
    result; the library expands both bitmaps to 8 KiB bitsets per container,
    so a query whose container exists also reads one 128-byte line of its
    bitset.  "spec"'s 88 KiB image stays in L2 (5 B/query from HBM); "wide"'s
    512 MiB image does not (4 + 1 + 128 B/query: all 65,536 containers
    exist)."""
    import numpy as np
    from mxco import workloads as W
    n = 1 << 28
    g = torch.Generator(device="cuda")
    g.manual_seed(41)
    out = torch.empty(n, dtype=torch.bool, device="cuda")
    res = {}
    spec = (list(range(0, 100000, 1000)) + [3 * x for x in range(100000, 200000)]
            + list(range(700000, 800000)))
    wide = np.random.default_rng(41).integers(0, 1 << 32, 1 << 26, dtype=np.uint64)
    for name, keys, lo_hi, qbytes in (("spec", spec, (0, 1 << 20), 5), ("wide", wide, None, 133)):
        bm = W.roaring_portable32(keys, "optimize")
        rb = mxco.RoaringBitmap(bm, torch.int32)
        if lo_hi:
            q = torch.randint(lo_hi[0], lo_hi[1], (n,), dtype=torch.int64, device="cuda", generator=g)
        else:
            kd = torch.from_numpy(np.unique(keys).astype(np.int64)).cuda()
            pick = torch.randint(0, kd.numel(), (n // 2,), device="cuda", generator=g)
            q = torch.cat([kd[pick], torch.randint(0, 1 << 32, (n - n // 2,), dtype=torch.int64, device="cuda",
                                                   generator=g)])
            del kd
        q = torch.where(q >= 1 << 31, q - (1 << 32), q).to(torch.int32)
        rb.contains(q, out=out)
        ms = _timed(torch, lambda: rb.contains(q, out=out, sync=False), steps)
        # spot check of 2^16 results against the key set
        qs = q[: 1 << 16].cpu().numpy().view(np.uint32).astype(np.uint64)
        ok = bool(np.array_equal(out[: 1 << 16].cpu().numpy(), np.isin(qs, np.asarray(keys, dtype=np.uint64))))
        gbs = n * qbytes / (ms * 1e-3) / 1e9
        res[name] = {"bitmap_bytes": len(bm), "keys": rb.size(), "queries": n, "contains_ms": round(ms, 4),
                     "contains_gkeys_s": round(n / ms / 1e6, 3), "hit_rate": round(out.float().mean().item(), 4),
                     "spot_check": ok,
                     "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                  "frac": round(gbs / PEAK_HBM_GBS, 4), "bytes_per_key": qbytes}}
        del rb, q
    return res


def run_containers(torch, mxco, steps=3):
    """The other containers at one shape each, int64 keys, verified:
    - StaticMultimap (int64 -> int64): capacity 2^26, Insert 2^25 pairs over
      2^23 distinct keys (Poisson(4) multiplicities), Count and Find of 2^24
      queries (half present), Retrieve of 2^20 present queries;
    - StaticMultiset: the same keys, Insert, CountEach and Contains of 2^24;
    - DynamicMap: initial capacity 2^22, 2^26 distinct pairs inserted in 16
      batches of 2^22 (the map grows), then Find of 2^26 (half present);
    - StaticMap Erase and Count on a 2^26-slot table at LF 0.8.
    Rates are keys (or queries) per second of the op's launches; the lower-
    bound bytes count one 128-byte line per key or query beside its input
    and output."""
    g = torch.Generator(device="cuda")
    g.manual_seed(77)
    res = {}

    def rate(n, ms):
        return round(n / ms / 1e6, 3)

    def lb(n, per, ms):  # lower-bound GB/s and fraction of the HBM peak
        gbs = n * per / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved_lower_bound": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac_lower_bound": round(gbs / PEAK_HBM_GBS, 4), "bytes_per_key": per}

    # keys with multiplicities: index r in [0, D) -> key r * ZMUL + 1
    D, N, Q = 1 << 23, 1 << 25, 1 << 24
    ZMUL = 0x9E3779B97F4A7C1
    r = torch.randint(0, D, (N,), device="cuda", generator=g)
    mult = torch.bincount(r, minlength=D)
    keys = r * ZMUL + 1
    vals = keys * 3 + r
    qi = torch.cat([torch.randint(0, D, (Q // 2,), device="cuda", generator=g),
                    torch.randint(D, 2 * D, (Q - Q // 2,), device="cuda", generator=g)])
    q = qi * ZMUL + 1
    exp_mult = torch.where(qi < D, mult[qi.clamp(max=D - 1)], torch.zeros_like(qi))

    mm = mxco.StaticMultimap(1 << 26, -1, -1, torch.int64, torch.int64)
    pairs = mxco.make_pairs(keys, vals)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    failed = mm.insert(pairs)
    ev1.record()
    torch.cuda.synchronize()
    t_ins = ev0.elapsed_time(ev1)
    t_cnt = _timed(torch, lambda: mm.count(q), steps)
    total = mm.count(q)
    out = torch.empty(Q, dtype=torch.int64, device="cuda")
    t_find = _timed(torch, lambda: mm.find(q, out=out), steps)
    present = qi < D
    rq = q[present][: 1 << 20].contiguous()
    cap_out = int(mm.count(rq))
    po = torch.empty(cap_out, dtype=torch.int64, device="cuda")
    vo = torch.empty(cap_out, dtype=torch.int64, device="cuda")
    t_ret = _timed(torch, lambda: mm.retrieve(rq, po, vo, cap_out), 1)
    got = mm.retrieve(rq, po, vo, cap_out)
    exp_find = torch.where(exp_mult > 0, q * 3 + qi, torch.full_like(q, -1))  # one value per key
    ok = (failed == 0 and total == int(exp_mult.sum().item()) and got == cap_out
          and bool(torch.equal(out, exp_find)))
    res["multimap"] = {"capacity": mm.capacity(), "inserted": N, "distinct": int((mult > 0).sum().item()),
                       "queries": Q, "insert_ms": round(t_ins, 3), "insert_gkeys_s": rate(N, t_ins),
                       "count_ms": round(t_cnt, 3), "count_gkeys_s": rate(Q, t_cnt),
                       "find_ms": round(t_find, 3), "find_gkeys_s": rate(Q, t_find),
                       "retrieve_queries": rq.numel(), "retrieve_matches": got, "retrieve_ms": round(t_ret, 3),
                       "verified": bool(ok), "roofline_find": lb(Q, 128 + 16, t_find),
                       "roofline_insert": lb(N, 128 + 128 + 16, t_ins)}
    del mm, po, vo

    log("bench: containers: multiset")
    ms = mxco.StaticMultiset(1 << 26, -1, torch.int64)
    ev0.record()
    failed = ms.insert(keys)
    ev1.record()
    torch.cuda.synchronize()
    t_ins = ev0.elapsed_time(ev1)
    ce = torch.empty(Q, dtype=torch.int64, device="cuda")
    t_ce = _timed(torch, lambda: ms.count_each(q, out=ce), steps)
    con = torch.empty(Q, dtype=torch.bool, device="cuda")
    t_con = _timed(torch, lambda: ms.contains(q, out=con), steps)
    ok = failed == 0 and bool(torch.equal(ce.to(torch.int64), exp_mult)) and bool(torch.equal(con, exp_mult > 0))
    res["multiset"] = {"capacity": ms.capacity(), "inserted": N, "queries": Q, "insert_ms": round(t_ins, 3),
                       "insert_gkeys_s": rate(N, t_ins), "count_each_ms": round(t_ce, 3),
                       "count_each_gkeys_s": rate(Q, t_ce), "contains_ms": round(t_con, 3),
                       "contains_gkeys_s": rate(Q, t_con), "verified": bool(ok),
                       "roofline_count_each": lb(Q, 128 + 16, t_ce)}
    del ms, ce, con, keys, vals, pairs, r, mult

    # DynamicMap growth
    log("bench: containers: dynamic map")
    NB, B = 16, 1 << 22
    dk = torch.unique(torch.randint(0, 1 << 62, (NB * B + (1 << 16),), device="cuda", generator=g))[: NB * B]
    dk = dk[torch.randperm(dk.numel(), device="cuda", generator=g)]
    dpairs = mxco.make_pairs(dk, dk * 7 + 1)
    dm = mxco.DynamicMap(1 << 22, -1, -1, -2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for b in range(NB):
        dm.insert(dpairs[b * B:(b + 1) * B])
    torch.cuda.synchronize()
    t_dins = (time.perf_counter() - t0) * 1e3
    dq = torch.cat([dk[torch.randint(0, dk.numel(), (dk.numel() // 2,), device="cuda", generator=g)],
                    torch.randint(1 << 62, (1 << 63) - 1, (dk.numel() // 2,), device="cuda", generator=g)])
    dout = torch.empty(dq.numel(), dtype=torch.int64, device="cuda")
    t_dfind = _timed(torch, lambda: dm.find(dq, out=dout), steps)
    ok = bool(torch.equal(dout, torch.where(dq < (1 << 62), dq * 7 + 1, torch.full_like(dq, -1)))) and \
        dm.size() == dk.numel()
    res["dynamic_map"] = {"initial_capacity": 1 << 22, "batches": NB, "batch": B, "final_capacity": dm.capacity(),
                          "submaps": dm.num_submaps(), "insert_ms": round(t_dins, 3),
                          "insert_gkeys_s": rate(NB * B, t_dins), "find_ms": round(t_dfind, 3),
                          "find_gkeys_s": rate(dq.numel(), t_dfind), "verified": ok,
                          "note": "insert time is wall clock over the 16 synchronous batches, growth included"}
    del dm, dk, dpairs, dq, dout

    # StaticMap Erase + Count
    log("bench: containers: erase, count, insert_and_find, contains")
    m = mxco.StaticMap(1 << 26, -1, -1, torch.int64, torch.int64)
    m.set_erased_key(-2)
    ek, ev, eq, e_hit = make_workload(torch, m.capacity(), 0.8, seed=99)
    m.insert(mxco.make_pairs(ek, ev))
    t_c = _timed(torch, lambda: m.count(eq), steps)
    present = int(m.count(eq))
    cq = torch.empty(eq.numel(), dtype=torch.bool, device="cuda")
    t_con = _timed(torch, lambda: m.contains(eq, out=cq, sync=False), steps)
    ok_con = int(cq.sum().item()) == eq.numel() // 2
    # Contains, then InsertAndFind with half the pairs already present, on
    # the table at load factor 0.8 (before the Erase).  Only 2^22 new keys are
    # added (0.86 after): a batch larger than the EMPTY slots left would walk
    # full cycles.
    iaf_keys = torch.cat([ek[ek.numel() // 2:ek.numel() // 2 + (1 << 22)], eq[~e_hit][:1 << 22]])
    iaf_pairs = mxco.make_pairs(iaf_keys, iaf_keys * 7 + 1)
    of = torch.empty(iaf_keys.numel(), dtype=torch.int64, device="cuda")
    oi = torch.empty(iaf_keys.numel(), dtype=torch.bool, device="cuda")
    ev0.record()
    m.insert_and_find(iaf_pairs, out_find=of, out_inserted=oi)
    ev1.record()
    torch.cuda.synchronize()
    t_iaf = ev0.elapsed_time(ev1)
    ok_iaf = bool(torch.equal(of, iaf_keys * 7 + 1)) and int(oi.sum().item()) == (1 << 22)
    res["insert_and_find"] = {"keys": iaf_keys.numel(), "already_present": 1 << 22, "ms": round(t_iaf, 3),
                              "gkeys_s": rate(iaf_keys.numel(), t_iaf), "verified": ok_iaf}
    res["contains"] = {"queries": eq.numel(), "ms": round(t_con, 3), "gkeys_s": rate(eq.numel(), t_con),
                       "verified": ok_con}
    half = ek[: ek.numel() // 2].contiguous()
    ev0.record()
    failed = m.erase(half)
    ev1.record()
    torch.cuda.synchronize()
    t_e = ev0.elapsed_time(ev1)
    left = int(m.count(ek))
    ok = failed == 0 and present == eq.numel() // 2 and left == ek.numel() - half.numel()
    res["erase_count"] = {"capacity": m.capacity(), "inserted": ek.numel(), "count_queries": eq.numel(),
                          "count_ms": round(t_c, 3), "count_gkeys_s": rate(eq.numel(), t_c),
                          "erase_keys": half.numel(), "erase_ms": round(t_e, 3),
                          "erase_gkeys_s": rate(half.numel(), t_e), "verified": bool(ok)}
    del m
    torch.cuda.empty_cache()
    return res


def run_batched(args, torch, mxco, steps=3, warmup=1):
    """1024 StaticMap int32 tables, capacities log-uniform 2^10..2^22, LF
    U[0.5,0.85], finds log-uniform 2^6..2^20 at 50 % hit: one graph replay
    (a BUILD plan -- Clear + Insert built region by region -- and one
    persistent Find launch) against 1024 plain per-table Clear/Insert/Find
    calls.  The graph of Clear + the persistent CAS Insert + Find is timed
    beside it (insert_plan_ms, against the plain CAS calls), and the plain
    calls again with the region build open to every table large enough
    (plain_region_ms, against the BUILD plan)."""
    from mxco import workloads
    w = workloads.batched_tables(1024, (10, 22), (0.5, 0.85), (6, 20), seed=4,
                                 valid_extent=lambda c: mxco.valid_extent(c, 5, "linear"))
    T = len(w["caps"])
    pairs, q = w["pairs"], w["queries"]
    b = mxco.BatchedStaticMap(w["caps"], -1, -1)
    items = [(t, w["ins_off"][t], w["n_ins"][t]) for t in range(T)]
    ins = b.plan(mxco.BATCH_INSERT, items)
    bld = b.plan(mxco.BATCH_BUILD, items)
    fnd = b.plan(mxco.BATCH_FIND, [(t, w["find_off"][t], w["n_find"][t]) for t in range(T)])
    out_b = torch.empty_like(q)
    # the insert tuner leaves the tables filled, so the find plan is tuned on them
    tuned = {"insert": ins.tune(pairs), "find": fnd.tune(q, out_b), "build": bld.tune(pairs)}

    def timed_graph(record):
        out_b.fill_(0)
        g = mxco.Graph(record)
        for _ in range(warmup):
            g.launch()
        g.stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.launch()
        g.stream.synchronize()
        t = (time.perf_counter() - t0) / steps
        ok = bool(torch.equal(out_b, w["expected"]))
        return t, ok, g

    # BUILD (Clear + Insert in one region-built pass) + Find
    def record_build(s):
        bld.run(pairs, stream=s)
        fnd.run(q, out_b, stream=s)

    t_batched, ok_b, g = timed_graph(record_build)
    ok_b = ok_b and bld.failures(g.stream) == [0] * T
    del g

    # Clear + Insert (CAS walk) + Find
    def record_insert(s):
        b.clear(s)
        ins.run(pairs, stream=s)
        fnd.run(q, out_b, stream=s)

    t_cas, ok_c, g = timed_graph(record_insert)
    ok_b = ok_b and ok_c and ins.failures(g.stream) == [0] * T
    del g
    # plain: one StaticMap per table, the reference's synchronous calls
    maps = [mxco.StaticMap(w["caps"][t], -1, -1) for t in range(T)]
    out_p = torch.empty_like(q)
    views = [(pairs[w["ins_off"][t]:w["ins_off"][t] + w["n_ins"][t]],
              q[w["find_off"][t]:w["find_off"][t] + w["n_find"][t]],
              out_p[w["find_off"][t]:w["find_off"][t] + w["n_find"][t]]) for t in range(T)]

    def plain_step():
        failed = 0
        for m, (pv, qv, ov) in zip(maps, views):
            m.clear_async()
            failed += m.insert(pv)
            m.find(qv, out=ov)
        return failed

    def timed_plain():
        out_p.fill_(0)
        for _ in range(warmup):
            plain_step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        failed = 0
        for _ in range(steps):
            failed += plain_step()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / steps
        return t, failed == 0 and bool(torch.equal(out_p, w["expected"]))

    # the plain calls as they come (CAS walk: these tables are below the
    # region build's size threshold), then with the region build opened to
    # every table (it still needs n >= 2^20 and n >= capacity / 2), the
    # same insert algorithm the BUILD plan runs
    t_plain, ok_p = timed_plain()
    key = "MXCO_OA_REGION_INSERT_MIN_BYTES"
    saved = os.environ.get(key)
    os.environ[key] = "0"
    try:
        t_plain_region, ok_pr = timed_plain()
    finally:
        if saved is None:
            del os.environ[key]
        else:
            os.environ[key] = saved
    region_tables = sum(1 for t in range(T) if w["n_ins"][t] >= (1 << 20) and 2 * w["n_ins"][t] >= w["caps"][t])
    keys = sum(w["n_ins"]) + sum(w["n_find"])
    cfg = {name: {"group": pl.config()[0], "blocks_per_cu": pl.config()[1], "block": pl.launch_shape()[0],
                  "wave_staged": pl.launch_shape()[1] if name == "find" else None}
           for name, pl in (("insert", ins), ("find", fnd))}
    del maps, b, ins, bld, fnd
    torch.cuda.empty_cache()
    return {"tables": T, "slots": sum(w["caps"]), "inserts": sum(w["n_ins"]), "finds": sum(w["n_find"]),
            "batched_ms": round(t_batched * 1e3, 3), "plain_ms": round(t_plain * 1e3, 3),
            "batched_gkeys_s": round(keys / t_batched / 1e9, 4), "plain_gkeys_s": round(keys / t_plain / 1e9, 4),
            "speedup": round(t_plain / t_batched, 3), "identical_results": ok_b and ok_p and ok_pr,
            "insert_plan_ms": round(t_cas * 1e3, 3), "insert_plan_speedup": round(t_plain / t_cas, 3),
            "plain_region_ms": round(t_plain_region * 1e3, 3), "plain_region_tables": region_tables,
            "speedup_vs_plain_region": round(t_plain_region / t_batched, 3),
            "tuned_config": cfg,
            "tune_ms": {k: round(v, 3) for k, v in tuned.items()}}


def run_serving(args, torch, mxco, qps_list=(20000.0, 100000.0, 200000.0), num_requests=3000):
    """64 resident tables (48 StaticMap + 16 StaticMultimap int64, 2^20..2^26
    slots, half full), open-loop Poisson request stream (80 % Find/Contains,
    15 % Insert on the maps, 5 % Retrieve on the multimaps; batches
    log-uniform 1..2^16 keys, 70 % hit): the continuously batching server
    (StaticMap ticks fed to the resident kernel), the same server with one
    insert and one read launch per tick, and one-at-a-time plain calls;
    latency percentiles over all requests and per table kind; then the closed
    loop: every request of the 200k-QPS stream submitted at once, keys over
    the makespan (saturation)."""
    import numpy as np
    from mxco import workloads
    ops = {"insert": mxco.REQ_INSERT, "find": mxco.REQ_FIND, "contains": mxco.REQ_CONTAINS,
           "retrieve": mxco.REQ_MM_RETRIEVE}

    def serve(w, mode, arrivals):
        # "batched": the default server (StaticMap ticks in the resident kernel);
        # "launched": one insert and one read launch per tick; "plain": one by one
        torch.cuda.synchronize()  # requests carry device pointers: their inputs are written before submit
        srv = mxco.Server(resident=(mode != "launched"))
        tabs = [srv.add_map(c) if t < w["num_maps"] else srv.add_multimap(c) for t, c in enumerate(w["caps"])]
        for t, pairs in enumerate(w["init"]):
            op = mxco.REQ_INSERT if t < w["num_maps"] else mxco.REQ_MM_INSERT
            srv.wait(srv.submit(tabs[t], op, pairs))
        outs, reqs = [], []
        for r in w["requests"]:
            n = r["n"]
            if r["op"] == "retrieve":
                o1 = torch.zeros(9 * n, dtype=torch.int64, device="cuda")
                o2 = torch.zeros(9 * n, dtype=torch.int64, device="cuda")
            elif r["op"] == "find":
                o1, o2 = torch.empty(n, dtype=torch.int64, device="cuda"), None
            elif r["op"] == "contains":
                o1, o2 = torch.empty(n, dtype=torch.uint8, device="cuda"), None
            else:
                o1 = o2 = None
            outs.append((o1, o2))
            reqs.append((tabs[r["table"]], ops[r["op"]], r["input"].data_ptr(), n,
                         o1.data_ptr() if o1 is not None else None, o2.data_ptr() if o2 is not None else None,
                         o1.numel() if r["op"] == "retrieve" else 0))
        torch.cuda.synchronize()
        launches0 = srv.resident_stats()[1]
        ticks0 = srv.stats()[0]
        results, lat = srv.replay(reqs, arrivals, plain=(mode == "plain"))
        ticks, folded = srv.stats()
        resident, launches, grid = srv.resident_stats()
        serve.kernel = {"resident_kernel": resident, "resident_launches": launches - launches0}
        if resident and mode == "batched" and ticks > ticks0:
            # the resident kernel's own timing of this replay's ticks
            tr = srv.resident_trace(min(ticks - ticks0, 1024))
            hand = tr["handoff_us"]  # ticks staged before the previous one finished
            serve.kernel.update({"resident_grid": grid, "tick_run_us_p50": round(float(np.median(tr["run_us"])), 1),
                                 "tick_handoff_us_p50": round(float(np.median(hand)), 2) if len(hand) else None,
                                 "ticks_staged_ahead": int(len(hand))})
        srv.close()
        del srv
        return results, outs, np.array(lat) / 1e3, ticks, folded

    def same(a, b):
        ok = a[0] == b[0]
        for (a1, a2), (b1, b2) in zip(a[1], b[1]):
            if a1 is not None:
                ok &= bool(torch.equal(a1, b1))
            if a2 is not None:
                ok &= bool(torch.equal(a2, b2))
        return bool(ok)

    def pct(x, q):
        return round(float(np.percentile(x, q)), 4) if len(x) else None

    out = {}
    for qps in qps_list:
        w = workloads.serving(num_requests=num_requests, qps=qps, seed=5)
        keys = sum(r["n"] for r in w["requests"])
        mm = np.array([r["op"] == "retrieve" for r in w["requests"]])
        res, runs = {}, {}
        for mode in ("batched", "launched", "plain"):
            results, outs, lat, ticks, folded = serve(w, mode, w["arrivals_us"])
            done = np.array(w["arrivals_us"]) / 1e3 + lat
            res[mode] = {"p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                         "map_p99_ms": pct(lat[~mm], 99), "multimap_p99_ms": pct(lat[mm], 99),
                         "gkeys_s": round(keys / (done.max() * 1e-3) / 1e9, 4),
                         "ticks": ticks, "requests_per_tick": round(folded / max(ticks, 1), 2)}
            if mode != "plain":
                res[mode].update(serve.kernel)
            runs[mode] = (results, outs)
        res["identical_results"] = same(runs["batched"], runs["plain"]) and same(runs["launched"], runs["plain"])
        res["keys"] = keys
        out["qps_%d" % int(qps)] = res
        del runs
        if qps == qps_list[-1]:
            # closed loop: the same requests, all arriving at once
            closed, runs = {"requests": len(w["requests"]), "keys": keys}, {}
            for mode in ("batched", "launched", "plain"):
                results, outs, lat, ticks, folded = serve(w, mode, [0.0] * len(w["requests"]))
                closed[mode] = {"makespan_ms": round(float(lat.max()), 3),
                                "gkeys_s": round(keys / (lat.max() * 1e-3) / 1e9, 4),
                                "map_makespan_ms": round(float(lat[~mm].max()), 3),
                                "multimap_makespan_ms": round(float(lat[mm].max()), 3),
                                "requests_per_tick": round(folded / max(ticks, 1), 2)}
                if mode != "plain":
                    closed[mode].update(serve.kernel)
                runs[mode] = (results, outs)
            closed["identical_results"] = same(runs["batched"], runs["plain"]) and same(runs["launched"],
                                                                                          runs["plain"])
            out["closed_loop"] = closed
            del runs
        del w
        torch.cuda.empty_cache()
    return out


def _cpu_description():
    """What /proc/cpuinfo and sysfs say about the host CPU."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if info:
                    break  # the first processor's block
                continue
            k, v = (x.strip() for x in line.split(":", 1))
            if k in ("vendor_id", "model name", "cpu family", "model", "stepping", "cache size", "cpu MHz",
                     "cpu cores", "siblings") and k not in info:
                info[k] = v
            if k == "flags":
                info["flags_count"] = len(v.split())
    except OSError:
        pass
    try:
        info["max_mhz"] = int(open("/sys/devices/system/cpu/cpu0/cpufreq/cpuinfo_max_freq").read()) // 1000
    except (OSError, ValueError):
        pass
    info["logical_cpus_visible"] = os.cpu_count()
    info["affinity_cpus"] = len(os.sched_getaffinity(0))
    return info


def cpu_baseline(args, runs=5):
    """The oracle's restatements of the reference algorithms on the host
    cores, each leg the median of `runs` timed runs on a bounded sample of
    the measured workload (about 20 s together):
      - StaticMap int64: Insert of round(0.8 * cap) distinct keys then Find of
        cap queries (50 % hit) on a 2^25-slot table, T threads (pthreads,
        __atomic 64-bit CAS), and on a 2^21-slot table with one thread;
      - the reference tests' std::unordered_map golden path, 2^22 sample;
      - BloomFilter: Add of 2^26 keys and Contains of 2^26 (half absent) on
        2^22 blocks (the 128 MiB filter), T threads;
      - HyperLogLog p = 15: Add of 2^27 int64 keys, T threads (private
        registers merged by maximum).
    T is the host's CPU share for this process: OMP_NUM_THREADS when set
    (the measurement host sets it to the share of one GPU), else every CPU
    in the affinity mask."""
    import statistics
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    share = os.environ.get("OMP_NUM_THREADS")
    threads = args.cpu_threads or (int(share) if share and share.isdigit() else len(os.sched_getaffinity(0)))

    def median_run(fn):
        """fn() -> (seconds, verified); the median time and all verified."""
        res = [fn() for _ in range(runs)]
        return statistics.median(r[0] for r in res), all(r[1] for r in res), [round(r[0], 4) for r in res]

    def map_leg(cap_log2, nthreads):
        cap_req = 1 << cap_log2
        rng = np.random.default_rng(5)
        probe = O.OpenAddressing(np.int64, np.int64, cap_req, -1, -1)
        n_ins = int(round(0.8 * probe.capacity))
        del probe
        # distinct keys below 2^62: an odd multiplier is a bijection mod 2^62
        keys = ((np.arange(n_ins, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)) & np.uint64((1 << 62) - 1))
        keys = keys.astype(np.int64)
        vals = keys * 7 + 1
        q = np.concatenate([keys[rng.integers(0, n_ins, cap_req // 2)],
                            rng.integers(1 << 62, (1 << 63) - 1, cap_req // 2)])
        rng.shuffle(q)
        exp = np.where(q < (1 << 62), q * 7 + 1, -1)

        def once():
            t = O.OpenAddressing(np.int64, np.int64, cap_req, -1, -1)
            t0 = time.perf_counter()
            failed = t.mt_insert(keys, vals, nthreads)
            out = t.mt_find(q, nthreads)
            dt = time.perf_counter() - t0
            return dt, failed == 0 and np.array_equal(out, exp)

        sec, ok, all_s = median_run(once)
        return {"value": (n_ins + len(q)) / sec / 1e9, "unit": "Gkeys/s", "cores": nthreads,
                "median_s": round(sec, 4), "runs_s": all_s, "verified": ok,
                "sample": "Insert of %d keys + Find of %d queries (50%% hit) on a 2^%d-slot table at LF 0.8"
                          % (n_ins, len(q), cap_log2)}

    def golden_leg(cap_log2):
        """The reference tests' host path: a std::unordered_map filled and
        queried on one thread (oracle/libgolden_map.so)."""
        import ctypes
        lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "libgolden_map.so"))
        cap_req = 1 << cap_log2
        n_ins = int(round(0.8 * cap_req))
        rng = np.random.default_rng(5)
        keys = ((np.arange(n_ins, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
                & np.uint64((1 << 62) - 1)).astype(np.int64)
        vals = keys * 7 + 1
        q = np.concatenate([keys[rng.integers(0, n_ins, cap_req // 2)],
                            rng.integers(1 << 62, (1 << 63) - 1, cap_req // 2)])
        rng.shuffle(q)
        P = ctypes.c_void_p

        def once():
            out = np.empty_like(q)
            ins_s, find_s = ctypes.c_double(), ctypes.c_double()
            lib.golden_map_run(P(keys.ctypes.data), P(vals.ctypes.data), ctypes.c_uint64(n_ins),
                               P(q.ctypes.data), ctypes.c_uint64(len(q)), ctypes.c_int64(-1), P(out.ctypes.data),
                               ctypes.byref(ins_s), ctypes.byref(find_s))
            return ins_s.value + find_s.value, bool(np.array_equal(out, np.where(q < (1 << 62), q * 7 + 1, -1)))

        sec, ok, all_s = median_run(once)
        return {"value": (n_ins + len(q)) / sec / 1e9, "unit": "Gkeys/s", "cores": 1, "median_s": round(sec, 4),
                "runs_s": all_s, "verified": ok,
                "sample": "std::unordered_map Insert of %d keys + Find of %d queries (50%% hit)" % (n_ins, len(q))}

    def bloom_leg():
        nb, n = 1 << 22, 1 << 26
        rng = np.random.default_rng(21)
        keys = rng.integers(-(1 << 63), (1 << 63) - 1, n, dtype=np.int64)
        q = np.concatenate([keys[: n // 2], rng.integers(-(1 << 63), (1 << 63) - 1, n - n // 2, dtype=np.int64)])
        words = np.zeros(nb * 8, np.uint32)

        def add():
            words[:] = 0
            t0 = time.perf_counter()
            O.bloom_add_mt(words, nb, keys, threads)
            return time.perf_counter() - t0, True

        def contains():
            t0 = time.perf_counter()
            out = O.bloom_contains_mt(words, nb, q, threads)
            return time.perf_counter() - t0, bool(out[: n // 2].all())

        a_s, a_ok, a_all = median_run(add)
        c_s, c_ok, c_all = median_run(contains)
        return {"add_gkeys_s": round(n / a_s / 1e9, 4), "contains_gkeys_s": round(n / c_s / 1e9, 4),
                "add_median_s": round(a_s, 4), "contains_median_s": round(c_s, 4), "add_runs_s": a_all,
                "contains_runs_s": c_all, "cores": threads, "verified": a_ok and c_ok,
                "sample": "BloomFilter of 2^22 blocks: Add of 2^26 keys, Contains of 2^26 (half absent)"}

    def hll_leg():
        n, p = 1 << 27, 15
        keys = np.random.default_rng(22).integers(0, 10 ** 8, n, dtype=np.int64)

        def once():
            regs = np.zeros(1 << p, np.uint8)
            t0 = time.perf_counter()
            O.hll_add_mt(regs, p, keys, threads)
            return time.perf_counter() - t0, bool(regs.any())

        sec, ok, all_s = median_run(once)
        return {"add_gkeys_s": round(n / sec / 1e9, 4), "median_s": round(sec, 4), "runs_s": all_s,
                "cores": threads, "verified": ok, "sample": "HyperLogLog p=15 Add of 2^27 int64 keys"}

    mt = map_leg(25, threads)
    st = map_leg(21, 1)
    gm = golden_leg(22)
    bl = bloom_leg()
    hl = hll_leg()
    return {"value": mt["value"], "unit": "Gkeys/s", "cores": threads, "kind": "port",
            "sample": "oracle restatement (pthreads, __atomic 64-bit CAS) of StaticMap int64 " + mt["sample"]
                      + ", median of %d runs" % runs,
            "median_s": mt["median_s"], "runs_s": mt["runs_s"], "verified": mt["verified"],
            "single_thread": st, "golden_unordered_map": gm, "bloom": bl, "hll": hl,
            "threads_from": "OMP_NUM_THREADS" if (not args.cpu_threads and share and share.isdigit())
                            else ("--cpu-threads" if args.cpu_threads else "sched_getaffinity"),
            "cpu": _cpu_description()}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: one GPU only; --gpus %d is not supported (the tables shard by key hash across GPUs, not built here)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.blocks_per_cu:
        os.environ["MXCO_OA_BLOCKS_PER_CU"] = str(args.blocks_per_cu)
    if args.inflight:
        os.environ["MXCO_OA_INFLIGHT"] = str(args.inflight)
    if args.probe_flags >= 0:
        os.environ["MXCO_OA_PROBE_FLAGS"] = str(args.probe_flags)
    if args.store_mode >= 0:
        os.environ["MXCO_OA_STORE_MODE"] = str(args.store_mode)
    import torch
    if args.copy_calibration:
        copy_calibration()
        return
    import mxco
    log("bench: StaticMap int64 LF %.2f, 2^%d slots requested, group %d" % (args.lf, args.slots_log2, args.group))
    r = run_gpu(args, torch, mxco, args.lf, args.group, args.steps, args.warmup)
    log("bench: gpu done %s" % json.dumps(r))
    n_keys = r["n_insert"] + r["n_find"]
    ms = r["wall_ms_per_step"]
    value = n_keys / (ms * 1e-3) / 1e9
    # Algorithmic bytes (128 B per distinct line the reference's walk reads,
    # plus the streamed key/value in and out).
    find_bytes = 128.0 * r["find_lines"] + r["n_find"] * (8 + 8)
    ins_bytes = 128.0 * r["insert_lines"] + r["n_insert"] * (128 + 16)
    find_gbs = find_bytes / (r["find_ms"] * 1e-3) / 1e9
    ins_gbs = ins_bytes / (r["insert_ms"] * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "oa_traffic_i64_lf08.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf))
    dominant_insert = r["insert_ms"] >= r["find_ms"]
    g = args.group or 4
    find_kernel = "oa_wave_staged_kernel<8,8,%d,OP_FIND>" % g
    roof_find = {"kernel": find_kernel, "bound": "hbm", "achieved": round(find_gbs, 1),
                 "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(find_gbs / PEAK_HBM_GBS, 4),
                 "traffic": traffic.get("find") if traffic else None,
                 "lines_per_query": r["find_lines"] / r["n_find"], "launch_ms": r["find_ms"],
                 "algorithmic_bytes": find_bytes}
    if r["insert_path"] == "region build":
        # the bytes the build's passes move: scatter 32, regroup 32 and apply
        # 16 per pair, plus each region written once (the step's Clear leaves
        # the table known empty, so regions are not read)
        build_bytes = 80.0 * r["n_insert"] + 16.0 * r["capacity"]
        build_gbs = build_bytes / (r["insert_ms"] * 1e-3) / 1e9
        roof_ins = {"kernel": "build_bounded_{scatter,regroup}_kernel, build_{apply,overflow,spill}_kernel<8,8> "
                              "(region build)",
                    "bound": "hbm", "achieved": round(build_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(build_gbs / PEAK_HBM_GBS, 4),
                    "traffic": traffic.get("insert") if traffic else None, "launch_ms": r["insert_ms"],
                    "algorithmic_bytes": build_bytes,
                    "cas_walk_bytes": ins_bytes, "cas_walk_equivalent_frac": round(ins_gbs / PEAK_HBM_GBS, 4),
                    "cas_walk_ms": r.get("insert_cas_walk_ms")}
    else:
        roof_ins = {"kernel": "oa_kernel<8,8,%d,OP_INSERT>" % g, "bound": "hbm", "achieved": round(ins_gbs, 1),
                    "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(ins_gbs / PEAK_HBM_GBS, 4),
                    "traffic": traffic.get("insert") if traffic else None,
                    "lines_per_key": r["insert_lines"] / r["n_insert"], "launch_ms": r["insert_ms"],
                    "algorithmic_bytes": ins_bytes}
    line = {
        "metric": "Gkeys/s StaticMap Insert/Find int64 @LF0.8; batched vs plain speedup; p99 ms",
        "value": round(value, 4), "unit": "Gkeys/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "synthetic (uniform 64-bit keys, seeded torch generator)",
        "config": {"workload": "StaticMap int64->int64, %d slots (2^%d requested, bucket 5, linear probing, "
                               "murmurhash3_32), LF %.2f: Clear + Insert %d + Find %d (50%% hit) per step"
                               % (r["capacity"], args.slots_log2, args.lf, r["n_insert"], r["n_find"]),
                   "parallelism": "single GPU"},
        "insert_gkeys_s": round(r["n_insert"] / (r["insert_ms"] * 1e-3) / 1e9, 4),
        "find_gkeys_s": round(r["n_find"] / (r["find_ms"] * 1e-3) / 1e9, 4),
        "clear_ms": round(r["clear_ms"], 4),
        "insert_path": r["insert_path"],
        "cas_walk_atomics": r.get("cas_walk_atomics"),
        "verified": r["verified"] and r["insert_failed"] == 0,
        "roofline": roof_ins if dominant_insert else roof_find,
        "roofline_other": roof_find if dominant_insert else roof_ins,
        "group": r["group"],
    }
    if args.sweep:
        # Load factors of the measured configuration at the default group
        # width, with their algorithmic bytes; then group widths at args.lf.
        sw = []
        for lf in (0.5, 0.7, 0.8, 0.9):
            x = run_gpu(args, torch, mxco, lf, args.group, 3, 1, verify=True)
            lq = x["find_lines"] / x["n_find"]
            lk = x["insert_lines"] / x["n_insert"]
            f_gbs = x["n_find"] * (128 * lq + 16) / (x["find_ms"] * 1e-3) / 1e9
            i_gbs = x["n_insert"] * (128 * lk + 128 + 16) / (x["insert_ms"] * 1e-3) / 1e9
            sw.append({"lf": lf, "group": args.group, "insert_ms": round(x["insert_ms"], 3),
                       "find_ms": round(x["find_ms"], 3),
                       "insert_gkeys_s": round(x["n_insert"] / x["insert_ms"] / 1e6, 3),
                       "find_gkeys_s": round(x["n_find"] / x["find_ms"] / 1e6, 3),
                       "lines_per_query": round(lq, 4), "lines_per_key": round(lk, 4),
                       "find_frac": round(f_gbs / PEAK_HBM_GBS, 4), "insert_frac": round(i_gbs / PEAK_HBM_GBS, 4),
                       "verified": x["verified"]})
            log("sweep %s" % json.dumps(sw[-1]))
        for g in (1, 2, 4, 8):
            x = run_gpu(args, torch, mxco, args.lf, g, 3, 1, verify=False)
            sw.append({"lf": args.lf, "group": g, "insert_ms": round(x["insert_ms"], 3),
                       "find_ms": round(x["find_ms"], 3),
                       "insert_gkeys_s": round(x["n_insert"] / x["insert_ms"] / 1e6, 3),
                       "find_gkeys_s": round(x["n_find"] / x["find_ms"] / 1e6, 3)})
            log("sweep %s" % json.dumps(sw[-1]))
        line["sweep"] = sw
    if not args.no_streams:
        log("bench: zipf insert stream")
        line["zipf_insert"] = run_zipf_insert(args, torch, mxco)
        log("bench: zipf %s" % json.dumps(line["zipf_insert"]))
    line["insert_or_assign"] = r["insert_or_assign"]
    if not args.no_filters:
        log("bench: bloom filter")
        line["bloom"] = run_bloom(torch, mxco)
        log("bench: bloom %s" % json.dumps(line["bloom"]))
        log("bench: hyperloglog")
        line["hll"] = run_hll(torch, mxco, sketches=args.hll_sketches)
        log("bench: hll %s" % json.dumps(line["hll"]))
    log("bench: ceilings")
    line["ceilings"] = run_ceilings(torch, mxco)
    log("bench: ceilings %s" % json.dumps(line["ceilings"]))
    if not args.no_roaring:
        log("bench: roaring bitmap")
        line["roaring"] = run_roaring(torch, mxco)
        log("bench: roaring %s" % json.dumps(line["roaring"]))
    if not args.no_containers:
        log("bench: multimap, multiset, dynamic map, erase/count")
        line["containers"] = run_containers(torch, mxco)
        log("bench: containers %s" % json.dumps(line["containers"]))
    if not args.no_batched:
        log("bench: batched 1024 tables")
        line["batched"] = run_batched(args, torch, mxco)
        log("bench: batched %s" % json.dumps(line["batched"]))
    if not args.no_serving:
        log("bench: serving")
        line["serving"] = run_serving(args, torch, mxco)
        log("bench: serving %s" % json.dumps(line["serving"]))
    if not args.no_cpu:
        log("bench: cpu baseline")
        line["cpu_baseline"] = cpu_baseline(args)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
