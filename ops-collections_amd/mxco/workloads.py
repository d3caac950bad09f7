"""Synthetic, seeded workloads of the measured configurations (device tensors).

Keys of one table are distinct by construction: key = ((j * 2654435761) mod
2^31) xor salt, a bijection of the index j on [0, 2^31), so no key equals the
sentinel -1 and keys with j >= the table's insert count are guaranteed absent.
"""
import math

import numpy as np
import torch

MULT = 2654435761


def _mix(j, salt):
    return ((j * MULT) & 0x7FFFFFFF) ^ salt


def uniform_map(n_ins, n_find, hit=0.5, seed=1234, device="cuda"):
    """Distinct keys drawn uniformly from the whole signed 64-bit range with
    the sentinel -1 excluded, in random order, and n_find queries of which
    round(hit * n_find) are drawn from those keys and the rest are uniform
    64-bit keys guaranteed absent (rejected against the sorted key set),
    shuffled.

    Returns (keys, queries, is_hit) as device tensors."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    lo, hi = -(1 << 63), (1 << 63) - 1

    def draw(count):
        c = torch.randint(lo, hi, (count,), device=device, dtype=torch.int64, generator=g)
        return c[c != -1]

    cand = torch.unique(draw(n_ins + n_ins // 1000 + 64))  # sorted; a collision is rare
    if cand.numel() < n_ins:
        raise RuntimeError("too few distinct keys drawn")
    keys = cand[torch.randperm(cand.numel(), device=device, generator=g)[:n_ins]]
    del cand
    srt = torch.sort(keys).values
    n_hit = int(round(hit * n_find))
    hits = keys[torch.randint(0, n_ins, (n_hit,), device=device, generator=g)]
    n_miss = n_find - n_hit
    miss = draw(n_miss + n_miss // 100 + 64)
    pos = torch.searchsorted(srt, miss).clamp_(max=n_ins - 1)
    miss = miss[srt[pos] != miss][:n_miss]
    del srt, pos
    if miss.numel() < n_miss:
        raise RuntimeError("too few absent keys drawn")
    q = torch.cat([hits, miss])
    is_hit = torch.cat([torch.ones(n_hit, dtype=torch.bool, device=device),
                        torch.zeros(n_miss, dtype=torch.bool, device=device)])
    perm = torch.randperm(n_find, device=device, generator=g)
    return keys, q[perm], is_hit[perm]


def batched_tables(num_tables=1024, cap_log2=(10, 22), lf=(0.5, 0.85), find_log2=(6, 20), hit=0.5, seed=4,
                   valid_extent=None):
    """Capacities log-uniform in 2^cap_log2, per-table load factors uniform in
    `lf`, round(load*capacity) distinct inserts per table (value = key*7+1 in
    uint32), find batches log-uniform in 2^find_log2 at `hit` hit rate.

    Returns a dict of host lists (caps, n_ins, n_find, offsets) and device
    tensors (pairs int32 (N,2), queries int32 (Q,))."""
    rng = np.random.default_rng(seed)
    req = np.floor(2.0 ** rng.uniform(cap_log2[0], cap_log2[1], num_tables)).astype(np.int64)
    caps = [valid_extent(int(c)) if valid_extent else int(c) for c in req]
    lfs = rng.uniform(lf[0], lf[1], num_tables)
    n_ins = [int(round(l * c)) for l, c in zip(lfs, caps)]
    n_find = [int(2.0 ** rng.uniform(find_log2[0], find_log2[1])) for _ in range(num_tables)]
    salts = rng.integers(0, 2**31, num_tables)
    dev = "cuda"
    ni = torch.tensor(n_ins, dtype=torch.int64, device=dev)
    nf = torch.tensor(n_find, dtype=torch.int64, device=dev)
    salt = torch.tensor(salts, dtype=torch.int64, device=dev)
    ins_off = torch.cumsum(ni, 0) - ni
    tid = torch.repeat_interleave(torch.arange(num_tables, device=dev), ni)
    j = torch.arange(int(ni.sum().item()), device=dev) - ins_off[tid]
    keys = _mix(j, salt[tid])
    vals = (keys * 7 + 1) & 0xFFFFFFFF
    vals = torch.where(vals >= 2**31, vals - 2**32, vals)
    pairs = torch.stack([keys, vals], 1).to(torch.int32).contiguous()
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    qt = torch.repeat_interleave(torch.arange(num_tables, device=dev), nf)
    u = torch.rand(qt.numel(), device=dev, generator=g, dtype=torch.float64)
    is_hit = torch.rand(qt.numel(), device=dev, generator=g) < hit
    jq = torch.where(is_hit, (u * ni[qt]).long().clamp(max=(ni[qt] - 1).clamp(min=0)),
                     ni[qt] + (u * 2**30).long())
    queries = _mix(jq, salt[qt]).to(torch.int32).contiguous()
    ins_off_l = [0] + list(np.cumsum(n_ins)[:-1])
    find_off_l = [0] + list(np.cumsum(n_find)[:-1])
    return {"caps": caps, "n_ins": n_ins, "n_find": n_find, "ins_off": [int(x) for x in ins_off_l],
            "find_off": [int(x) for x in find_off_l], "pairs": pairs, "queries": queries,
            "expected": torch.where(is_hit & (ni[qt] > 0), queries * 7 + 1, torch.full_like(queries, -1))}


def serving(num_maps=48, num_multimaps=16, slots_log2=(20, 26), fill=0.5, num_requests=2000, qps=20000.0,
            batch_log2=(0, 16), hit=0.7, seed=5):
    """Resident tables and an open-loop request stream: Poisson arrivals at
    `qps`; 80 % Find/Contains, 15 % Insert on the StaticMaps, 5 % Retrieve on
    the multimaps; batch sizes log-uniform 1..2^16 keys at `hit` hit rate.

    Keys of table t are the bijection of (t, j) used by batched_tables; map
    tables start `fill` full, multimap keys repeat 1..8 times (value = key
    * 16 + copy).  Queries draw present keys from the initial fill and
    absent keys from indices no insert reaches, so a query's hit flag
    (request["hit"]) holds whatever the order of the inserts.  Returns host
    plans and device buffers for every request."""
    rng = np.random.default_rng(seed)
    dev = "cuda"
    T = num_maps + num_multimaps
    caps = [int(2.0 ** rng.uniform(*slots_log2)) for _ in range(T)]
    salts = rng.integers(0, 2**31, T)

    def keys_of(t, j):
        return (((j * MULT) & 0x7FFFFFFF) ^ int(salts[t])) + (int(t) << 32)

    init = []  # per table: device pairs to pre-insert
    n_present = []
    for t in range(T):
        if t < num_maps:
            n = int(fill * caps[t])
            j = torch.arange(n, device=dev, dtype=torch.int64)
            k = keys_of(t, j)
            init.append(torch.stack([k, k * 7 + 1], 1).contiguous())
            n_present.append(n)
        else:
            n_keys = max(1, int(fill * caps[t] / 4.5))
            rep = torch.from_numpy(rng.integers(1, 9, n_keys)).to(dev)
            j = torch.repeat_interleave(torch.arange(n_keys, device=dev, dtype=torch.int64), rep)
            k = keys_of(t, j)
            copy = torch.arange(j.numel(), device=dev) - torch.repeat_interleave(torch.cumsum(rep, 0) - rep, rep)
            init.append(torch.stack([k, k * 16 + copy], 1).contiguous())
            n_present.append(n_keys)
    gaps = rng.exponential(1e6 / qps, num_requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    reqs = []
    next_new = list(n_present)
    for r in range(num_requests):
        u = rng.random()
        n = int(2.0 ** rng.uniform(*batch_log2))
        if u < 0.05:
            t = int(rng.integers(num_maps, T))
            op = "retrieve"
        else:
            t = int(rng.integers(0, num_maps))
            op = "find" if u < 0.45 else ("contains" if u < 0.85 else "insert")
        if op == "insert":
            j = torch.arange(next_new[t], next_new[t] + n, device=dev, dtype=torch.int64)
            next_new[t] += n
            k = keys_of(t, j)
            inp = torch.stack([k, k * 7 + 1], 1).contiguous()
        else:
            hitm = torch.from_numpy(rng.random(n) < hit).to(dev)
            jh = torch.from_numpy(rng.integers(0, max(1, n_present[t]), n)).to(dev)
            jm = torch.from_numpy(rng.integers(2**30, 2**31 - 1, n)).to(dev)
            inp = keys_of(t, torch.where(hitm, jh, jm)).contiguous()
        reqs.append({"table": t, "op": op, "input": inp, "n": n, "hit": None if op == "insert" else hitm})
    return {"caps": caps, "num_maps": num_maps, "init": init, "requests": reqs, "arrivals_us": arrivals.tolist()}


# Odd 64-bit multiplier (golden ratio): rank -> rank * ZIPF_MULT mod 2^64 is a
# bijection, so distinct ranks give distinct keys, and the ranks drawn here
# (< 2^40) never map to the sentinel -1 (that needs rank = -1/ZIPF_MULT mod 2^64).
ZIPF_MULT = 0x9E3779B97F4A7C15 - (1 << 64)


class Zipf:
    """Exact bounded Zipf(s) over ranks 1..n by inverse CDF: the cumulative
    weights k^-s (FP64, on the device) are searched with uniform FP64 draws.

    draw(count, seed) returns rank-1 in [0, n) as int64; keys(...) maps ranks to
    64-bit keys by the bijection above."""

    def __init__(self, n, s, device="cuda"):
        w = torch.arange(1, n + 1, dtype=torch.float64, device=device).pow_(-s)
        self.cdf = torch.cumsum(w, 0)
        del w
        self.cdf /= self.cdf[-1].clone()
        self.n = n
        self.device = device

    def draw(self, count, seed):
        g = torch.Generator(device=self.device)
        g.manual_seed(seed)
        u = torch.rand(count, dtype=torch.float64, device=self.device, generator=g)
        r = torch.searchsorted(self.cdf, u)
        del u
        return r.clamp_(max=self.n - 1)

    @staticmethod
    def keys(ranks):
        return ranks * ZIPF_MULT


def roaring_portable32(keys, runs=False):
    """Roaring portable serialization (RoaringFormatSpec) of a set of uint32
    keys, vectorised for large key sets (bench inputs).  runs=False writes
    array (<= 4096 values) and bitset containers with cookie 12346; with
    runs="optimize" a container is written as runs when 2 + 4 * runs is
    below its array (2 + 2 * cardinality) or bitset (8192) size."""
    k = np.unique(np.asarray(keys, dtype=np.uint64) & 0xFFFFFFFF).astype(np.uint32)
    hi, lo = (k >> 16).astype(np.int64), (k & 0xFFFF).astype(np.int64)
    hks, first, cards = np.unique(hi, return_index=True, return_counts=True)
    n = len(hks)
    brk = np.ones(len(k), dtype=bool)  # first value of each run
    if len(k):
        brk[1:] = (hi[1:] != hi[:-1]) | (lo[1:] != lo[:-1] + 1)
    run_counts = np.add.reduceat(brk.astype(np.int64), first) if n else np.zeros(0, np.int64)
    if runs == "optimize":
        as_run = np.where(cards <= 4096, 4 * run_counts < 2 * cards, 2 + 4 * run_counts < 8192)
    else:
        as_run = np.zeros(n, dtype=bool)
    has_run = bool(as_run.any())
    head = bytearray()
    if has_run:
        head += np.uint32(12347 | ((n - 1) << 16)).tobytes()
        head += np.packbits(as_run, bitorder="little").tobytes()
    else:
        head += np.array([12346, n], dtype="<u4").tobytes()
    head += np.stack([hks, cards - 1], 1).astype("<u2").tobytes()
    offsets = (not has_run) or n >= 4
    bodies = []
    for i in range(n):
        v = lo[first[i]:first[i] + cards[i]]
        if as_run[i]:
            idx = np.flatnonzero(brk[first[i]:first[i] + cards[i]])
            starts, ends = v[idx], v[np.append(idx[1:] - 1, len(v) - 1)]
            bodies.append(np.uint16(len(starts)).tobytes() + np.stack([starts, ends - starts], 1).astype("<u2").tobytes())
        elif cards[i] <= 4096:
            bodies.append(v.astype("<u2").tobytes())
        else:
            bits = np.zeros(65536, dtype=bool)
            bits[v] = True
            bodies.append(np.packbits(bits, bitorder="little").tobytes())
    sizes = np.array([len(b) for b in bodies], dtype=np.int64)
    if offsets:
        base = len(head) + 4 * n
        head += (base + np.concatenate([[0], np.cumsum(sizes)[:-1]]) if n else np.zeros(0)).astype("<u4").tobytes()
    return bytes(head) + b"".join(bodies)


def roaring_portable64(keys, runs=False):
    """64-bit portable format: uint64 bucket count, then per distinct high
    32-bit half its key and the 32-bit bitmap of the low halves."""
    k = np.unique(np.asarray(keys, dtype=np.uint64))
    his = np.unique(k >> np.uint64(32))
    out = bytearray(np.uint64(len(his)).tobytes())
    for h in his:
        out += np.uint32(h).tobytes() + roaring_portable32(k[(k >> np.uint64(32)) == h], runs)
    return bytes(out)
