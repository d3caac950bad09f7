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
