"""Roaring portable serialization (TEST INFRASTRUCTURE ONLY: used by tests/
to build bitmaps and as the membership checker, never by the product).

Writes a key set in the RoaringFormatSpec layout the reference reads
(include/detail/roaring_bitmap/format.h): 32-bit bitmaps with cookie 12346
(no runs, explicit offsets) or 12347 (run bitmap; offsets only with 4 or more
containers), array containers up to 4096 values, bitset containers above,
run containers where asked; 64-bit bitmaps as a uint64 bucket count followed
by (uint32 high key, 32-bit bitmap) per bucket.  Membership is checked
against the key set itself (brute force), not against this writer.
"""
import struct

import numpy as np

COOKIE_NO_RUN = 12346
COOKIE_RUN = 12347


def _runs_of(values):
    runs = []
    for v in values:
        if runs and v == runs[-1][0] + runs[-1][1] + 1:
            runs[-1][1] += 1
        else:
            runs.append([v, 0])
    return runs


def serialize32(keys, run_containers=False, force_offsets=None):
    """keys: iterable of uint32.  run_containers: True (every container as
    runs), False (no runs), a set of container keys to write as runs, or
    "optimize": a container becomes runs when that is strictly smaller than
    its cardinality-chosen form, 2 + 4 * runs < 2 + 2 * cardinality for an
    array and < 8192 for a bitset (the rule of CRoaring's run_optimize, whose
    output the reference's RoaringFormatSpec test files are)."""
    keys = np.unique(np.asarray(list(keys), dtype=np.uint64)).astype(np.uint64)
    groups = {}
    for k in keys.tolist():
        groups.setdefault(k >> 16, []).append(k & 0xFFFF)
    hks = sorted(groups)
    def optimized(vals):
        r = len(_runs_of(vals))
        return 4 * r < 2 * len(vals) if len(vals) <= 4096 else 2 + 4 * r < 8192

    as_run = [(run_containers is True) or (isinstance(run_containers, (set, frozenset)) and h in run_containers)
              or (run_containers == "optimize" and optimized(groups[h])) for h in hks]
    has_run = any(as_run)
    out = bytearray()
    n = len(hks)
    if has_run:
        out += struct.pack("<I", COOKIE_RUN | ((n - 1) << 16))
        bits = bytearray((n + 7) // 8)
        for i, r in enumerate(as_run):
            if r:
                bits[i // 8] |= 1 << (i % 8)
        out += bits
    else:
        out += struct.pack("<II", COOKIE_NO_RUN, n)
    for h in hks:
        out += struct.pack("<HH", h, len(groups[h]) - 1)
    offsets = (not has_run) or n >= 4
    if force_offsets is not None:
        offsets = force_offsets
    table_at = len(out)
    if offsets:
        out += bytes(4 * n)
    bodies = []
    for h, r in zip(hks, as_run):
        vals = groups[h]
        if r:
            runs = _runs_of(vals)
            body = struct.pack("<H", len(runs)) + b"".join(struct.pack("<HH", s, l) for s, l in runs)
        elif len(vals) <= 4096:
            body = b"".join(struct.pack("<H", v) for v in vals)
        else:
            bs = bytearray(8192)
            for v in vals:
                bs[v // 8] |= 1 << (v % 8)
            body = bytes(bs)
        bodies.append(body)
    for i, body in enumerate(bodies):
        if offsets:
            struct.pack_into("<I", out, table_at + 4 * i, len(out))
        out += body
    return bytes(out)


def serialize64(keys, run_containers=False):
    keys = np.unique(np.asarray(list(keys), dtype=np.uint64))
    buckets = {}
    for k in keys.tolist():
        buckets.setdefault(k >> 32, []).append(k & 0xFFFFFFFF)
    out = bytearray(struct.pack("<Q", len(buckets)))
    for hk in sorted(buckets):
        out += struct.pack("<I", hk) + serialize32(buckets[hk], run_containers)
    return bytes(out)
