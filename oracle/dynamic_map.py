"""DynamicMap restatement (TEST INFRASTRUCTURE ONLY: used by tests/ as the
checker, never by the product).

Follows include/dynamic_map.h: the growth rules (Usable :366-370,
GrowTarget :512-521, StartPrepareNext :470-486, ActivateNext :488-510,
Reserve :142-165), the size accounting of every call (Insert :168-219,
InsertAsync :221-235, InsertOrAssign :237-259, Erase :318-335, Clear
:349-357) and map semantics (a key lives in at most one submap; Insert is
first-wins, InsertOrAssign last-wins).  Capacities of the StaticMap submaps
come from the oracle's MakeValidExtent (pinned in tests/test_oracle.py).
"""
import math

import numpy as np

import oracle as O

MIN_INSERT = 10000
MAX_LOAD = float(np.float32(0.6))  # the reference's float 0.60f, promoted to double
GROWTH = 2
SUB_BATCH = 800000


def usable(cap):
    u = math.floor(MAX_LOAD * cap)
    return u - MIN_INSERT if u > MIN_INSERT else 0


class DynamicMapOracle:
    def __init__(self, initial_capacity, empty_key, erased_key, bucket=5, probing=O.PROBE_LINEAR):
        if empty_key == erased_key:
            raise ValueError("DynamicMap: emptyKey and erasedKey must differ")
        self.bucket, self.probing = bucket, probing
        self.empty, self.erased = empty_key, erased_key
        self.caps, self.sub_size = [], []
        self.total_cap = 0
        self._push(initial_capacity)
        self.next_cap = self.caps[0] * GROWTH
        self.prep = None  # requested capacity of the prepared submap
        self.size = 0
        self.vals = {}
        self.where = {}

    def _extent(self, c):
        return O.valid_extent(c, self.bucket, self.probing)

    def _push(self, requested):
        cap = self._extent(requested)
        self.caps.append(cap)
        self.sub_size.append(0)
        self.total_cap += cap

    def num_submaps(self):
        return len(self.caps)

    def _active_room(self):
        a = len(self.caps) - 1
        u = usable(self.caps[a])
        return u - self.sub_size[a] if u > self.sub_size[a] else 0

    def _grow_target(self, m):
        return max(self.size * GROWTH, m)

    def _prepare(self, m):
        if self.prep is not None:
            return
        cap = self.next_cap
        while usable(cap) < self._grow_target(m):
            cap *= GROWTH
        self.prep = cap

    def _activate(self, m):
        if self.prep is not None and usable(self.prep) >= m:
            self._push(self.prep)
            self.next_cap = self.prep * GROWTH
            self.prep = None
            return
        self.prep = None
        cap = self.next_cap
        while usable(cap) < m:
            cap *= GROWTH
        self._push(cap)
        self.next_cap = cap * GROWTH

    def _ensure(self, n):
        if self._active_room() < n:
            self._activate(self._grow_target(n))

    def _nonempty_before(self, a):
        return any(self.sub_size[i] for i in range(a))

    def reserve(self, n):
        remaining, idx = n, 0
        while remaining > 0:
            if idx < len(self.caps):
                cap = self.caps[idx]
            else:
                cap = self.next_cap
                self._push(cap)
                self.next_cap = cap * GROWTH
            u = usable(cap)
            if u >= remaining:
                break
            remaining -= u
            idx += 1

    def _put(self, keys, vals, a, assign):
        """Sequential insert of the chunk into submap a; returns new keys."""
        new = 0
        for k, v in zip(keys, vals):
            if k == self.empty:
                continue
            if k in self.vals:
                if assign:
                    self.vals[k] = v
                continue
            self.vals[k] = v
            self.where[k] = a
            new += 1
        return new

    def insert(self, keys, vals, mode="exact", sync=True):
        keys, vals = list(keys), list(vals)
        n = len(keys)
        if n == 0:
            return 0
        self._ensure(n)
        a = len(self.caps) - 1
        if not sync or mode == "append":
            if not sync and self._active_room() < n * 2:
                self._prepare(n)
            self._put(keys, vals, a, False)
            self.sub_size[a] += n
            self.size += n
            return n
        total, off = 0, 0
        first = not self._nonempty_before(a)
        while off < n:
            if not first:
                room = self._active_room()
                if room == 0:
                    self._activate(self._grow_target(n - off))
                    continue
                if room < (n - off) * 2:
                    self._prepare(n - off)
                chunk = min(n - off, room)
            else:
                chunk = n
            a = len(self.caps) - 1
            ins = self._put(keys[off:off + chunk], vals[off:off + chunk], a, False)
            self.sub_size[a] += ins
            self.size += ins
            total += ins
            off += chunk
        return total

    def insert_or_assign(self, keys, vals):
        keys, vals = list(keys), list(vals)
        n = len(keys)
        if n == 0:
            return 0
        self._ensure(n)
        total, off = 0, 0
        while off < n:
            room = self._active_room()
            if room == 0:
                self._activate(self._grow_target(n - off))
                continue
            if room < (n - off) * 2:
                self._prepare(n - off)
            chunk = min(n - off, room)
            a = len(self.caps) - 1
            ins = self._put(keys[off:off + chunk], vals[off:off + chunk], a, True)
            self.sub_size[a] += ins
            self.size += ins
            total += ins
            off += chunk
        return total

    def find(self, keys, empty_value):
        return [self.vals.get(k, empty_value) for k in keys]

    def contains(self, keys):
        return [k in self.vals for k in keys]

    def erase(self, keys):
        gone = {}
        for k in keys:
            if k in self.vals and k not in gone:
                gone[k] = self.where[k]
        per = [0] * len(self.caps)
        for k, s in gone.items():
            del self.vals[k]
            del self.where[k]
            per[s] += 1
        total = 0
        for s in range(len(self.caps)):
            if self.sub_size[s] == 0:
                continue
            self.sub_size[s] -= min(per[s], self.sub_size[s])
            total += per[s]
        self.size -= min(total, self.size)
        return total

    def clear(self):
        self.prep = None
        self.sub_size = [0] * len(self.caps)
        self.size = 0
        self.vals, self.where = {}, {}
