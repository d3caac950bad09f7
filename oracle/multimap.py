"""Numpy restatement of the StaticMultimap contract (TEST INFRASTRUCTURE ONLY).

Follows include/static_multimap.h:323-554 and
docs/static_multimap_API文档和使用示例.md:21-45:
  Insert / InsertIf: every selected pair is checked for the reserved key or
    value first (ValueError, nothing inserted); then the input-order prefix
    of the selected pairs that fits the remaining capacity is admitted and
    the number of selected-but-not-admitted pairs is returned.
  Find: smallest value of the key, else the empty value.  Contains: bool.
  Count: total matches, a repeated query counting again.
  Retrieve: matches grouped by query index, values ascending per query.
  RetrieveAll: the stored multiset.
Parity pins: tests/test_oracle.py checks it against a dict-of-lists brute
force on small inputs.
"""
import numpy as np


class MultimapOracle:
    def __init__(self, capacity, empty_key, empty_value, key_dtype=np.int64, value_dtype=np.int64):
        if capacity <= 0:
            raise ValueError("capacity must be positive")
        self.capacity = capacity
        self.ek, self.ev = empty_key, empty_value
        self.kd, self.vd = np.dtype(key_dtype), np.dtype(value_dtype)
        self.keys = np.zeros(0, self.kd)
        self.values = np.zeros(0, self.vd)
        self._index = None

    def size(self):
        return len(self.keys)

    def clear(self):
        self.keys = np.zeros(0, self.kd)
        self.values = np.zeros(0, self.vd)
        self._index = None

    def insert(self, keys, values, selected=None):
        keys = np.asarray(keys, self.kd)
        values = np.asarray(values, self.vd)
        sel = np.ones(len(keys), bool) if selected is None else np.asarray(selected, bool)
        if np.any((keys[sel] == self.ek) | (values[sel] == self.ev)):
            raise ValueError("selected input contains a reserved key or value")
        idx = np.nonzero(sel)[0]
        room = self.capacity - self.size()
        take = idx[:room]
        self.keys = np.concatenate([self.keys, keys[take]])
        self.values = np.concatenate([self.values, values[take]])
        self._index = None
        return len(idx) - len(take)

    def _build(self):
        if self._index is None:
            order = np.lexsort((self.values, self.keys))  # by key, then value
            k, v = self.keys[order], self.values[order]
            uk, start, cnt = np.unique(k, return_index=True, return_counts=True)
            self._index = (uk, start, cnt, v)
        return self._index

    def _lookup(self, q):
        uk, start, cnt, v = self._build()
        q = np.asarray(q, self.kd)
        pos = np.searchsorted(uk, q)
        pos_c = np.minimum(pos, max(len(uk) - 1, 0))
        hit = (len(uk) > 0) & (uk[pos_c] == q) if len(uk) else np.zeros(len(q), bool)
        hit &= q != self.ek
        return hit, pos_c, start, cnt, v

    def contains(self, q):
        return self._lookup(q)[0]

    def find(self, q):
        hit, pos, start, cnt, v = self._lookup(q)
        out = np.full(len(hit), self.ev, self.vd)
        if len(v):
            out[hit] = v[start[pos[hit]]]
        return out

    def counts(self, q):
        hit, pos, start, cnt, v = self._lookup(q)
        c = np.zeros(len(hit), np.int64)
        if len(v):
            c[hit] = cnt[pos[hit]]
        return c

    def count(self, q):
        return int(self.counts(q).sum())

    def retrieve(self, q):
        q = np.asarray(q, self.kd)
        hit, pos, start, cnt, v = self._lookup(q)
        c = self.counts(q)
        total = int(c.sum())
        probe = np.repeat(q, c)
        vals = np.zeros(total, self.vd)
        off = np.concatenate([[0], np.cumsum(c)])
        for i in np.nonzero(c)[0]:
            vals[off[i]:off[i + 1]] = v[start[pos[i]]:start[pos[i]] + c[i]]
        return probe, vals

    def retrieve_all(self):
        order = np.lexsort((self.values, self.keys))
        return self.keys[order], self.values[order]
