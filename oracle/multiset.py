"""StaticMultiset restatement (TEST INFRASTRUCTURE ONLY: used by tests/ as the
checker, never by the product).

Follows the reference's contract (include/static_multiset.h:27-34, 159-345):
Insert keeps, in input order, the earliest selected non-sentinel keys that
fit the remaining capacity and returns the selected keys not kept plus the
selected sentinel keys; RetrieveAll returns the kept keys in insertion order;
CountEach / CountEachOuter / Count / Retrieve read exact multiplicities,
Retrieve in query order.  Pinned by the reference's own test vectors
(tests/static_multiset/semantics_test.cpp:87-129, insert_test.cpp:60-74) in
tests/test_oracle.py.
"""
import collections

import numpy as np


class MultisetOracle:
    def __init__(self, capacity, empty_key):
        if capacity <= 0:
            raise ValueError("StaticMultiset: invalid capacity")
        self.capacity = capacity
        self.empty = empty_key
        self.ordered = []
        self.counts = collections.Counter()

    def size(self):
        return len(self.ordered)

    def clear(self):
        self.ordered = []
        self.counts = collections.Counter()

    def insert(self, keys, selected=None):
        keys = list(np.asarray(keys).tolist())
        sel = [True] * len(keys) if selected is None else list(np.asarray(selected, bool).tolist())
        room = self.capacity - len(self.ordered)
        failed = 0
        for k, s in zip(keys, sel):
            if not s:
                continue
            if k == self.empty:
                failed += 1
            elif room > 0:
                self.ordered.append(k)
                self.counts[k] += 1
                room -= 1
            else:
                failed += 1
        return failed

    def count_each(self, keys, selected=None):
        keys = np.asarray(keys).tolist()
        sel = [True] * len(keys) if selected is None else list(np.asarray(selected, bool).tolist())
        return np.array([self.counts.get(k, 0) if s else 0 for k, s in zip(keys, sel)], np.uint64)

    def count_each_outer(self, keys):
        c = self.count_each(keys)
        return np.where(c == 0, np.uint64(1), c)

    def contains(self, keys, selected=None):
        return self.count_each(keys, selected) > 0

    def find(self, keys, selected=None):
        k = np.asarray(keys)
        return np.where(self.contains(keys, selected), k, np.array(self.empty, k.dtype))

    def count(self, keys):
        return int(self.count_each(keys).sum())

    def retrieve(self, keys):
        out = [k for k in np.asarray(keys).tolist() for _ in range(self.counts.get(k, 0))]
        return np.array(out), np.array(out)

    def retrieve_all(self):
        return np.array(self.ordered)
