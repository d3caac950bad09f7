"""The walk-free placement of the region build, restated on the CPU.

Inserting keys into an empty region in the order of their first slots, each
linear-probing walk ends at the first slot past the previous key's that is
not below its own first slot, so the i-th kept key lands at
    i + max_{j <= i}(first_j - j)
(a prefix maximum).  These tests pin that identity against a sequential
slot-by-slot walk, and check that the placement is one the CAS walk can
produce: every key sits at the first empty slot of its walk at the time it is
inserted in that order, and the occupied slots equal those of any other
insertion order (open_addressing_ref_impl.h:490-532 of the reference: the
first slot that is EMPTY or holds the key ends the walk)."""
import numpy as np
import pytest


def sequential(firsts, rlen):
    """Slot of each key, inserted one by one in the given order by a walk
    from its first slot over the region; -1 past the region's end."""
    used = np.zeros(rlen, bool)
    out = []
    for f in firsts:
        p = f
        while p < rlen and used[p]:
            p += 1
        if p < rlen:
            used[p] = True
            out.append(p)
        else:
            out.append(-1)
    return np.array(out), used


def prefix_max(firsts_sorted, rlen):
    """The placement's closed form over keys sorted by first slot."""
    i = np.arange(len(firsts_sorted))
    q = i + np.maximum.accumulate(firsts_sorted - i)
    return np.where(q < rlen, q, -1)


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("load", [0.3, 0.8, 0.97, 1.2])
def test_prefix_max_equals_sequential_walk(seed, load):
    rng = np.random.default_rng(seed)
    rlen = int(rng.integers(16, 4097))
    bucket = int(rng.choice([1, 2, 5, 8]))
    n = int(load * rlen)
    firsts = np.sort(rng.integers(0, rlen // bucket, n) * bucket)
    want, _ = sequential(firsts, rlen)
    assert np.array_equal(prefix_max(firsts, rlen), want)


@pytest.mark.parametrize("seed", range(8))
def test_placement_is_a_cas_walk_outcome(seed):
    """Keys placed by the closed form sit after full slots only, and any
    insertion order of the same keys occupies the same slots."""
    rng = np.random.default_rng(100 + seed)
    rlen = 2048
    n = int(0.9 * rlen)
    firsts = np.sort(rng.integers(0, rlen // 5, n) * 5)
    q = prefix_max(firsts, rlen)
    placed = q >= 0
    occupied = np.zeros(rlen, bool)
    occupied[q[placed]] = True
    for f, p in zip(firsts[placed], q[placed]):
        assert f <= p and occupied[f:p + 1].all()
    _, used = sequential(rng.permutation(firsts), rlen)
    # overflowing keys aside, the occupied set does not depend on the order
    # of a region whose keys all fit
    if placed.all():
        assert np.array_equal(used, occupied)


def test_duplicates_are_dropped_before_ranking():
    """A second copy of a key in the same first-slot group takes no slot:
    ranking only the first copies gives the sequential result where copies
    stop at the slot holding their key."""
    rng = np.random.default_rng(7)
    rlen = 1000
    keys = rng.integers(0, 600, 900)
    firsts = (keys * 2654435761 % (rlen // 5)) * 5
    order = np.lexsort((keys, firsts))
    fk, kk = firsts[order], keys[order]
    keep = np.ones(len(kk), bool)
    keep[1:] = ~((kk[1:] == kk[:-1]) & (fk[1:] == fk[:-1]))
    q = prefix_max(fk[keep], rlen)
    # sequential CAS walk with duplicate detection
    slot_key = np.full(rlen, -1)
    for f, k in zip(fk, kk):
        p = f
        while p < rlen and slot_key[p] != -1 and slot_key[p] != k:
            p += 1
        if p < rlen and slot_key[p] == -1:
            slot_key[p] = k
    placed = q >= 0
    got = np.full(rlen, -1)
    got[q[placed]] = kk[keep][placed]
    assert np.array_equal(got, slot_key)
