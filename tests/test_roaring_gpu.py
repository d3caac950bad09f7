"""RoaringBitmap on the GPU through the C ABI: membership of uint32 / uint64
keys in portable-serialization bitmaps written by the oracle's writer, against
the key sets themselves (brute force).  The first test queries the reference's
three RoaringFormatSpec files (tests/roaring_bitmap/contains_test.cpp:428-476),
which the writer rebuilds byte for byte (their SHA-256 is checked on the CPU)."""
import struct

import numpy as np
import pytest
import torch

import roaring as R

pytestmark = pytest.mark.gpu


def mx():
    import mxco
    return mxco


def query(rb, keys, dtype):
    arr = np.asarray(keys, dtype=np.uint32 if dtype == torch.int32 else np.uint64)
    t = torch.from_numpy(arr.view(np.int32 if dtype == torch.int32 else np.int64)).cuda()
    return rb.contains(t).cpu().numpy()


def spec_keys32():
    k = list(range(0, 100000, 1000)) + [3 * x for x in range(100000, 200000)] + list(range(700000, 800000))
    return np.array(k, np.uint64)


def spec_keys64():
    low = list(range(0, 0x9000)) + [0x9000] + list(range(0xA000, 0x10000)) + [0x10000, 0x20000, 0x20005]
    low += [0x80000 + k for k in range(0, 0x10000, 2)]
    return np.array(low + [(1 << 32) | k for k in low], np.uint64)


@pytest.mark.parametrize("runs", [False, "optimize"])
def test_spec_file_key_sets(runs, image):
    keys = spec_keys32()
    rb = mx().RoaringBitmap(R.serialize32(keys, run_containers=runs), torch.int32)
    assert rb.size() == len(np.unique(keys))
    assert query(rb, keys, torch.int32).all()
    absent = np.setdiff1d(np.arange(0, 900000, 7, dtype=np.uint64), keys)
    assert not query(rb, absent, torch.int32).any()
    keys64 = spec_keys64()
    rb64 = mx().RoaringBitmap(R.serialize64(keys64, run_containers="optimize"), torch.int64)
    assert rb64.size() == len(np.unique(keys64))
    assert query(rb64, keys64, torch.int64).all()
    absent64 = np.setdiff1d(np.concatenate([keys64 + 1, keys64 + (2 << 32)]), keys64)
    assert not query(rb64, absent64, torch.int64).any()


@pytest.fixture(params=["bitset_image", "container_image"])
def image(request, monkeypatch):
    """The library expands containers to bitsets within a byte budget
    (MXCO_RB_BITSET_MAX_MB); 0 keeps array and run containers as they are."""
    if request.param == "container_image":
        monkeypatch.setenv("MXCO_RB_BITSET_MAX_MB", "0")
    return request.param


@pytest.mark.parametrize("layout", ["array", "bitset", "run", "mixed", "few_runs", "empty"])
def test_random_bitmaps_vs_brute_force(layout, image):
    rng = np.random.default_rng(len(layout))
    if layout == "empty":
        keys = np.array([], np.uint64)
    elif layout == "bitset":
        keys = np.unique(rng.integers(0, 1 << 19, 60000)).astype(np.uint64)
    elif layout in ("run", "few_runs"):
        starts = rng.integers(0, 1 << 22 if layout == "run" else 1 << 17, 300)
        keys = np.unique(np.concatenate([np.arange(s, s + rng.integers(1, 500)) for s in starts])).astype(np.uint64)
    else:
        keys = np.unique(rng.integers(0, 1 << 32, 20000)).astype(np.uint64)
    runs = {"run": True, "few_runs": True, "mixed": set(range(0, 1 << 16, 3))}.get(layout, False)
    data = R.serialize32(keys, run_containers=runs)
    rb = mx().RoaringBitmap(data, torch.int32)
    assert rb.size() == len(keys) and rb.size_bytes() == len(data)
    q = np.concatenate([keys, rng.integers(0, 1 << 32, 50000).astype(np.uint64), keys + 1, np.maximum(keys, 1) - 1])
    q = q.astype(np.uint64) & 0xFFFFFFFF
    assert np.array_equal(query(rb, q, torch.int32), np.isin(q, keys))


def test_malformed_inputs_raise():
    mxco = mx()
    good = R.serialize32([1, 2, 3, 70000])
    with pytest.raises(ValueError):
        mxco.RoaringBitmap(struct.pack("<II", 99, 0), torch.int32)  # bad cookie
    with pytest.raises(ValueError):
        mxco.RoaringBitmap(good[:-1], torch.int32)  # truncated
    bad = bytearray(R.serialize32([1, 70000]))
    struct.pack_into("<H", bad, 12, 0)  # second container key 0 <= first key 0
    with pytest.raises(ValueError):
        mxco.RoaringBitmap(bytes(bad), torch.int32)
    bad = bytearray(good)
    struct.pack_into("<I", bad, 16, 3)  # container offset inside the header
    with pytest.raises(ValueError):
        mxco.RoaringBitmap(bytes(bad), torch.int32)
    two = bytearray(R.serialize64([5, (1 << 32) + 5]))
    hk2 = 8 + 4 + len(R.serialize32([5]))
    struct.pack_into("<I", two, hk2, 0)  # bucket keys 0, 0
    with pytest.raises(ValueError):
        mxco.RoaringBitmap(bytes(two), torch.int64)
    rb = mxco.RoaringBitmap(good, torch.int32)
    assert rb.contains(torch.empty(0, dtype=torch.int32, device="cuda")).numel() == 0


def test_search_paths_dense_budget_and_long_run_lists(image):
    """Buckets past the dense-table budget (binary search of the key list)
    and a run container of 32,768 runs (fence list longer than one line)."""
    rng = np.random.default_rng(12)
    hk = np.arange(300, dtype=np.uint64) << np.uint64(32)
    lo = (np.arange(100, dtype=np.uint64) << np.uint64(16)) + np.uint64(5)
    keys = np.unique((hk[:, None] + lo[None, :]).ravel())
    rb = mx().RoaringBitmap(R.serialize64(keys), torch.int64)
    q = np.concatenate([keys, keys + 1, rng.integers(0, 1 << 41, 100000, dtype=np.uint64)])
    assert np.array_equal(query(rb, q, torch.int64), np.isin(q, keys))
    alt = np.arange(0, 65536, 2, dtype=np.uint64)
    keys = np.concatenate([alt, np.uint64(1 << 16) + alt[:3000]])
    rb = mx().RoaringBitmap(R.serialize32(keys, run_containers=True), torch.int32)
    q = np.concatenate([keys, keys + 1, rng.integers(0, 1 << 18, 100000, dtype=np.uint64)]) & 0xFFFFFFFF
    assert np.array_equal(query(rb, q, torch.int32), np.isin(q, keys))


def test_one_bucket_64bit_bitmap(image):
    """A 64-bit bitmap of one bucket (high half 7): keys of other high halves
    are absent even when their low half is present."""
    rng = np.random.default_rng(13)
    low = np.unique(rng.integers(0, 1 << 20, 30000, dtype=np.uint64))
    keys = low | np.uint64(7 << 32)
    rb = mx().RoaringBitmap(R.serialize64(keys, run_containers="optimize"), torch.int64)
    q = np.concatenate([keys, low, low | np.uint64(8 << 32), keys + 1])
    assert np.array_equal(query(rb, q, torch.int64), np.isin(q, keys))


@pytest.mark.parametrize("offset", [0, 1])
def test_small_bitmap_lds_path_and_ragged_tail(offset):
    """A uint32 bitmap of a few containers and a large query batch (the
    LDS-resident path when the key pointer is 16-byte aligned, the general
    path at an offset of one key), with a ragged tail."""
    rng = np.random.default_rng(14)
    keys = np.unique(np.concatenate([rng.integers(k << 16, (k << 16) + 65536, 9000) for k in (1, 4, 7, 8, 300)]))
    keys = keys.astype(np.uint64)
    rb = mx().RoaringBitmap(R.serialize32(keys, run_containers="optimize"), torch.int32)
    n = (1 << 17) + 3
    q = np.concatenate([keys[rng.integers(0, len(keys), n // 2)], rng.integers(0, 1 << 25, n - n // 2)])
    q = q.astype(np.uint64)
    t = torch.from_numpy(np.concatenate([np.zeros(offset, np.uint32), q.astype(np.uint32)]).view(np.int32)).cuda()
    got = rb.contains(t[offset:]).cpu().numpy()
    assert np.array_equal(got, np.isin(q, keys))
