"""mxco_bloom_add_noalloc, the Add that allocates nothing (its partitioned
pass runs in the library's fixed per-GPU scratch), gives the filter words of
mxco_bloom_add, bit for bit, on consecutive calls (the second smaller than
the first, so it reuses scratch the first left behind), with and without a
stencil, on both Add paths; oversized calls raise a length error."""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 22


def mx():
    import mxco
    return mxco


def words(f):
    return f.words().cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("kd", [torch.int64, torch.int32])
def test_noalloc_add_equals_add(kd):
    mxco = mx()
    print("no-alloc Add, key dtype", kd, flush=True)
    nb = 1 << 16  # 2 MiB filter, 32 ranges
    g = torch.Generator(device="cuda")
    g.manual_seed(31)
    lo, hi = (-(1 << 63), (1 << 63) - 1) if kd == torch.int64 else (-(1 << 31), (1 << 31) - 1)
    ref = mxco.BloomFilter(nb, kd, seed=5)
    nal = mxco.BloomFilter(nb, kd, seed=5)
    # MAX_KEYS then fewer keys (partitioned), a stencil call, then a small
    # batch (direct path)
    for n, with_stencil in ((MAX_KEYS, False), (300001, False), (200000, True), (1000, False)):
        keys = torch.randint(lo, hi, (n,), dtype=kd, device="cuda", generator=g)
        st = (torch.rand(n, device="cuda", generator=g) < 0.5).to(torch.uint8) if with_stencil else None
        if st is None:
            ref.add(keys)
        else:
            ref.add_if(keys, st)
        nal.add_noalloc(keys, stencil=st)
        assert np.array_equal(words(nal), words(ref)), n
    assert words(nal).any()


def test_noalloc_add_matches_oracle_and_contains():
    mxco = mx()
    nb = 3000  # not a multiple of the 2048-block range
    rng = np.random.default_rng(4)
    keys = rng.integers(-2**63, 2**63 - 1, 1 << 20, dtype=np.int64)
    f = mxco.BloomFilter(nb, torch.int64)
    f.add_noalloc(torch.from_numpy(keys).cuda())
    w = np.zeros(nb * 8, np.uint32)
    O.bloom_add(w, nb, keys)
    assert np.array_equal(words(f), w)
    assert f.contains(torch.from_numpy(keys).cuda()).all()


def test_noalloc_add_rejects_oversized_batches():
    mxco = mx()
    f = mxco.BloomFilter(1 << 16, torch.int64)
    keys = torch.zeros(MAX_KEYS + 1, dtype=torch.int64, device="cuda")
    with pytest.raises(mxco.LengthError):
        f.add_noalloc(keys)
    assert not words(f).any()
    f.add_noalloc(keys, n=0)
