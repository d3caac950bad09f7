"""StaticMap / StaticSet with 1- and 2-byte keys or payloads (Pair layouts of
2 to 16 bytes with padding, the 2-byte and 1-byte CAS done on the 32-bit
word that holds the slot), through the C ABI, against the oracle:
- failure counts, Find / Contains / Count / InsertAndFind / InsertOrAssign /
  Erase outputs bit-exact for 17 layouts x 2 probing schemes x 4 group widths;
- slots placed one key per call equal the oracle's slot array for each hash
  family (this pins the device murmurhash3_32 / xxhash_32 / xxhash_64 of 1-
  and 2-byte keys);
- the reference's own narrow cases: two int16 keys sharing one 32-bit word
  (tests/static_set/signed_cas_test.cpp:19-61) and an insert into a slot
  with nonzero padding (tests/static_map/insert_test.cpp:35-72);
- binary16 keys compare with IEEE == (NaN sentinel, +0 == -0)."""
import ctypes

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

TD = {np.int8: torch.int8, np.uint8: torch.uint8, np.int16: torch.int16, np.int32: torch.int32,
      np.int64: torch.int64, np.float16: torch.float16}
PROBE = {O.PROBE_LINEAR: "linear", O.PROBE_DOUBLE: "double"}
HASH = {O.HASH_MURMUR3_32: "murmurhash3_32", O.HASH_XXHASH_32: "xxhash_32", O.HASH_XXHASH_64: "xxhash_64"}

# (key, payload) dtypes; None = a set
LAYOUTS = [(np.uint8, None), (np.int8, None), (np.int16, None),
           (np.int8, np.int8), (np.uint8, np.int16), (np.uint8, np.int32), (np.int8, np.int32), (np.int8, np.int64),
           (np.int16, np.int8), (np.int16, np.int16), (np.int16, np.int32), (np.int16, np.int64),
           (np.int32, np.int8), (np.int32, np.int16), (np.int64, np.int8), (np.int64, np.int16),
           (np.uint8, np.uint8)]


def mx():
    import mxco
    return mxco


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def sentinel(dt):
    return dt(255) if dt == np.uint8 else dt(-1)


def value_of(keys, vd):
    """The payload of a key, so duplicates of a key carry equal payloads;
    never the empty payload (a reader of a racing InsertAndFind waits for
    the payload to differ from it)."""
    v = (keys.astype(np.int64) * 7 + 1)
    if np.dtype(vd).itemsize < 8:
        bits = 8 * np.dtype(vd).itemsize
        v = v & ((1 << bits) - 1)
        if np.issubdtype(vd, np.signedinteger):
            v = np.where(v >= 1 << (bits - 1), v - (1 << bits), v)
    v = np.where(v == sentinel(vd).astype(np.int64), 0, v)
    return v.astype(vd)


def keys_for(rng, kd):
    """Distinct non-sentinel keys: all of them for 1-byte keys, 12000 for
    wider ones; plus the absent keys for queries."""
    if np.dtype(kd).itemsize == 1:
        allk = np.arange(256).astype(np.int64).astype(kd)
        allk = allk[allk != sentinel(kd)]
        rng.shuffle(allk)
        return allk[:200], allk[200:]
    hi = 1 << 15 if kd == np.int16 else 1 << 30
    k = rng.choice(hi, 16000, replace=False).astype(kd)
    return k[:12000], k[12000:]


def make(kd, vd, cap, probing=O.PROBE_LINEAR, bucket=5, hash_family=O.HASH_MURMUR3_32, group=0):
    mxco = mx()
    if vd is None:
        m = mxco.StaticSet(cap, sentinel(kd).item(), TD[kd], bucket_size=bucket, probing=PROBE[probing],
                           hash=HASH[hash_family], group_size=group)
    else:
        m = mxco.StaticMap(cap, sentinel(kd).item(), sentinel(vd).item(), TD[kd], TD[vd], bucket_size=bucket,
                           probing=PROBE[probing], hash=HASH[hash_family], group_size=group)
    ref = O.OpenAddressing(kd, vd, cap, sentinel(kd), None if vd is None else sentinel(vd), bucket=bucket,
                           probing=probing, hash_family=hash_family)
    assert m.capacity() == ref.capacity and m.slot_size == ref.slot_size
    return m, ref


def pairs_of(keys, vd):
    return dev(O.pack_pairs(keys, None if vd is None else value_of(keys, vd)))


@pytest.mark.parametrize("kd,vd", LAYOUTS, ids=lambda x: getattr(x, "__name__", str(x)))
@pytest.mark.parametrize("probing", [O.PROBE_LINEAR, O.PROBE_DOUBLE])
@pytest.mark.parametrize("group", [1, 2, 4, 8])
def test_narrow_layouts_vs_oracle(kd, vd, probing, group):
    rng = np.random.default_rng(group * 7 + probing)
    present, absent = keys_for(rng, kd)
    cap = int(len(present) * 1.4)
    m, ref = make(kd, vd, cap, probing, group=group)
    ins = np.concatenate([present, present[rng.integers(0, len(present), len(present) // 3)],
                          np.full(3, sentinel(kd), kd)])
    rng.shuffle(ins)
    assert m.insert(pairs_of(ins, vd)) == ref.insert(ins, None if vd is None else value_of(ins, vd))
    q = np.concatenate([present, absent, np.full(2, sentinel(kd), kd)])
    rng.shuffle(q)
    rv, rf = ref.find(q)
    assert m.find(dev(q)).cpu().numpy().tobytes() == rv.tobytes()
    assert m.contains(dev(q)).cpu().numpy().astype(bool).tolist() == rf.tolist()
    assert m.count(dev(q)) == int(rf.sum())
    # every present key sits in one slot with its payload
    s = m.slots().cpu().numpy()
    ks = np.dtype(kd).itemsize
    stored = np.ascontiguousarray(s[:, :ks]).view(kd).reshape(-1)
    stored = stored[stored != sentinel(kd)]
    assert np.array_equal(np.sort(stored), np.sort(present))


@pytest.mark.parametrize("kd,vd", [(np.int8, np.int32), (np.int16, None), (np.int16, np.int16),
                                   (np.uint8, np.int64), (np.int64, np.int8)],
                         ids=lambda x: getattr(x, "__name__", str(x)))
@pytest.mark.parametrize("bucket", [1, 8])
def test_narrow_insert_and_find_assign_erase(kd, vd, bucket):
    rng = np.random.default_rng(bucket)
    present, absent = keys_for(rng, kd)
    er = np.dtype(kd).type(-2) if kd != np.uint8 else np.uint8(254)
    present, absent = present[present != er], absent[absent != er]
    cap = int(len(present) * 1.6)
    m, ref = make(kd, vd, cap, bucket=bucket, group=4)
    half = present[: len(present) // 2]
    assert m.insert(pairs_of(half, vd)) == ref.insert(half, None if vd is None else value_of(half, vd)) == 0
    # InsertAndFind: the first half is present, the second half new (twice)
    rest = present[len(present) // 2:]
    iaf = np.concatenate([half[:50], rest, rest[:40]])
    rng.shuffle(iaf)
    of, oi = m.insert_and_find(pairs_of(iaf, vd))
    rf_, ri_ = ref.insert_and_find(iaf, None if vd is None else value_of(iaf, vd))
    assert of.cpu().numpy().tobytes() == rf_.tobytes()
    assert int(oi.sum().item()) == int(ri_.sum()) == len(rest)
    if vd is not None:
        # InsertOrAssign: new payloads for a third of the keys plus some new keys
        ak = present[::3]
        av = (value_of(ak, vd).astype(np.int64) + 3).astype(vd)
        f, a = m.insert_or_assign(dev(O.pack_pairs(ak, av)))
        assert (f, a) == ref.insert_or_assign(ak, av)
    # Erase with a tombstone sentinel
    m.set_erased_key(er.item())
    gone = present[::4]
    assert m.erase(dev(gone)) == ref.erase(gone, er) == 0
    q = np.concatenate([present, absent])
    rv, rfl = ref.find(q)
    assert m.find(dev(q)).cpu().numpy().tobytes() == rv.tobytes()
    assert m.contains(dev(q)).cpu().numpy().astype(bool).tolist() == rfl.tolist()


@pytest.mark.parametrize("kd,vd,probing,hash_family,bucket", [
    (np.uint8, np.int32, O.PROBE_LINEAR, O.HASH_MURMUR3_32, 1),
    (np.int8, None, O.PROBE_DOUBLE, O.HASH_XXHASH_32, 5),
    (np.int16, None, O.PROBE_DOUBLE, O.HASH_XXHASH_32, 5),
    (np.int16, np.int16, O.PROBE_LINEAR, O.HASH_XXHASH_64, 3),
    (np.int16, np.int64, O.PROBE_DOUBLE, O.HASH_MURMUR3_32, 2),
], ids=str)
def test_sequential_slots_equal_oracle(kd, vd, probing, hash_family, bucket):
    """One key per Insert call, so the order is the oracle's: the slot arrays
    are equal field by field (padding aside), which requires the device hash
    of 1- and 2-byte keys to equal the oracle's."""
    rng = np.random.default_rng(3)
    present, _ = keys_for(rng, kd)
    present = present[:150]
    m, ref = make(kd, vd, 200, probing, bucket, hash_family)
    pairs = pairs_of(present, vd)
    for i in range(len(present)):
        assert m.insert(pairs[i:i + 1]) == 0
    assert ref.insert(present, None if vd is None else value_of(present, vd)) == 0
    got, exp = m.slots().cpu().numpy(), ref.slots()
    ks = np.dtype(kd).itemsize
    assert np.array_equal(got[:, :ks], exp[:, :ks])
    if vd is not None:
        vs, voff = np.dtype(vd).itemsize, max(ks, np.dtype(vd).itemsize)
        assert np.array_equal(got[:, voff:voff + vs], exp[:, voff:voff + vs])


def test_int16_neighbours_in_one_cas_word():
    """tests/static_set/signed_cas_test.cpp:19-61: capacity 2, bucket 1,
    LinearProbing<murmurhash3_32<int16_t>>: key 2 starts at slot 1 and -2 at
    slot 0, so both live in one 32-bit word; inserting and erasing one keeps
    the other."""
    m, ref = make(np.int16, None, 2, O.PROBE_LINEAR, 1, O.HASH_MURMUR3_32)
    assert m.capacity() == 2
    assert ref.probe_start(2)[0] == 1 and ref.probe_start(-2)[0] == 0
    pos, neg = np.array([2], np.int16), np.array([-2], np.int16)
    assert m.insert(dev(pos)) == 0
    assert m.insert(dev(neg)) == 0
    q = dev(np.array([2, -2], np.int16))
    assert m.contains(q).cpu().numpy().tolist() == [True, True]
    assert m.erase(dev(neg)) == 0
    assert m.contains(q).cpu().numpy().tolist() == [True, False]
    assert m.slots().cpu().numpy().view(np.int16).reshape(-1).tolist() == [-1, 2]


@pytest.mark.parametrize("kd,vd", [(np.uint8, np.uint32), (np.int8, np.int32)], ids=str)
@pytest.mark.parametrize("op", ["insert", "insert_or_assign"])
def test_insert_into_slot_with_nonzero_padding(kd, vd, op):
    """tests/static_map/insert_test.cpp:35-72: Pair<uint8_t, uint32_t> has 3
    padding bytes.  The slots' padding is set to 0xA5 (the fields keep the
    sentinels); the whole-slot CAS compares against the slot's own padding,
    so the insert succeeds."""
    mxco = mx()
    vdt = np.uint32 if vd == np.uint32 else np.int32
    m = mxco.StaticMap(128, sentinel(kd).item(), -1, TD[kd], torch.int32, bucket_size=1, probing="linear")
    cap, ss = m.capacity(), m.slot_size
    assert ss == 8
    img = m.slots().cpu().numpy().copy()
    img[:, 1:4] = 0xA5
    src = torch.from_numpy(img.reshape(-1)).cuda()
    torch.cuda.synchronize()
    assert mxco._hip.hipMemcpy(ctypes.c_void_p(m.data()), ctypes.c_void_p(src.data_ptr()),
                               ctypes.c_size_t(cap * ss), 3) == 0
    key = np.array([7], kd)
    pairs = dev(O.pack_pairs(key, np.array([31], vdt)))
    if op == "insert":
        assert m.insert(pairs) == 0
        want = 31
    else:
        assert m.insert_or_assign(pairs) == (0, 0)
        assert m.insert_or_assign(dev(O.pack_pairs(key, np.array([42], vdt)))) == (0, 1)
        want = 42
    assert m.count(dev(key)) == 1
    assert m.find(dev(key)).cpu().numpy().view(vdt).tolist() == [want]
    # the other slots keep their padding
    after = m.slots().cpu().numpy()
    assert (after[:, 1:4] == 0xA5).sum() >= 3 * (cap - 1)


def test_half_keys_compare_with_ieee_equality():
    kd, vd = np.float16, np.int32
    keys = np.array([0.0, 1.5, -2.25, 65504.0, -0.0, 6e-8, np.inf, -np.inf], np.float16)
    vals = np.arange(len(keys), dtype=np.int32)
    q = np.array([-0.0, 0.0, 1.5, 3.0, 65504.0, 6e-8, np.inf, -np.inf, np.nan], np.float16)
    for g in (1, 2, 4, 8):
        mxco = mx()
        m = mxco.StaticMap(64, float("nan"), -1, torch.float16, torch.int32, group_size=g)
        ref = O.OpenAddressing(kd, vd, 64, np.float16("nan"), -1)
        assert m.insert(dev(O.pack_pairs(keys, vals))) == ref.insert(keys, vals)
        rv, rf = ref.find(q)
        assert m.find(dev(q)).cpu().numpy().tobytes() == rv.tobytes()
        assert m.contains(dev(q)).cpu().numpy().astype(bool).tolist() == rf.tolist()


def test_narrow_creation_rules():
    mxco = mx()
    with pytest.raises(Exception):
        mxco.StaticSet(100, 0, torch.int8, hash="murmurhash3_fmix32")  # fmix32 hashes 4-byte keys only
    # a 1-byte floating key type does not exist: the C ABI refuses it
    d = mxco.OaDesc(1, 0, 1, 5, 0, 0, 0, 0)
    h = ctypes.c_void_p()
    ek = ctypes.create_string_buffer(b"\xff", 1)
    assert mxco.lib().mxco_oa_create(ctypes.byref(d), 100, ek, None, None, ctypes.byref(h)) != 0
