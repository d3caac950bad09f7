"""Python front end of libmxco.so (ctypes over the C ABI in include/mxco.h).

The classes mirror the reference's container API (aclco::StaticMap,
StaticSet, StaticMultimap, BloomFilter, HyperLogLog): same method names
(snake_case), same argument meaning, same error behaviour.  Device buffers
are torch tensors on the GPU; streams default to torch's current stream.
The library must be built (ops-collections_amd/build.py): importing this
module without it raises ImportError -- there is no CPU fallback.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libmxco.so")

if not os.path.exists(LIB_PATH):
    raise ImportError("libmxco.so is not built (run ops-collections_amd/build.py): %s" % LIB_PATH)

_lib = ctypes.CDLL(LIB_PATH)
_hip = ctypes.CDLL("libamdhip64.so")  # already loaded as a dependency of libmxco.so

vp = ctypes.c_void_p
u32 = ctypes.c_uint32
u64 = ctypes.c_uint64
i32 = ctypes.c_int
pu64 = ctypes.POINTER(ctypes.c_uint64)

OK, INVALID_ARGUMENT, LENGTH_ERROR, OUT_OF_MEMORY, RUNTIME_ERROR, LOGIC_ERROR, UNSUPPORTED, OVERFLOW_ERROR = range(8)
PROBING = {"linear": 0, "double": 1}
HASH = {"murmurhash3_32": 0, "xxhash_32": 1, "xxhash_64": 2, "murmurhash3_fmix32": 3, "murmurhash3_fmix64": 4}
PRED = {"always": 0, "nonzero": 1, "odd": 2, "even": 3}


def _release(obj, destroy):
    """Frees obj._h with the named C-ABI destructor (no-op at interpreter exit,
    when the module globals are already gone)."""
    h = getattr(obj, "_h", None)
    obj._h = None
    try:
        if h and _lib is not None:
            getattr(_lib, destroy)(h)
    except (NameError, AttributeError, TypeError):
        pass  # interpreter shutdown: the library handle is gone


class LengthError(ValueError):
    """std::length_error: batch beyond the kernel index range."""


class LogicError(RuntimeError):
    """std::logic_error: moved-from or uninitialised object."""


_EXC = {INVALID_ARGUMENT: ValueError, LENGTH_ERROR: LengthError, OUT_OF_MEMORY: MemoryError,
        RUNTIME_ERROR: RuntimeError, LOGIC_ERROR: LogicError, UNSUPPORTED: NotImplementedError,
        OVERFLOW_ERROR: OverflowError}


class OaDesc(ctypes.Structure):
    _fields_ = [("key_size", u32), ("value_size", u32), ("key_kind", u32), ("bucket_size", u32),
                ("probing", u32), ("hash", u32), ("group_size", u32), ("value_kind", u32)]


class Stencil(ctypes.Structure):
    _fields_ = [("data", vp), ("elem_size", u32), ("pred", u32)]


def _sig(name, args, res=i32):
    f = getattr(_lib, name)
    f.argtypes = args
    f.restype = res
    return f


_lib.mxco_last_error.restype = ctypes.c_char_p
_sig("mxco_valid_extent", [u64, u32, u32, pu64])
_sig("mxco_device_cu_count", [ctypes.POINTER(i32)])
_sig("mxco_oa_create", [ctypes.POINTER(OaDesc), u64, vp, vp, vp, ctypes.POINTER(vp)])
_sig("mxco_oa_destroy", [vp])
_sig("mxco_oa_capacity", [vp, pu64])
_sig("mxco_oa_data", [vp, ctypes.POINTER(vp)])
_sig("mxco_oa_slot_size", [vp, ctypes.POINTER(u32)])
_sig("mxco_oa_set_erased_key", [vp, vp, vp])
_sig("mxco_oa_clear", [vp, vp, i32])
_sig("mxco_oa_insert", [vp, vp, ctypes.POINTER(Stencil), u64, vp, i32, pu64])
_sig("mxco_oa_insert_or_assign", [vp, vp, u64, vp, i32, pu64, pu64])
_sig("mxco_oa_assign_if", [vp, vp, ctypes.POINTER(Stencil), u64, vp, i32])
_sig("mxco_oa_insert_and_find", [vp, vp, vp, vp, u64, vp, i32])
_sig("mxco_oa_find", [vp, vp, ctypes.POINTER(Stencil), vp, u64, vp, i32])
_sig("mxco_oa_contains", [vp, vp, ctypes.POINTER(Stencil), vp, u64, vp, i32])
_sig("mxco_oa_count", [vp, vp, u64, vp, pu64])
_sig("mxco_oa_erase", [vp, vp, u64, vp, i32, pu64])
_sig("mxco_oa_insert_stats", [vp, vp, u64, vp, pu64, pu64])
_sig("mxco_oa_probe_stats", [vp, vp, u64, vp, pu64, pu64])


def lib():
    return _lib


def check(status):
    if status != OK:
        msg = _lib.mxco_last_error().decode()
        raise _EXC.get(status, RuntimeError)(msg)


def _stream(stream):
    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, torch.cuda.Stream):
        return stream.cuda_stream
    return stream


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _scalar_bytes(value, dtype):
    t = torch.tensor([value], dtype=dtype)
    return ctypes.create_string_buffer(t.numpy().tobytes(), t.element_size())


def _stencil(stencil, pred):
    if stencil is None:
        return None
    s = Stencil()
    s.data = _ptr(stencil)
    s.elem_size = stencil.element_size() if isinstance(stencil, torch.Tensor) else 1
    s.pred = PRED[pred] if isinstance(pred, str) else int(pred)
    return ctypes.byref(s)


def valid_extent(capacity, bucket_size=5, probing="linear"):
    out = u64()
    check(_lib.mxco_valid_extent(capacity, bucket_size, PROBING[probing], ctypes.byref(out)))
    return out.value


def cu_count():
    n = i32()
    check(_lib.mxco_device_cu_count(ctypes.byref(n)))
    return n.value


def slot_layout(key_dtype, value_dtype):
    """(slot size, payload offset) of Pair<Key, T> (include/pair.h:41)."""
    ks = torch.empty(0, dtype=key_dtype).element_size()
    if value_dtype is None:
        return ks, 0
    vs = torch.empty(0, dtype=value_dtype).element_size()
    align = 1
    while align < ks + vs:
        align <<= 1
    return min(align, 16), max(ks, vs)


def make_pairs(keys, values, key_dtype=None, value_dtype=None):
    """Packs device keys/values into the slot layout: uint8 tensor (n, slot)."""
    key_dtype = key_dtype or keys.dtype
    value_dtype = value_dtype or values.dtype
    ss, off = slot_layout(key_dtype, value_dtype)
    n = keys.shape[0]
    out = torch.zeros((n, ss), dtype=torch.uint8, device=keys.device)
    ks = torch.empty(0, dtype=key_dtype).element_size()
    vs = torch.empty(0, dtype=value_dtype).element_size()
    out[:, :ks] = keys.to(key_dtype).contiguous().view(torch.uint8).view(n, ks)
    out[:, off:off + vs] = values.to(value_dtype).contiguous().view(torch.uint8).view(n, vs)
    return out


class _OpenAddressing:
    _set = False

    def __init__(self, capacity, empty_key, empty_value, key_dtype, value_dtype, bucket_size, probing, hash,
                 group_size, stream):
        self.key_dtype = key_dtype
        self.value_dtype = value_dtype
        ks = torch.empty(0, dtype=key_dtype).element_size()
        vs = 0 if value_dtype is None else torch.empty(0, dtype=value_dtype).element_size()
        d = OaDesc(ks, vs, 1 if key_dtype.is_floating_point else 0, bucket_size, PROBING[probing], HASH[hash],
                   group_size, 1 if value_dtype is not None and value_dtype.is_floating_point else 0)
        self._ek = _scalar_bytes(empty_key, key_dtype)
        self._ev = _scalar_bytes(empty_value, value_dtype) if value_dtype is not None else None
        h = vp()
        check(_lib.mxco_oa_create(ctypes.byref(d), capacity, self._ek, self._ev, _stream(stream), ctypes.byref(h)))
        self._h = h
        self.empty_key = empty_key
        self.empty_value = empty_value
        c, s = u64(), u32()
        check(_lib.mxco_oa_capacity(self._h, ctypes.byref(c)))
        check(_lib.mxco_oa_slot_size(self._h, ctypes.byref(s)))
        self._capacity = c.value
        self.slot_size = s.value

    def __del__(self, _release=_release):
        _release(self, "mxco_oa_destroy")

    def capacity(self):
        return self._capacity

    def data(self):
        p = vp()
        check(_lib.mxco_oa_data(self._h, ctypes.byref(p)))
        return p.value

    def set_erased_key(self, erased_key, stream=None):
        self._erk = _scalar_bytes(erased_key, self.key_dtype)
        check(_lib.mxco_oa_set_erased_key(self._h, self._erk, _stream(stream)))

    def clear(self, stream=None):
        check(_lib.mxco_oa_clear(self._h, _stream(stream), 1))

    def clear_async(self, stream=None):
        check(_lib.mxco_oa_clear(self._h, _stream(stream), 0))

    def insert(self, values, n=None, stream=None):
        n = values.shape[0] if n is None else n
        f = u64()
        check(_lib.mxco_oa_insert(self._h, _ptr(values), None, n, _stream(stream), 1, ctypes.byref(f)))
        return f.value

    def insert_async(self, values, n=None, stream=None):
        n = values.shape[0] if n is None else n
        check(_lib.mxco_oa_insert(self._h, _ptr(values), None, n, _stream(stream), 0, None))

    def insert_if(self, values, stencil, pred, n=None, stream=None):
        n = values.shape[0] if n is None else n
        f = u64()
        check(_lib.mxco_oa_insert(self._h, _ptr(values), _stencil(stencil, pred), n, _stream(stream), 1,
                                  ctypes.byref(f)))
        return f.value

    def insert_if_async(self, values, stencil, pred, n=None, stream=None):
        n = values.shape[0] if n is None else n
        check(_lib.mxco_oa_insert(self._h, _ptr(values), _stencil(stencil, pred), n, _stream(stream), 0, None))

    def insert_and_find(self, values, out_find=None, out_inserted=None, n=None, stream=None, sync=True):
        n = values.shape[0] if n is None else n
        dev = values.device if isinstance(values, torch.Tensor) else "cuda"
        od = self.value_dtype if self.value_dtype is not None else self.key_dtype
        if out_find is None:
            out_find = torch.empty(n, dtype=od, device=dev)
        if out_inserted is None:
            out_inserted = torch.empty(n, dtype=torch.bool, device=dev)
        check(_lib.mxco_oa_insert_and_find(self._h, _ptr(values), _ptr(out_find), _ptr(out_inserted), n,
                                           _stream(stream), int(sync)))
        return out_find, out_inserted

    def _out(self, keys, out, n, dtype):
        if out is None:
            dev = keys.device if isinstance(keys, torch.Tensor) else "cuda"
            out = torch.empty(n, dtype=dtype, device=dev)
        return out

    def find(self, keys, out=None, n=None, stream=None, stencil=None, pred="always", sync=True):
        n = keys.shape[0] if n is None else n
        od = self.value_dtype if self.value_dtype is not None else self.key_dtype
        out = self._out(keys, out, n, od)
        check(_lib.mxco_oa_find(self._h, _ptr(keys), _stencil(stencil, pred), _ptr(out), n, _stream(stream),
                                int(sync)))
        return out

    def find_async(self, keys, out=None, n=None, stream=None, stencil=None, pred="always"):
        return self.find(keys, out, n, stream, stencil, pred, sync=False)

    def contains(self, keys, out=None, n=None, stream=None, stencil=None, pred="always", sync=True):
        n = keys.shape[0] if n is None else n
        out = self._out(keys, out, n, torch.bool)
        check(_lib.mxco_oa_contains(self._h, _ptr(keys), _stencil(stencil, pred), _ptr(out), n, _stream(stream),
                                    int(sync)))
        return out

    def contains_async(self, keys, out=None, n=None, stream=None, stencil=None, pred="always"):
        return self.contains(keys, out, n, stream, stencil, pred, sync=False)

    def count(self, keys, n=None, stream=None):
        n = keys.shape[0] if n is None else n
        c = u64()
        check(_lib.mxco_oa_count(self._h, _ptr(keys), n, _stream(stream), ctypes.byref(c)))
        return c.value

    def erase(self, keys, n=None, stream=None):
        n = keys.shape[0] if n is None else n
        f = u64()
        check(_lib.mxco_oa_erase(self._h, _ptr(keys), n, _stream(stream), 1, ctypes.byref(f)))
        return f.value

    def erase_async(self, keys, n=None, stream=None):
        n = keys.shape[0] if n is None else n
        check(_lib.mxco_oa_erase(self._h, _ptr(keys), n, _stream(stream), 0, None))

    def probe_stats(self, keys, n=None, stream=None):
        """(lines, slots) the reference's Find walk reads for `keys`."""
        n = keys.shape[0] if n is None else n
        ln, sl = u64(), u64()
        check(_lib.mxco_oa_probe_stats(self._h, _ptr(keys), n, _stream(stream), ctypes.byref(ln), ctypes.byref(sl)))
        return ln.value, sl.value

    def insert_stats(self, values, n=None, stream=None):
        """Insert by the CAS walk, returning (CAS lost to another key, CAS made)."""
        n = values.shape[0] if n is None else n
        lost, made = u64(), u64()
        check(_lib.mxco_oa_insert_stats(self._h, _ptr(values), n, _stream(stream), ctypes.byref(lost),
                                        ctypes.byref(made)))
        return lost.value, made.value

    def insert_uses_region_build(self, n, stream=None):
        """Whether an Insert of n pairs takes the region build (a diagnostic)."""
        r = i32()
        check(_lib.mxco_oa_insert_path(self._h, n, _stream(stream), ctypes.byref(r)))
        return bool(r.value)

    def slots(self):
        """Device copy of the raw slot array as uint8 (capacity, slot_size)."""
        n = self._capacity * self.slot_size
        out = torch.empty(n, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        check(0 if _hip.hipMemcpy(vp(out.data_ptr()), vp(self.data()), ctypes.c_size_t(n), 3) == 0
              else RUNTIME_ERROR)
        return out.view(self._capacity, self.slot_size)

    def dump(self):
        """Live slots in slot order, read from Data() as the reference's
        DumpTable does (tests/common/dump_table.h:20-57): Capacity() slots of
        Pair<Key, Value> (key at byte 0, value at the next multiple of its
        own size) or of Key for a set, skipping slots whose key is bitwise the
        empty or, once set, the erased sentinel.  Returns host numpy arrays:
        keys for a set, (keys, values) for a map."""
        import numpy as np
        raw = self.slots().cpu().numpy()
        ks = torch.empty(0, dtype=self.key_dtype).element_size()
        kb = raw[:, :ks]
        live = ~np.all(kb == np.frombuffer(self._ek, np.uint8), axis=1)
        erk = getattr(self, "_erk", None)
        if erk is not None:
            live &= ~np.all(kb == np.frombuffer(erk, np.uint8), axis=1)
        keys = np.ascontiguousarray(kb[live]).view(torch.empty(0, dtype=self.key_dtype).numpy().dtype).reshape(-1)
        if self.value_dtype is None:
            return keys
        _, off = slot_layout(self.key_dtype, self.value_dtype)
        vs = torch.empty(0, dtype=self.value_dtype).element_size()
        vals = np.ascontiguousarray(raw[live, off:off + vs]).view(
            torch.empty(0, dtype=self.value_dtype).numpy().dtype).reshape(-1)
        return keys, vals


class StaticMap(_OpenAddressing):
    """aclco::StaticMap<Key, T> (include/static_map.h:29-35).

    Defaults: LinearProbing<murmurhash3_32<Key>>, Storage<5>."""

    def __init__(self, capacity, empty_key, empty_value, key_dtype=torch.int32, value_dtype=torch.int32,
                 bucket_size=5, probing="linear", hash="murmurhash3_32", group_size=0, stream=None):
        super().__init__(capacity, empty_key, empty_value, key_dtype, value_dtype, bucket_size, probing, hash,
                         group_size, stream)

    def insert_or_assign(self, values, n=None, stream=None):
        """Returns (failed, assigned)."""
        n = values.shape[0] if n is None else n
        f, a = u64(), u64()
        check(_lib.mxco_oa_insert_or_assign(self._h, _ptr(values), n, _stream(stream), 1, ctypes.byref(f),
                                            ctypes.byref(a)))
        return f.value, a.value

    def insert_or_assign_async(self, values, n=None, stream=None):
        n = values.shape[0] if n is None else n
        check(_lib.mxco_oa_insert_or_assign(self._h, _ptr(values), n, _stream(stream), 0, None, None))

    def assign_if_async(self, values, stencil, pred, n=None, stream=None):
        n = values.shape[0] if n is None else n
        check(_lib.mxco_oa_assign_if(self._h, _ptr(values), _stencil(stencil, pred), n, _stream(stream), 0))


class StaticSet(_OpenAddressing):
    """aclco::StaticSet<Key> (include/static_set.h:27-32).

    Defaults: DoubleHashing<xxhash_32<Key>>, Storage<5>."""
    _set = True

    def __init__(self, capacity, empty_key, key_dtype=torch.int32, bucket_size=5, probing="double",
                 hash="xxhash_32", group_size=0, stream=None):
        super().__init__(capacity, empty_key, None, key_dtype, None, bucket_size, probing, hash, group_size,
                         stream)


# ---------------------------------------------------------------------------
# BloomFilter / HyperLogLog
# ---------------------------------------------------------------------------

class BloomPolicy(ctypes.Structure):
    _fields_ = [("words_per_block", u32), ("pattern_bits", u32), ("seed", u64)]


_sig("mxco_bloom_create", [u64, u32, u64, vp, ctypes.POINTER(vp)])
_sig("mxco_bloom_create_with_policy", [u64, u32, ctypes.POINTER(BloomPolicy), vp, vp, ctypes.POINTER(vp)])
_sig("mxco_bloom_destroy", [vp])
_sig("mxco_bloom_data", [vp, ctypes.POINTER(vp), pu64])
_sig("mxco_bloom_clear", [vp, vp, i32])
_sig("mxco_bloom_add", [vp, vp, vp, u64, vp, i32])
_sig("mxco_bloom_add_noalloc", [vp, vp, vp, u64, vp, i32])
_sig("mxco_bloom_contains", [vp, vp, vp, vp, u64, vp, i32])
_sig("mxco_bloom_merge", [vp, vp, vp, i32])
_sig("mxco_bloom_intersect", [vp, vp, vp, i32])
_sig("mxco_hll_precision_for_sketch_kb", [u32, ctypes.POINTER(u32)])
_sig("mxco_hll_precision_for_stddev", [ctypes.c_double, ctypes.POINTER(u32)])
_sig("mxco_hll_create", [u32, u32, vp, ctypes.POINTER(vp)])
_sig("mxco_hll_destroy", [vp])
_sig("mxco_hll_data", [vp, ctypes.POINTER(vp), ctypes.POINTER(u32)])
_sig("mxco_hll_clear", [vp, vp])
_sig("mxco_hll_add", [vp, vp, u64, vp])
_sig("mxco_hll_merge", [vp, vp, vp])
_sig("mxco_hll_histogram", [vp, vp, ctypes.POINTER(u32)])
_sig("mxco_hll_estimate", [vp, vp, pu64])


def _dev_view(ptr, nbytes, dtype=torch.uint8):
    """Device copy of `nbytes` at device address `ptr`."""
    out = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    if nbytes:
        check(0 if _hip.hipMemcpy(vp(out.data_ptr()), vp(ptr), ctypes.c_size_t(nbytes), 3) == 0 else RUNTIME_ERROR)
    return out.view(dtype)


class BloomFilter:
    """aclco::BloomFilter<Key, Extent, BloomFilterPolicy<Key, xxhash_64<Key>,
    uint32_t, W, P>> (include/bloom_filter.h:37, bloom_filter_policy.h:54):
    the default policy is W = P = 8.  `storage`: a caller-owned device tensor
    of at least num_blocks * W int32 words to hold the filter (the
    reference's Allocator), kept alive by this object."""

    def __init__(self, num_blocks, key_dtype=torch.int64, seed=0, words_per_block=8, pattern_bits=None,
                 storage=None, stream=None):
        self.key_dtype = key_dtype
        self.seed = seed
        self.words_per_block = words_per_block
        self.pattern_bits = words_per_block if pattern_bits is None else pattern_bits
        self._storage = storage
        if storage is not None and storage.numel() * storage.element_size() < num_blocks * words_per_block * 4:
            raise ValueError("BloomFilter: storage holds fewer than num_blocks * words_per_block words")
        h = vp()
        pol = BloomPolicy(words_per_block, self.pattern_bits, seed)
        check(_lib.mxco_bloom_create_with_policy(num_blocks, torch.empty(0, dtype=key_dtype).element_size(),
                                                 ctypes.byref(pol), _ptr(storage), _stream(stream), ctypes.byref(h)))
        self._h = h
        self._nb = num_blocks

    def __del__(self, _release=_release):
        _release(self, "mxco_bloom_destroy")

    def block_extent(self):
        return self._nb

    def num_words(self):
        return self._nb * self.words_per_block

    def size_bytes(self):
        return self.num_words() * 4

    def data(self):
        p, nb = vp(), u64()
        check(_lib.mxco_bloom_data(self._h, ctypes.byref(p), ctypes.byref(nb)))
        return p.value

    def words(self):
        """Device copy of the filter words (uint32 viewed as int32)."""
        return _dev_view(self.data(), self.size_bytes(), torch.int32)

    def clear(self, stream=None, sync=True):
        check(_lib.mxco_bloom_clear(self._h, _stream(stream), int(sync)))

    def add(self, keys, n=None, stream=None, sync=True):
        n = keys.shape[0] if n is None else n
        check(_lib.mxco_bloom_add(self._h, _ptr(keys), None, n, _stream(stream), int(sync)))

    def add_async(self, keys, n=None, stream=None):
        self.add(keys, n, stream, sync=False)

    def add_noalloc(self, keys, n=None, stream=None, sync=True, stencil=None):
        """Add through the entry point that allocates nothing (at most
        4,194,304 keys per call; calls on one GPU share its scratch)."""
        n = keys.shape[0] if n is None else n
        check(_lib.mxco_bloom_add_noalloc(self._h, _ptr(keys), _ptr(stencil), n, _stream(stream), int(sync)))

    def add_if(self, keys, stencil, n=None, stream=None, sync=True):
        n = keys.shape[0] if n is None else n
        if n and stencil is None:
            raise ValueError("BloomFilter::AddIf: stencil must not be null when keyNum is non-zero")
        check(_lib.mxco_bloom_add(self._h, _ptr(keys), _ptr(stencil), n, _stream(stream), int(sync)))

    def contains(self, keys, out=None, n=None, stream=None, sync=True, stencil=None):
        n = keys.shape[0] if n is None else n
        if out is None and n:
            out = torch.empty(n, dtype=torch.bool, device="cuda")
        check(_lib.mxco_bloom_contains(self._h, _ptr(keys), _ptr(stencil), _ptr(out), n, _stream(stream),
                                       int(sync)))
        return out

    def contains_if(self, keys, stencil, out=None, n=None, stream=None, sync=True):
        n = keys.shape[0] if n is None else n
        if n and stencil is None:
            raise ValueError("BloomFilter::ContainsIfAsync: stencil must not be null when keyNum is non-zero")
        return self.contains(keys, out, n, stream, sync, stencil)

    def merge(self, other, stream=None, sync=True):
        check(_lib.mxco_bloom_merge(self._h, other._h, _stream(stream), int(sync)))

    def intersect(self, other, stream=None, sync=True):
        check(_lib.mxco_bloom_intersect(self._h, other._h, _stream(stream), int(sync)))


class HyperLogLog:
    """aclco::HyperLogLog<Key> (include/hyperloglog.h:24-59)."""

    def __init__(self, precision, key_dtype=torch.int64, stream=None):
        self.key_dtype = key_dtype
        h = vp()
        check(_lib.mxco_hll_create(precision, torch.empty(0, dtype=key_dtype).element_size(), _stream(stream),
                                   ctypes.byref(h)))
        self._h = h
        self.precision = precision

    @classmethod
    def create_with_precision(cls, p, key_dtype=torch.int64, stream=None):
        return cls(p, key_dtype, stream)

    @classmethod
    def create_with_sketch_size_kb(cls, kb, key_dtype=torch.int64, stream=None):
        p = u32()
        check(_lib.mxco_hll_precision_for_sketch_kb(kb, ctypes.byref(p)))
        return cls(p.value, key_dtype, stream)

    @classmethod
    def create_with_standard_deviation(cls, sd, key_dtype=torch.int64, stream=None):
        p = u32()
        check(_lib.mxco_hll_precision_for_stddev(sd, ctypes.byref(p)))
        return cls(p.value, key_dtype, stream)

    def __del__(self, _release=_release):
        _release(self, "mxco_hll_destroy")

    def sketch_size_kb(self):
        return (1 << self.precision) // 1024

    def data(self):
        p, pr = vp(), u32()
        check(_lib.mxco_hll_data(self._h, ctypes.byref(p), ctypes.byref(pr)))
        return p.value

    def registers(self):
        return _dev_view(self.data(), 1 << self.precision)

    def clear(self, stream=None):
        check(_lib.mxco_hll_clear(self._h, _stream(stream)))

    def add(self, keys, n=None, stream=None):
        n = keys.shape[0] if n is None else n
        check(_lib.mxco_hll_add(self._h, _ptr(keys), n, _stream(stream)))

    def merge(self, other, stream=None):
        check(_lib.mxco_hll_merge(self._h, other._h, _stream(stream)))

    def histogram(self, stream=None):
        h = (u32 * 65)()
        check(_lib.mxco_hll_histogram(self._h, _stream(stream), h))
        return list(h)

    def estimate(self, stream=None):
        e = u64()
        check(_lib.mxco_hll_estimate(self._h, _stream(stream), ctypes.byref(e)))
        return e.value


# ---------------------------------------------------------------------------
# StaticMultimap
# ---------------------------------------------------------------------------

_sig("mxco_mm_create", [u32, u32, u64, vp, vp, vp, ctypes.POINTER(vp)])
_sig("mxco_mm_destroy", [vp])
_sig("mxco_mm_capacity", [vp, pu64])
_sig("mxco_mm_size", [vp, vp, pu64])
_sig("mxco_mm_clear", [vp, vp])
_sig("mxco_mm_insert", [vp, vp, ctypes.POINTER(Stencil), u64, vp, pu64])
_sig("mxco_mm_contains", [vp, vp, ctypes.POINTER(Stencil), vp, u64, vp])
_sig("mxco_mm_find", [vp, vp, ctypes.POINTER(Stencil), vp, u64, vp])
_sig("mxco_mm_count", [vp, vp, u64, vp, pu64])
_sig("mxco_mm_retrieve", [vp, vp, u64, vp, vp, u64, vp, pu64])
_sig("mxco_mm_retrieve_all", [vp, vp, vp, u64, vp, pu64])


class StaticMultimap:
    """aclco::StaticMultimap<Key, Value> (include/static_multimap.h:88)."""

    def __init__(self, capacity, empty_key, empty_value, key_dtype=torch.int64, value_dtype=torch.int64,
                 stream=None):
        self.key_dtype, self.value_dtype = key_dtype, value_dtype
        self._ek = _scalar_bytes(empty_key, key_dtype)
        self._ev = _scalar_bytes(empty_value, value_dtype)
        self.empty_key, self.empty_value = empty_key, empty_value
        h = vp()
        check(_lib.mxco_mm_create(torch.empty(0, dtype=key_dtype).element_size(),
                                  torch.empty(0, dtype=value_dtype).element_size(), capacity, self._ek, self._ev,
                                  _stream(stream), ctypes.byref(h)))
        self._h = h

    def __del__(self, _release=_release):
        _release(self, "mxco_mm_destroy")

    def capacity(self):
        c = u64()
        check(_lib.mxco_mm_capacity(self._h, ctypes.byref(c)))
        return c.value

    def size(self, stream=None):
        c = u64()
        check(_lib.mxco_mm_size(self._h, _stream(stream), ctypes.byref(c)))
        return c.value

    def clear(self, stream=None):
        check(_lib.mxco_mm_clear(self._h, _stream(stream)))

    def insert(self, pairs, n=None, stream=None):
        n = pairs.shape[0] if n is None else n
        f = u64()
        check(_lib.mxco_mm_insert(self._h, _ptr(pairs), None, n, _stream(stream), ctypes.byref(f)))
        return f.value

    def insert_if(self, pairs, stencil, pred, n=None, stream=None):
        n = pairs.shape[0] if n is None else n
        f = u64()
        if n and stencil is None:
            raise ValueError("StaticMultimap: null nonempty buffer")
        check(_lib.mxco_mm_insert(self._h, _ptr(pairs), _stencil(stencil, pred), n, _stream(stream),
                                  ctypes.byref(f)))
        return f.value

    def contains(self, keys, out=None, n=None, stream=None, stencil=None, pred="always"):
        n = keys.shape[0] if n is None else n
        if out is None and n:
            out = torch.empty(n, dtype=torch.bool, device="cuda")
        check(_lib.mxco_mm_contains(self._h, _ptr(keys), _stencil(stencil, pred), _ptr(out), n, _stream(stream)))
        return out

    def find(self, keys, out=None, n=None, stream=None, stencil=None, pred="always"):
        n = keys.shape[0] if n is None else n
        if out is None and n:
            out = torch.empty(n, dtype=self.value_dtype, device="cuda")
        check(_lib.mxco_mm_find(self._h, _ptr(keys), _stencil(stencil, pred), _ptr(out), n, _stream(stream)))
        return out

    def count(self, keys, n=None, stream=None):
        n = keys.shape[0] if n is None else n
        c = u64()
        check(_lib.mxco_mm_count(self._h, _ptr(keys), n, _stream(stream), ctypes.byref(c)))
        return c.value

    def retrieve(self, keys, probe_out, value_out, output_capacity=None, n=None, stream=None):
        n = keys.shape[0] if n is None else n
        cap = value_out.shape[0] if output_capacity is None else output_capacity
        c = u64()
        check(_lib.mxco_mm_retrieve(self._h, _ptr(keys), n, _ptr(probe_out), _ptr(value_out), cap, _stream(stream),
                                    ctypes.byref(c)))
        return c.value

    def retrieve_all(self, key_out, value_out, output_capacity=None, stream=None):
        cap = value_out.shape[0] if output_capacity is None else output_capacity
        c = u64()
        check(_lib.mxco_mm_retrieve_all(self._h, _ptr(key_out), _ptr(value_out), cap, _stream(stream),
                                        ctypes.byref(c)))
        return c.value


# ---------------------------------------------------------------------------
# RoaringBitmap
# ---------------------------------------------------------------------------

_sig("mxco_rb_create", [u32, vp, ctypes.c_size_t, vp, ctypes.POINTER(vp)])
_sig("mxco_rb_destroy", [vp])
class RbView(ctypes.Structure):
    """mxco_rb_view: device copy of the serialized bytes and the search image."""
    _fields_ = [("bytes", vp), ("buckets", vp), ("dense", vp), ("container_keys", vp), ("container_descs", vp),
                ("image", vp), ("num_buckets", u32), ("reserved", u32), ("num_keys", u64), ("size_bytes", u64)]


_sig("mxco_rb_info", [vp, ctypes.POINTER(RbView)])
_sig("mxco_rb_contains", [vp, vp, vp, u64, vp, i32])


class RoaringBitmap:
    """aclco::RoaringBitmap<Key> (include/roaring_bitmap.h:24): `bitmap` is
    the portable serialization (bytes); key_dtype int32 (uint32 keys) or int64
    (uint64 keys)."""

    def __init__(self, bitmap, key_dtype=torch.int32, stream=None):
        self.key_dtype = key_dtype
        self._bytes = bytes(bitmap)
        buf = ctypes.create_string_buffer(self._bytes, len(self._bytes))
        h = vp()
        check(_lib.mxco_rb_create(torch.empty(0, dtype=key_dtype).element_size(), buf, len(self._bytes),
                                  _stream(stream), ctypes.byref(h)))
        self._h = h

    def __del__(self, _release=_release):
        _release(self, "mxco_rb_destroy")

    def view(self):
        v = RbView()
        check(_lib.mxco_rb_info(self._h, ctypes.byref(v)))
        return v

    def size(self):
        return self.view().num_keys

    def size_bytes(self):
        return self.view().size_bytes

    def contains(self, keys, out=None, n=None, stream=None, sync=True):
        n = keys.shape[0] if n is None else n
        if out is None:
            out = torch.empty(n, dtype=torch.bool, device="cuda")
        check(_lib.mxco_rb_contains(self._h, _ptr(keys), _ptr(out), n, _stream(stream), int(sync)))
        return out


# ---------------------------------------------------------------------------
# DynamicMap
# ---------------------------------------------------------------------------

DM_EXACT_DEDUP, DM_APPEND_UNIQUE = 0, 1
_sig("mxco_dm_create", [ctypes.POINTER(OaDesc), u64, vp, vp, vp, vp, ctypes.POINTER(vp)])
_sig("mxco_dm_destroy", [vp])
_sig("mxco_dm_info", [vp, pu64, pu64, pu64])
_sig("mxco_dm_reserve", [vp, u64, vp])
_sig("mxco_dm_insert", [vp, vp, u64, i32, i32, vp, pu64])
_sig("mxco_dm_insert_or_assign", [vp, vp, u64, vp, pu64])
_sig("mxco_dm_find", [vp, vp, vp, u64, vp])
_sig("mxco_dm_contains", [vp, vp, vp, u64, vp])
_sig("mxco_dm_erase", [vp, vp, u64, i32, vp, pu64])
_sig("mxco_dm_clear", [vp, vp])
_sig("mxco_dm_submap", [vp, u32, ctypes.POINTER(vp)])


class DynamicMap:
    """aclco::DynamicMap<Key, T> (include/dynamic_map.h:26): StaticMap
    submaps (LinearProbing<murmurhash3_32>, bucket 5) grown on demand."""

    def __init__(self, initial_capacity, empty_key, empty_value, erased_key, key_dtype=torch.int64,
                 value_dtype=torch.int64, bucket_size=5, probing="linear", hash="murmurhash3_32", stream=None):
        self.key_dtype, self.value_dtype = key_dtype, value_dtype
        ks = torch.empty(0, dtype=key_dtype).element_size()
        vs = torch.empty(0, dtype=value_dtype).element_size()
        d = OaDesc(ks, vs, 1 if key_dtype.is_floating_point else 0, bucket_size, PROBING[probing], HASH[hash], 0,
                   1 if value_dtype.is_floating_point else 0)
        self._ek = _scalar_bytes(empty_key, key_dtype)
        self._ev = _scalar_bytes(empty_value, value_dtype)
        self._rk = _scalar_bytes(erased_key, key_dtype)
        h = vp()
        check(_lib.mxco_dm_create(ctypes.byref(d), initial_capacity, self._ek, self._ev, self._rk, _stream(stream),
                                  ctypes.byref(h)))
        self._h = h

    def __del__(self, _release=_release):
        _release(self, "mxco_dm_destroy")

    def _info(self):
        a, b, c = u64(), u64(), u64()
        check(_lib.mxco_dm_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def size(self):
        return self._info()[0]

    def capacity(self):
        return self._info()[1]

    def num_submaps(self):
        return self._info()[2]

    def reserve(self, n, stream=None):
        check(_lib.mxco_dm_reserve(self._h, n, _stream(stream)))

    def insert(self, pairs, n=None, stream=None, mode=DM_EXACT_DEDUP):
        n = (0 if pairs is None else pairs.shape[0]) if n is None else n
        c = u64()
        check(_lib.mxco_dm_insert(self._h, _ptr(pairs), n, mode, 1, _stream(stream), ctypes.byref(c)))
        return c.value

    def insert_async(self, pairs, n=None, stream=None):
        n = pairs.shape[0] if n is None else n
        c = u64()
        check(_lib.mxco_dm_insert(self._h, _ptr(pairs), n, DM_EXACT_DEDUP, 0, _stream(stream), ctypes.byref(c)))
        return c.value

    def insert_or_assign(self, pairs, n=None, stream=None):
        n = pairs.shape[0] if n is None else n
        c = u64()
        check(_lib.mxco_dm_insert_or_assign(self._h, _ptr(pairs), n, _stream(stream), ctypes.byref(c)))
        return c.value

    def find(self, keys, out=None, n=None, stream=None):
        n = keys.shape[0] if n is None else n
        if out is None and n:
            out = torch.empty(n, dtype=self.value_dtype, device="cuda")
        check(_lib.mxco_dm_find(self._h, _ptr(keys), _ptr(out), n, _stream(stream)))
        return out

    def contains(self, keys, out=None, n=None, stream=None):
        n = keys.shape[0] if n is None else n
        if out is None and n:
            out = torch.empty(n, dtype=torch.bool, device="cuda")
        check(_lib.mxco_dm_contains(self._h, _ptr(keys), _ptr(out), n, _stream(stream)))
        return out

    def erase(self, keys, n=None, stream=None):
        n = keys.shape[0] if n is None else n
        c = u64()
        check(_lib.mxco_dm_erase(self._h, _ptr(keys), n, 1, _stream(stream), ctypes.byref(c)))
        return c.value

    def erase_async(self, keys, n=None, stream=None):
        n = keys.shape[0] if n is None else n
        check(_lib.mxco_dm_erase(self._h, _ptr(keys), n, 0, _stream(stream), None))

    def clear(self, stream=None):
        check(_lib.mxco_dm_clear(self._h, _stream(stream)))


# ---------------------------------------------------------------------------
# StaticMultiset
# ---------------------------------------------------------------------------

MS_CONTAINS, MS_FIND, MS_COUNT_EACH, MS_COUNT_EACH_OUTER = 0, 1, 2, 3
_sig("mxco_ms_create", [u32, u64, vp, vp, ctypes.POINTER(vp)])
_sig("mxco_ms_destroy", [vp])
_sig("mxco_ms_capacity", [vp, pu64])
_sig("mxco_ms_size", [vp, vp, pu64])
_sig("mxco_ms_clear", [vp, vp])
_sig("mxco_ms_insert", [vp, vp, ctypes.POINTER(Stencil), u64, vp, pu64])
_sig("mxco_ms_query", [vp, i32, vp, ctypes.POINTER(Stencil), vp, u64, vp])
_sig("mxco_ms_count", [vp, vp, u64, vp, pu64])
_sig("mxco_ms_retrieve", [vp, vp, u64, vp, vp, u64, vp, pu64])
_sig("mxco_ms_retrieve_all", [vp, vp, u64, vp, pu64])


class StaticMultiset:
    """aclco::StaticMultiset<Key> (include/static_multiset.h:36): int32 or
    int64 keys with exact multiplicities, kept keys in insertion order."""

    def __init__(self, capacity, empty_key, key_dtype=torch.int64, stream=None):
        self.key_dtype = key_dtype
        self.empty_key = empty_key
        self._ek = _scalar_bytes(empty_key, key_dtype)
        h = vp()
        check(_lib.mxco_ms_create(torch.empty(0, dtype=key_dtype).element_size(), capacity, self._ek,
                                  _stream(stream), ctypes.byref(h)))
        self._h = h

    def __del__(self, _release=_release):
        _release(self, "mxco_ms_destroy")

    def capacity(self):
        c = u64()
        check(_lib.mxco_ms_capacity(self._h, ctypes.byref(c)))
        return c.value

    def size(self, stream=None):
        c = u64()
        check(_lib.mxco_ms_size(self._h, _stream(stream), ctypes.byref(c)))
        return c.value

    def clear(self, stream=None):
        check(_lib.mxco_ms_clear(self._h, _stream(stream)))

    def insert(self, keys, n=None, stream=None):
        n = (0 if keys is None else keys.shape[0]) if n is None else n
        f = u64()
        check(_lib.mxco_ms_insert(self._h, _ptr(keys), None, n, _stream(stream), ctypes.byref(f)))
        return f.value

    def insert_if(self, keys, stencil, pred, n=None, stream=None):
        n = keys.shape[0] if n is None else n
        if n and keys is not None and stencil is None:
            raise ValueError("StaticMultiset: null pointer for nonempty operation")
        f = u64()
        check(_lib.mxco_ms_insert(self._h, _ptr(keys), _stencil(stencil, pred), n, _stream(stream),
                                  ctypes.byref(f)))
        return f.value

    def _query(self, kind, keys, out, dtype, n, stream, stencil, pred):
        n = keys.shape[0] if n is None else n
        if out is None and n:
            out = torch.empty(n, dtype=dtype, device="cuda")
        check(_lib.mxco_ms_query(self._h, kind, _ptr(keys), _stencil(stencil, pred), _ptr(out), n,
                                 _stream(stream)))
        return out

    def contains(self, keys, out=None, n=None, stream=None, stencil=None, pred="always"):
        return self._query(MS_CONTAINS, keys, out, torch.bool, n, stream, stencil, pred)

    def find(self, keys, out=None, n=None, stream=None, stencil=None, pred="always"):
        return self._query(MS_FIND, keys, out, self.key_dtype, n, stream, stencil, pred)

    def count_each(self, keys, out=None, n=None, stream=None):
        """uint64 multiplicities (returned as int64)."""
        return self._query(MS_COUNT_EACH, keys, out, torch.int64, n, stream, None, "always")

    def count_each_outer(self, keys, out=None, n=None, stream=None):
        return self._query(MS_COUNT_EACH_OUTER, keys, out, torch.int64, n, stream, None, "always")

    def count(self, keys, n=None, stream=None):
        n = (0 if keys is None else keys.shape[0]) if n is None else n
        c = u64()
        check(_lib.mxco_ms_count(self._h, _ptr(keys), n, _stream(stream), ctypes.byref(c)))
        return c.value

    def retrieve(self, keys, probe_out, match_out, output_capacity=None, n=None, stream=None):
        n = keys.shape[0] if n is None else n
        cap = (0 if match_out is None else match_out.shape[0]) if output_capacity is None else output_capacity
        c = u64()
        check(_lib.mxco_ms_retrieve(self._h, _ptr(keys), n, _ptr(probe_out), _ptr(match_out), cap,
                                    _stream(stream), ctypes.byref(c)))
        return c.value

    def retrieve_all(self, out, output_capacity=None, stream=None):
        cap = (0 if out is None else out.shape[0]) if output_capacity is None else output_capacity
        c = u64()
        check(_lib.mxco_ms_retrieve_all(self._h, _ptr(out), cap, _stream(stream), ctypes.byref(c)))
        return c.value


# ---------------------------------------------------------------------------
# Batched tables, graphs
# ---------------------------------------------------------------------------

class BatchItem(ctypes.Structure):
    _fields_ = [("table", u32), ("reserved", u32), ("offset", u64), ("count", u64)]


BATCH_INSERT, BATCH_FIND, BATCH_CONTAINS, BATCH_BUILD = 0, 1, 2, 3  # BUILD: Clear of every table + Insert
_sig("mxco_oa_insert_path", [vp, u64, vp, ctypes.POINTER(i32)])
_sig("mxco_batch_create", [ctypes.POINTER(OaDesc), u32, ctypes.POINTER(u64), vp, vp, vp, ctypes.POINTER(vp)])
_sig("mxco_batch_destroy", [vp])
_sig("mxco_batch_table", [vp, u32, pu64, ctypes.POINTER(vp)])
_sig("mxco_batch_clear", [vp, vp])
_sig("mxco_batch_plan_create", [vp, i32, ctypes.POINTER(BatchItem), u32, u32, ctypes.POINTER(vp)])
_sig("mxco_batch_plan_destroy", [vp])
_sig("mxco_batch_run", [vp, vp, vp, vp, vp])
_sig("mxco_batch_failures", [vp, vp, pu64])
_sig("mxco_batch_plan_tune", [vp, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_float)])
_sig("mxco_batch_plan_config", [vp, ctypes.POINTER(u32), ctypes.POINTER(u32)])
_sig("mxco_batch_plan_set_config", [vp, u32, u32])
_sig("mxco_batch_plan_launch_shape", [vp, ctypes.POINTER(u32), ctypes.POINTER(u32)])
_sig("mxco_batch_plan_set_launch_shape", [vp, u32, u32])


class BatchPlan:
    def __init__(self, batch, op, items, group_size=0):
        self.batch = batch
        self.op = op
        arr = (BatchItem * len(items))()
        for j, (t, off, cnt) in enumerate(items):
            arr[j].table, arr[j].offset, arr[j].count = t, off, cnt
        self.nitems = len(items)
        h = vp()
        check(_lib.mxco_batch_plan_create(batch._h, op, arr, len(items), group_size, ctypes.byref(h)))
        self._h = h

    def __del__(self, _release=_release):
        _release(self, "mxco_batch_plan_destroy")

    def run(self, inp, out=None, out2=None, stream=None):
        check(_lib.mxco_batch_run(self._h, _ptr(inp), _ptr(out), _ptr(out2), _stream(stream)))

    def failures(self, stream=None):
        f = (u64 * self.nitems)()
        check(_lib.mxco_batch_failures(self._h, _stream(stream), f))
        return list(f)

    def tune(self, inp, out=None, out2=None, stream=None):
        ms = ctypes.c_float()
        check(_lib.mxco_batch_plan_tune(self._h, _ptr(inp), _ptr(out), _ptr(out2), _stream(stream),
                                        ctypes.byref(ms)))
        return ms.value

    def config(self):
        """(lanes per key, workgroups per CU)."""
        g, b = u32(), u32()
        check(_lib.mxco_batch_plan_config(self._h, ctypes.byref(g), ctypes.byref(b)))
        return g.value, b.value

    def set_config(self, group_size, blocks_per_cu):
        check(_lib.mxco_batch_plan_set_config(self._h, group_size, blocks_per_cu))

    def launch_shape(self):
        """(threads per workgroup, wave-staged Find / Contains kernel)."""
        blk, st = u32(), u32()
        check(_lib.mxco_batch_plan_launch_shape(self._h, ctypes.byref(blk), ctypes.byref(st)))
        return blk.value, bool(st.value)

    def set_launch_shape(self, block_size, wave_staged):
        check(_lib.mxco_batch_plan_set_launch_shape(self._h, block_size, int(wave_staged)))


class BatchedStaticMap:
    """Many independent StaticMap tables in one arena, driven by plans."""

    def __init__(self, capacities, empty_key, empty_value, key_dtype=torch.int32, value_dtype=torch.int32,
                 bucket_size=5, probing="linear", hash="murmurhash3_32", stream=None):
        self.key_dtype, self.value_dtype = key_dtype, value_dtype
        ks = torch.empty(0, dtype=key_dtype).element_size()
        vs = 0 if value_dtype is None else torch.empty(0, dtype=value_dtype).element_size()
        d = OaDesc(ks, vs, 1 if key_dtype.is_floating_point else 0, bucket_size, PROBING[probing], HASH[hash], 0, 0)
        self._ek = _scalar_bytes(empty_key, key_dtype)
        self._ev = _scalar_bytes(empty_value, value_dtype) if value_dtype is not None else None
        self.slot_size = ks if not vs else (8 if ks == 4 and vs == 4 else 16)
        caps = (u64 * len(capacities))(*capacities)
        h = vp()
        check(_lib.mxco_batch_create(ctypes.byref(d), len(capacities), caps, self._ek, self._ev, _stream(stream),
                                     ctypes.byref(h)))
        self._h = h
        self.num_tables = len(capacities)

    def __del__(self, _release=_release):
        _release(self, "mxco_batch_destroy")

    def capacity(self, t):
        c = u64()
        check(_lib.mxco_batch_table(self._h, t, ctypes.byref(c), None))
        return c.value

    def slots(self, t):
        """Device copy of table t's raw slot array as uint8 (capacity, slot_size)."""
        c, p = u64(), vp()
        check(_lib.mxco_batch_table(self._h, t, ctypes.byref(c), ctypes.byref(p)))
        return _dev_view(p.value, c.value * self.slot_size).view(c.value, self.slot_size)

    def clear(self, stream=None):
        check(_lib.mxco_batch_clear(self._h, _stream(stream)))

    def plan(self, op, items, group_size=0):
        return BatchPlan(self, op, items, group_size)


_hip.hipStreamBeginCapture.argtypes = [vp, i32]
_hip.hipStreamEndCapture.argtypes = [vp, ctypes.POINTER(vp)]
_hip.hipGraphInstantiate.argtypes = [ctypes.POINTER(vp), vp, vp, vp, ctypes.c_size_t]
_hip.hipGraphLaunch.argtypes = [vp, vp]
_hip.hipGraphExecDestroy.argtypes = [vp]
_hip.hipGraphDestroy.argtypes = [vp]


class Graph:
    """Records the library calls made by `fn(stream)` on a private stream into
    a hipGraph (thread-local capture) and replays them with launch()."""

    def __init__(self, fn):
        self.stream = torch.cuda.Stream()
        torch.cuda.synchronize()
        s = vp(self.stream.cuda_stream)
        if _hip.hipStreamBeginCapture(s, 1) != 0:
            raise RuntimeError("hipStreamBeginCapture failed")
        try:
            fn(self.stream)
        finally:
            g = vp()
            rc = _hip.hipStreamEndCapture(s, ctypes.byref(g))
        if rc != 0:
            raise RuntimeError("hipStreamEndCapture failed")
        self._graph = g
        e = vp()
        if _hip.hipGraphInstantiate(ctypes.byref(e), g, None, None, 0) != 0:
            raise RuntimeError("hipGraphInstantiate failed")
        self._exec = e

    def launch(self, stream=None):
        if _hip.hipGraphLaunch(self._exec, vp(_stream(stream if stream is not None else self.stream))) != 0:
            raise RuntimeError("hipGraphLaunch failed")

    def __del__(self):
        ex, self._exec = getattr(self, "_exec", None), None
        try:
            if ex:
                _hip.hipGraphExecDestroy(ex)
                _hip.hipGraphDestroy(self._graph)
        except (NameError, AttributeError, TypeError):
            pass  # interpreter shutdown: the HIP runtime handle is gone


# ---------------------------------------------------------------------------
# Serving engine
# ---------------------------------------------------------------------------

class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_reserve_bytes", u64), ("hbm_budget_bytes", u64), ("max_tick_keys", u64),
                ("group_size", u32), ("blocks_per_cu", u32), ("tick_mode", u32), ("resident_blocks_per_cu", u32),
                ("empty_key", ctypes.c_uint8 * 8), ("empty_value", ctypes.c_uint8 * 8)]


class Request(ctypes.Structure):
    _fields_ = [("table", u32), ("op", u32), ("input", vp), ("n", u64), ("out", vp), ("out2", vp),
                ("out_capacity", u64)]


REQ_INSERT, REQ_FIND, REQ_CONTAINS, REQ_MM_INSERT, REQ_MM_FIND, REQ_MM_COUNT, REQ_MM_RETRIEVE = range(7)
_sig("mxco_server_create", [ctypes.POINTER(ServerConfig), ctypes.POINTER(OaDesc), ctypes.POINTER(vp)])
_sig("mxco_server_destroy", [vp])
_sig("mxco_server_add_map", [vp, u64, ctypes.POINTER(u32)])
_sig("mxco_server_add_multimap", [vp, u64, vp, vp, ctypes.POINTER(u32)])
_sig("mxco_server_drop_table", [vp, u32])
_sig("mxco_server_submit", [vp, ctypes.POINTER(Request), pu64])
_sig("mxco_server_wait", [vp, u64, pu64, ctypes.POINTER(ctypes.c_double)])
_sig("mxco_server_drain", [vp])
_sig("mxco_server_replay", [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_double), u32, pu64,
                            ctypes.POINTER(ctypes.c_double)])
_sig("mxco_server_replay_plain", [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_double), u32, pu64,
                                  ctypes.POINTER(ctypes.c_double)])
_sig("mxco_server_stats", [vp, pu64, pu64])
_sig("mxco_server_resident_stats", [vp, ctypes.POINTER(u32), pu64, ctypes.POINTER(u32)])
_sig("mxco_server_resident_trace", [vp, pu64, u32, ctypes.POINTER(u32), ctypes.POINTER(ctypes.c_double)])


class Server:
    """Resident int64 StaticMap / StaticMultimap tables behind a continuously
    batching request loop (native dispatcher, completion and multimap
    threads)."""

    def __init__(self, empty_key=-1, empty_value=-1, key_dtype=torch.int64, value_dtype=torch.int64,
                 hbm_reserve_bytes=1 << 30, hbm_budget_bytes=0, max_tick_keys=0, group_size=0, blocks_per_cu=0,
                 resident=True, resident_blocks_per_cu=0):
        cfg = ServerConfig()
        cfg.hbm_reserve_bytes, cfg.hbm_budget_bytes, cfg.max_tick_keys = hbm_reserve_bytes, hbm_budget_bytes, \
            max_tick_keys
        cfg.group_size, cfg.blocks_per_cu = group_size, blocks_per_cu
        cfg.tick_mode, cfg.resident_blocks_per_cu = (0 if resident else 1), resident_blocks_per_cu
        ek = _scalar_bytes(empty_key, key_dtype).raw
        ev = _scalar_bytes(empty_value, value_dtype).raw
        for j in range(len(ek)):
            cfg.empty_key[j] = ek[j]
        for j in range(len(ev)):
            cfg.empty_value[j] = ev[j]
        ks = torch.empty(0, dtype=key_dtype).element_size()
        vs = torch.empty(0, dtype=value_dtype).element_size()
        d = OaDesc(ks, vs, 0, 5, 0, 0, 0, 0)
        h = vp()
        check(_lib.mxco_server_create(ctypes.byref(cfg), ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self._mm_ek = _scalar_bytes(empty_key, torch.int64)
        self._mm_ev = _scalar_bytes(empty_value, torch.int64)

    def close(self):
        if getattr(self, "_h", None):
            check(_lib.mxco_server_destroy(self._h))
            self._h = None

    def __del__(self, _release=_release):
        _release(self, "mxco_server_destroy")

    def add_map(self, capacity):
        t = u32()
        check(_lib.mxco_server_add_map(self._h, capacity, ctypes.byref(t)))
        return t.value

    def add_multimap(self, capacity):
        t = u32()
        check(_lib.mxco_server_add_multimap(self._h, capacity, self._mm_ek, self._mm_ev, ctypes.byref(t)))
        return t.value

    def drop_table(self, t):
        check(_lib.mxco_server_drop_table(self._h, t))

    def submit(self, table, op, inp, n=None, out=None, out2=None, out_capacity=0):
        r = Request(table, op, _ptr(inp), inp.shape[0] if n is None else n, _ptr(out), _ptr(out2), out_capacity)
        tk = u64()
        check(_lib.mxco_server_submit(self._h, ctypes.byref(r), ctypes.byref(tk)))
        return tk.value

    def wait(self, ticket):
        res, lat = u64(), ctypes.c_double()
        check(_lib.mxco_server_wait(self._h, ticket, ctypes.byref(res), ctypes.byref(lat)))
        return res.value, lat.value

    def drain(self):
        check(_lib.mxco_server_drain(self._h))

    @staticmethod
    def _requests(reqs):
        arr = (Request * len(reqs))()
        for j, r in enumerate(reqs):
            arr[j] = Request(*r)
        return arr

    def replay(self, reqs, arrivals_us, plain=False):
        """reqs: list of (table, op, in_ptr, n, out_ptr, out2_ptr, out_capacity)."""
        n = len(reqs)
        arr = self._requests(reqs)
        arrivals = (ctypes.c_double * n)(*arrivals_us)
        res = (u64 * n)()
        lat = (ctypes.c_double * n)()
        fn = _lib.mxco_server_replay_plain if plain else _lib.mxco_server_replay
        check(fn(self._h, arr, arrivals, n, res, lat))
        return list(res), list(lat)

    def stats(self):
        a, b = u64(), u64()
        check(_lib.mxco_server_stats(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def resident_trace(self, max_ticks=1024):
        """Timing trace of the last StaticMap ticks served by the resident
        kernel (drains and ends the kernel first), in microseconds: per tick
        `run_us` from opening it to publishing its results; for each tick
        that was staged before the previous one finished, `handoff_us` from
        that finish to its opening; its insert and read chunk counts."""
        import numpy as np
        buf = (u64 * (4 * max_ticks))()
        n, tpu = u32(), ctypes.c_double()
        check(_lib.mxco_server_resident_trace(self._h, buf, max_ticks, ctypes.byref(n), ctypes.byref(tpu)))
        a = np.frombuffer(buf, dtype=np.uint64, count=4 * n.value).reshape(-1, 4).astype(np.float64)
        if not len(a):
            return {}
        staged, opened, done = a[:, 0], a[:, 1], a[:, 2]
        cnt = a[:, 3].astype(np.uint64)
        waiting = staged[1:] <= done[:-1]
        return {"run_us": (done - opened) / tpu.value,
                "handoff_us": ((opened[1:] - done[:-1]) / tpu.value)[waiting],
                "ins_chunks": (cnt >> np.uint64(32)).astype(np.int64),
                "rd_chunks": (cnt & np.uint64(0xffffffff)).astype(np.int64)}

    def resident_stats(self):
        """(StaticMap ticks run in the resident kernel, its launches so far, its grid)."""
        r, n, g = u32(), u64(), u32()
        check(_lib.mxco_server_resident_stats(self._h, ctypes.byref(r), ctypes.byref(n), ctypes.byref(g)))
        return bool(r.value), n.value, g.value


# ---------------------------------------------------------------------------
# Measured ceilings (random reads, streaming copy)
# ---------------------------------------------------------------------------

_sig("mxco_diag_random_read", [vp, u64, u32, u64, vp, u64, vp])
_sig("mxco_diag_stream_copy", [vp, vp, u64, vp])


def random_read(buf, nbytes, access, n, sink, stream=None):
    """n reads of `access` bytes at hashed aligned offsets of buf's first nbytes."""
    check(_lib.mxco_diag_random_read(_ptr(buf), nbytes, access, n, _ptr(sink), sink.numel(), _stream(stream)))


def stream_copy(src, dst, nbytes, stream=None):
    check(_lib.mxco_diag_stream_copy(_ptr(src), _ptr(dst), nbytes, _stream(stream)))
