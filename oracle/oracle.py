"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker.  The product
library never loads it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MXCO_ORACLE_LIB selects another build of oracle.c (the planted-mistake
# builds of tests/test_oracle_mutations.py).
_LIB_PATH = os.environ.get("MXCO_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

HASH_MURMUR3_32, HASH_XXHASH_32, HASH_XXHASH_64, HASH_FMIX_32, HASH_FMIX_64 = range(5)
PROBE_LINEAR, PROBE_DOUBLE = 0, 1

u8p = ctypes.POINTER(ctypes.c_uint8)
vp = ctypes.c_void_p


def build():
    """Compile liboracle.so from oracle.c (gcc only)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_fmix32.restype = ctypes.c_uint32
        L.orc_fmix32.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.orc_fmix64.restype = ctypes.c_uint64
        L.orc_fmix64.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.orc_murmur3_32.restype = ctypes.c_uint32
        L.orc_murmur3_32.argtypes = [vp, ctypes.c_size_t, ctypes.c_uint32]
        L.orc_xxh32.restype = ctypes.c_uint32
        L.orc_xxh32.argtypes = [vp, ctypes.c_size_t, ctypes.c_uint32]
        L.orc_xxh64.restype = ctypes.c_uint64
        L.orc_xxh64.argtypes = [vp, ctypes.c_size_t, ctypes.c_uint64]
        L.orc_probe_hash.restype = ctypes.c_uint32
        L.orc_probe_hash.argtypes = [ctypes.c_int, vp, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_primes.restype = ctypes.c_size_t
        L.orc_primes.argtypes = [ctypes.POINTER(ctypes.POINTER(ctypes.c_int64))]
        L.orc_valid_extent.restype = ctypes.c_uint64
        L.orc_valid_extent.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int]
        L.orc_pair_layout.restype = None
        L.orc_pair_layout.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint32)]
        L.orc_oa_create.restype = vp
        L.orc_oa_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_uint64, vp, vp]
        L.orc_oa_destroy.argtypes = [vp]
        L.orc_oa_capacity.restype = ctypes.c_uint64
        L.orc_oa_capacity.argtypes = [vp]
        L.orc_oa_slot_size.restype = ctypes.c_uint32
        L.orc_oa_slot_size.argtypes = [vp]
        L.orc_oa_slots.restype = vp
        L.orc_oa_slots.argtypes = [vp]
        L.orc_oa_probe_start.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
        L.orc_oa_insert.restype = ctypes.c_uint64
        L.orc_oa_insert.argtypes = [vp, vp, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                    ctypes.POINTER(ctypes.c_uint64)]
        L.orc_oa_insert_if.restype = ctypes.c_uint64
        L.orc_oa_insert_if.argtypes = [vp, vp, vp, ctypes.c_uint64]
        L.orc_oa_insert_and_find.argtypes = [vp, vp, ctypes.c_uint64, vp, vp]
        L.orc_oa_assign_if.restype = ctypes.c_uint64
        L.orc_oa_assign_if.argtypes = [vp, vp, vp, ctypes.c_uint64]
        L.orc_oa_find.restype = ctypes.c_uint64
        L.orc_oa_find.argtypes = [vp, vp, ctypes.c_uint64, vp, vp]
        L.orc_oa_erase.restype = ctypes.c_uint64
        L.orc_oa_erase.argtypes = [vp, vp, ctypes.c_uint64, vp]
        L.orc_oa_mt.restype = ctypes.c_uint64
        L.orc_oa_mt.argtypes = [vp, ctypes.c_int, vp, ctypes.c_uint64, vp, vp, ctypes.c_int]
        L.orc_bloom_add.argtypes = [vp, ctypes.c_uint64, vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64]
        L.orc_bloom_contains.argtypes = [vp, ctypes.c_uint64, vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64,
                                         vp]
        L.orc_bloom_add_policy.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, vp, ctypes.c_uint32,
                                           ctypes.c_uint64, ctypes.c_uint64]
        L.orc_bloom_contains_policy.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, vp,
                                                ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, vp]
        L.orc_bloom_salts.restype = ctypes.POINTER(ctypes.c_uint32)
        L.orc_bloom_mt.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, vp, ctypes.c_uint32,
                                   ctypes.c_uint64, ctypes.c_uint64, vp, ctypes.c_int]
        L.orc_hll_add_mt.argtypes = [vp, ctypes.c_uint32, vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int]
        L.orc_hll_add.argtypes = [vp, ctypes.c_uint32, vp, ctypes.c_uint32, ctypes.c_uint64]
        L.orc_hll_histogram.argtypes = [vp, ctypes.c_uint32, vp]
        L.orc_hll_z.restype = ctypes.c_double
        L.orc_hll_z.argtypes = [vp]
        L.orc_hll_finalize.restype = ctypes.c_uint64
        L.orc_hll_finalize.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_int, vp, vp, ctypes.c_int,
                                       ctypes.c_double]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _bytes_of(x, dtype):
    return np.array([x], dtype=dtype).tobytes()


def primes():
    p = ctypes.POINTER(ctypes.c_int64)()
    n = lib().orc_primes(ctypes.byref(p))
    return np.ctypeslib.as_array(p, shape=(n,)).copy()


def valid_extent(cap, bucket, probing):
    return int(lib().orc_valid_extent(cap, bucket, probing))


def pair_layout(ks, vs):
    ss, off = ctypes.c_uint32(), ctypes.c_uint32()
    lib().orc_pair_layout(ks, vs, ctypes.byref(ss), ctypes.byref(off))
    return ss.value, off.value


def pack_pairs(keys, values):
    """Pack keys/values into the Pair<K,V> slot layout (bytes, n x slot)."""
    keys = np.ascontiguousarray(keys)
    n = keys.shape[0]
    if values is None:
        return keys.view(np.uint8).reshape(n, keys.itemsize)
    values = np.ascontiguousarray(values)
    ss, off = pair_layout(keys.itemsize, values.itemsize)
    out = np.zeros((n, ss), np.uint8)
    out[:, :keys.itemsize] = keys.view(np.uint8).reshape(n, keys.itemsize)
    out[:, off:off + values.itemsize] = values.view(np.uint8).reshape(n, values.itemsize)
    return out


class OpenAddressing:
    """Sequential restatement of StaticMap / StaticSet."""

    def __init__(self, key_dtype, value_dtype, capacity, empty_key, empty_value=None, bucket=5,
                 probing=PROBE_LINEAR, hash_family=HASH_MURMUR3_32):
        self.kd = np.dtype(key_dtype)
        self.vd = np.dtype(value_dtype) if value_dtype is not None else None
        ks = self.kd.itemsize
        vs = self.vd.itemsize if self.vd is not None else 0
        ek = _bytes_of(empty_key, self.kd)
        ev = _bytes_of(empty_value, self.vd) if self.vd is not None else b"\0" * 8
        self._ek, self._ev = ek, ev
        self.t = lib().orc_oa_create(ks, vs, bucket, probing, hash_family, int(self.kd.kind == "f"),
                                     capacity, ek, ev)
        if not self.t:
            raise ValueError("invalid capacity")
        self.capacity = int(lib().orc_oa_capacity(self.t))
        self.slot_size = int(lib().orc_oa_slot_size(self.t))
        self.empty_key, self.empty_value = empty_key, empty_value

    def __del__(self):
        if getattr(self, "t", None):
            lib().orc_oa_destroy(self.t)
            self.t = None

    def slots(self):
        buf = (ctypes.c_uint8 * (self.capacity * self.slot_size)).from_address(lib().orc_oa_slots(self.t))
        return np.frombuffer(buf, np.uint8).reshape(self.capacity, self.slot_size).copy()

    def load_slots(self, slots):
        """Overwrite the slot array with `slots` (same layout, e.g. a GPU table copied to host)."""
        b = np.ascontiguousarray(slots).view(np.uint8).reshape(-1)
        assert b.size == self.capacity * self.slot_size
        ctypes.memmove(lib().orc_oa_slots(self.t), b.ctypes.data, b.size)

    def probe_start(self, key):
        """(first slot, step) of the key's probe sequence."""
        k = np.array([key], self.kd)
        st, sp = ctypes.c_uint32(), ctypes.c_uint32()
        lib().orc_oa_probe_start(self.t, _ptr(k), ctypes.byref(st), ctypes.byref(sp))
        return st.value, sp.value

    def _pairs(self, keys, values):
        return np.ascontiguousarray(pack_pairs(keys.astype(self.kd), None if self.vd is None else
                                               values.astype(self.vd)))

    def insert(self, keys, values=None, lines=False):
        p = self._pairs(keys, values)
        ln = ctypes.c_uint64()
        f = lib().orc_oa_insert(self.t, _ptr(p), len(keys), 0, None, ctypes.byref(ln) if lines else None)
        return (int(f), ln.value) if lines else int(f)

    def insert_or_assign(self, keys, values):
        p = self._pairs(keys, values)
        a = ctypes.c_uint64()
        f = lib().orc_oa_insert(self.t, _ptr(p), len(keys), 1, ctypes.byref(a), None)
        return int(f), a.value

    def insert_if(self, keys, values, flags):
        p = self._pairs(keys, values)
        fl = np.ascontiguousarray(flags.astype(np.uint8))
        return int(lib().orc_oa_insert_if(self.t, _ptr(p), _ptr(fl), len(keys)))

    def assign_if(self, keys, values, flags=None):
        """AssignIf: payloads of present keys overwritten, nothing inserted.
        Returns the number of selected pairs whose key was found."""
        p = self._pairs(keys, values)
        fl = None if flags is None else np.ascontiguousarray(np.asarray(flags).astype(np.uint8))
        return int(lib().orc_oa_assign_if(self.t, _ptr(p), _ptr(fl), len(keys)))

    def insert_and_find(self, keys, values=None):
        p = self._pairs(keys, values)
        od = self.vd if self.vd is not None else self.kd
        out = np.zeros(len(keys), od)
        ins = np.zeros(len(keys), np.uint8)
        lib().orc_oa_insert_and_find(self.t, _ptr(p), len(keys), _ptr(out), _ptr(ins))
        return out, ins.astype(bool)

    def find(self, keys, lines=False):
        k = np.ascontiguousarray(keys.astype(self.kd))
        od = self.vd if self.vd is not None else self.kd
        out = np.zeros(len(k), od)
        c = np.zeros(len(k), np.uint8)
        ln = lib().orc_oa_find(self.t, _ptr(k), len(k), _ptr(out), _ptr(c))
        return (out, c.astype(bool), int(ln)) if lines else (out, c.astype(bool))

    def erase(self, keys, erased_key=None):
        k = np.ascontiguousarray(keys.astype(self.kd))
        ek = _bytes_of(erased_key, self.kd) if erased_key is not None else None
        return int(lib().orc_oa_erase(self.t, _ptr(k), len(k), ek))

    def mt_insert(self, keys, values, threads):
        p = self._pairs(keys, values)
        return int(lib().orc_oa_mt(self.t, 0, _ptr(p), len(keys), None, None, threads))

    def mt_find(self, keys, threads):
        k = np.ascontiguousarray(keys.astype(self.kd))
        out = np.zeros(len(k), self.vd)
        lib().orc_oa_mt(self.t, 1, _ptr(k), len(k), _ptr(out), None, threads)
        return out


def bloom_add(words, num_blocks, keys, seed=0, words_per_block=8, pattern_bits=None):
    k = np.ascontiguousarray(keys)
    p = words_per_block if pattern_bits is None else pattern_bits
    lib().orc_bloom_add_policy(_ptr(words), num_blocks, words_per_block, p, _ptr(k), k.itemsize, len(k), seed)


def bloom_contains(words, num_blocks, keys, seed=0, words_per_block=8, pattern_bits=None):
    k = np.ascontiguousarray(keys)
    p = words_per_block if pattern_bits is None else pattern_bits
    out = np.zeros(len(k), np.uint8)
    lib().orc_bloom_contains_policy(_ptr(words), num_blocks, words_per_block, p, _ptr(k), k.itemsize, len(k), seed,
                                    _ptr(out))
    return out.astype(bool)


def bloom_salts():
    return np.ctypeslib.as_array(lib().orc_bloom_salts(), shape=(64,)).copy()


def hll_add(regs, p, keys):
    k = np.ascontiguousarray(keys)
    lib().orc_hll_add(_ptr(regs), p, _ptr(k), k.itemsize, len(k))


def hll_histogram(regs, p):
    h = np.zeros(65, np.uint32)
    lib().orc_hll_histogram(_ptr(regs), p, _ptr(h))
    return h


def bloom_add_mt(words, num_blocks, keys, threads, seed=0, words_per_block=8, pattern_bits=None):
    k = np.ascontiguousarray(keys)
    p = words_per_block if pattern_bits is None else pattern_bits
    lib().orc_bloom_mt(_ptr(words), num_blocks, words_per_block, p, _ptr(k), k.itemsize, len(k), seed, None, threads)


def bloom_contains_mt(words, num_blocks, keys, threads, seed=0, words_per_block=8, pattern_bits=None):
    k = np.ascontiguousarray(keys)
    p = words_per_block if pattern_bits is None else pattern_bits
    out = np.zeros(len(k), np.uint8)
    lib().orc_bloom_mt(_ptr(words), num_blocks, words_per_block, p, _ptr(k), k.itemsize, len(k), seed, _ptr(out),
                       threads)
    return out.astype(bool)


def hll_add_mt(regs, p, keys, threads):
    k = np.ascontiguousarray(keys)
    lib().orc_hll_add_mt(_ptr(regs), p, _ptr(k), k.itemsize, len(k), threads)
