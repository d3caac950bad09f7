"""CPU checks of the C ABI: the library loads, exports every entry point
include/*.h declares, and its host-only logic (capacity rounding, error
strings) agrees with the oracle.  No kernel is launched here."""
import ctypes
import glob
import os
import re

import numpy as np

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ops-collections_amd", "lib", "libmxco.so")


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        names.update(re.findall(r"MXCO_API\s+[^;(]*?\b(mxco_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build libmxco.so first (ops-collections_amd/build.py)"
    lib = ctypes.CDLL(LIB)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_valid_extent_matches_oracle():
    import mxco
    rng = np.random.default_rng(0)
    caps = [0, 1, 2, 5, 6, 7, 20, 21, 1000, 1 << 20, 1 << 28, (1 << 32) - 5] + rng.integers(
        0, 1 << 34, 200).tolist()
    for cap in caps:
        for b in (1, 2, 3, 5, 8):
            for probing, name in ((O.PROBE_LINEAR, "linear"), (O.PROBE_DOUBLE, "double")):
                assert mxco.valid_extent(int(cap), b, name) == O.valid_extent(int(cap), b, probing), (cap, b, name)


def test_invalid_extent_raises_invalid_argument():
    import mxco
    try:
        mxco.valid_extent(17177758134 * 5, 5, "double")
    except ValueError as e:
        assert "Extent" in str(e)
    else:
        raise AssertionError("expected ValueError")


def test_status_strings():
    lib = ctypes.CDLL(LIB)
    lib.mxco_status_string.restype = ctypes.c_char_p
    assert lib.mxco_status_string(0) == b"ok"
    assert lib.mxco_status_string(2) == b"length error"


def test_create_rejects_capacity_beyond_32_bit_slot_index_before_allocating():
    """The device probe sequence indexes slots with 32 bits: a capacity that
    rounds above 2^32 - 1 fails in mxco_oa_create before any allocation, with
    MXCO_INVALID_ARGUMENT and a message naming the limit."""
    import mxco
    d = mxco.OaDesc(8, 8, 0, 5, 0, 0, 0, 0)
    h = ctypes.c_void_p()
    ek, ev = ctypes.c_int64(-1), ctypes.c_int64(-1)
    st = mxco._lib.mxco_oa_create(ctypes.byref(d), 1 << 33, ctypes.byref(ek), ctypes.byref(ev), None,
                                  ctypes.byref(h))
    assert st == mxco.INVALID_ARGUMENT and h.value is None
    assert b"32-bit slot index" in mxco._lib.mxco_last_error()
