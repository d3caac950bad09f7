"""The oracle's pins must reject plausible mistakes, not only accept the
oracle: each case compiles oracle/oracle.c with one planted mistake
(-DORC_PLANT=n) and runs the pinning tests of tests/test_oracle.py that should
catch it against that build, which must fail.  The same selection passes on
the normal build (checked once at the end)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "oracle.c")

# plant id -> (the mistake, the tests of test_oracle.py that must catch it)
PLANTS = {
    1: ("fmix64: first shift 32 instead of 33", "reference_hash_kats or smhasher_verification"),
    2: ("xxh64: length term dropped", "reference_hash_kats"),
    3: ("murmur3_32: length xor dropped", "reference_hash_kats or murmur3_published"),
    4: ("DoubleHashing step: h2 % nB + 1 instead of h2 % (nB-1) + 1", "probe_start_closed_form"),
    5: ("LinearProbing start one bucket off", "probe_start_closed_form"),
    6: ("Insert overwrites the value of an existing key (last-wins)", "map_semantics_vs_dict"),
    7: ("Bloom: salts of neighbouring words swapped", "bloom_vs_bitwise"),
    8: ("HLL: rank without the +1", "hll_registers_closed_form"),
    9: ("MakeValidExtent (linear): floor instead of ceil", "extent_kats"),
    10: ("Pair: second member not naturally aligned", "pair_layout"),
    11: ("HLL++ linear counting: m ln(m/(V+1))", "hll_estimate_linear_counting"),
    12: ("xxh32: stripe rotation 12 instead of 13", "xxhash_against_library"),
    13: ("HLL++ bias: mean over 5 neighbours instead of 6", "hll_bias_regime_nearest_six"),
    14: ("HLL++ bias added instead of subtracted", "hll_bias_regime_nearest_six"),
    15: ("HLL++: the h <= threshold branch dropped", "hll_threshold_branch"),
    16: ("AssignIf inserts absent keys (InsertOrAssign semantics)", "assign_if_vs_dict"),
    17: ("AssignIf walks past an empty slot", "assign_if_stops_at_the_first_empty_slot"),
}


def _build(plant, out_dir):
    so = os.path.join(out_dir, "liboracle_plant%d.so" % plant)
    subprocess.check_call(["gcc", "-O1", "-std=gnu11", "-fPIC", "-shared", "-DORC_PLANT=%d" % plant, "-o", so, SRC,
                           "-lpthread", "-lm"])
    return so


def _run(selection, lib=None):
    env = dict(os.environ)
    if lib:
        env["MXCO_ORACLE_LIB"] = lib
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_oracle.py"), "-q", "-x",
                        "-p", "no:cacheprovider", "-k", selection, "--timeout", "120"],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    return r.returncode, r.stdout[-2000:]


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_is_rejected(plant, tmp_path):
    what, selection = PLANTS[plant]
    rc, out = _run(selection, _build(plant, str(tmp_path)))
    assert rc == 1, "planted mistake %d (%s) passed %r:\n%s" % (plant, what, selection, out)
    assert "failed" in out


def test_unplanted_build_passes_the_same_selection():
    selection = " or ".join("(%s)" % s for _, s in PLANTS.values())
    rc, out = _run(selection)
    assert rc == 0, out
