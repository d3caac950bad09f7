"""Runs the C++ header-API programs (tests/cpp/header_api_test.cpp, the
ForEach program tests/cpp/for_each_test.hip, the RoaringBitmap /
RoaringBitmapRef program tests/cpp/roaring_ref_test.hip and the
caller-defined KeyEqual / hasher / predicate program
tests/cpp/caller_functor_test.hip), built against libmxco.so by
ops-collections_amd/build.py."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ops-collections_amd", "lib")
PROGRAMS = ["header_api_test", "for_each_test", "roaring_ref_test", "caller_functor_test"]


@pytest.mark.parametrize("name", PROGRAMS)
def test_header_api_program_built(name):
    assert os.path.exists(os.path.join(LIB, name)), "run ops-collections_amd/build.py"


@pytest.mark.gpu
@pytest.mark.parametrize("name", PROGRAMS)
def test_header_api_program(name):
    r = subprocess.run([os.path.join(LIB, name)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().endswith("OK")
