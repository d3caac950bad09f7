"""Runs the C++ header-API program (tests/cpp/header_api_test.cpp), built
against libmxco.so by ops-collections_amd/build.py."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ops-collections_amd", "lib", "header_api_test")


def test_header_api_program_built():
    assert os.path.exists(EXE), "run ops-collections_amd/build.py"


@pytest.mark.gpu
def test_header_api_program():
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().endswith("OK")
