"""Runs the aqs:: C++ suite (tests/cpp/aqs_tests.cpp, a mirror of the
reference's test/tests.cpp) against libaqs.so / libaqs_hip.so."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "aqs_tests")


def test_cpp_suite_is_built():
    assert os.path.exists(BIN), "build with __graft_entry__.build()"


@pytest.mark.gpu
def test_cpp_reference_suite():
    r = subprocess.run([BIN, "0"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
