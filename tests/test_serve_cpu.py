"""Host logic of the request server, without a GPU.

tests/cpp/serve_core_test.cpp drives the scheduling core (serve_core.cpp:
queue, admission by footprint against an injected budget, resident-state
eviction, starvation limit, cancellation, waiting) with a stand-in backend,
from many threads at once.  It runs twice: as built, and built with
ThreadSanitizer, where any data race or lock misuse ends the run.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLAIN = os.path.join(ROOT, "tests", "cpp", "serve_core_test")
TSAN = os.path.join(ROOT, "tests", "cpp", "serve_core_test_tsan")


@pytest.mark.parametrize("binary", [PLAIN, TSAN], ids=["plain", "tsan"])
def test_server_core(binary):
    assert os.path.exists(binary), "build with __graft_entry__.build()"
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([binary], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
    assert "ThreadSanitizer" not in r.stderr, r.stderr
