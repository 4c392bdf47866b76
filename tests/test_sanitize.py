"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

`make -C oracle sanitize` (run by __graft_entry__.build()) compiles every
CPU oracle source, the library's host table builders
(damascene_amd/csrc/tables.cpp) and its image generator with
-fsanitize=address,undefined into oracle/_san/sanitize.  The driver builds
every host table of the library and the oracle for several sizes and
rotation angles, requires the two to agree bit for bit, and runs the
oracle's whole gPb pipeline with 32 and 64 textons.  Any out-of-bounds
access, use after free, leak or undefined behaviour aborts it with the
sanitizer's report.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_san", "sanitize")


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_san/sanitize not built (make -C oracle sanitize)")
def test_host_code_under_sanitizers():
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    print(out[-2000:])
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out
    assert "sanitized run: 0 mismatches" in out
