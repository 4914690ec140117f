"""AddressSanitizer and UndefinedBehaviorSanitizer builds of the host-only
code that parses untrusted files or runs the reference loops: the pole-file
reader/writer (vca_amd/csrc/gfprime.cpp) on the golden files, round trips
and damaged or hostile files; the oracle's C loops (oracle/vca_oracle.c)
on two small models.  CPU only (gcc/g++); tests/sanitize/Makefile."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                    reason="no host C++ toolchain")
def test_sanitized_reader_and_oracle(tmp_path):
    # verify_asan_link_order=0: the environment may preload other libraries
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run(["make", "-s", "-C", os.path.join(HERE, "sanitize"), "run",
                        f"OUT={tmp_path}"], capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert out.count("ALL OK") == 2, out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out
