"""The library's host code and the CPU oracle under AddressSanitizer and
UndefinedBehaviorSanitizer (no GPU).

tests/sanitize/host_check.cpp drives the generators, the BVH builder and
encoder, the OBJ and stream readers (fuzzed, over buffers of exactly the
stated size), the frame arena and the oracle's walks, renderer and cache
gradients; the build (tests/sanitize/Makefile, g++/gcc with
-fsanitize=address,undefined -fno-sanitize-recover=all) makes any heap or
stack overflow, use after free, leak or undefined operation fatal.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")


def test_host_code_and_oracle_are_clean_under_sanitizers():
    b = subprocess.run(["make", "-C", SAN, "-j8"], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stdout[-3000:] + b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(SAN, "_build", "host_check")], capture_output=True, text=True, timeout=600,
                       env=env)
    assert r.returncode == 0 and "host_check: ok" in r.stdout, (r.returncode, r.stdout[-3000:], r.stderr[-6000:])
    assert "runtime error" not in r.stderr, r.stderr[-6000:]
