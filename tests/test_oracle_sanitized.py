"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

oracle/selftest.c drives every entry point of oracle.h on small grids; built
with -fsanitize=address,undefined (make -C oracle _build/selftest), any
out-of-bounds access, leak or undefined operation aborts it.
"""
import os
import subprocess

from conftest import ROOT


def test_oracle_selftest_is_clean_under_sanitizers():
    d = os.path.join(ROOT, "oracle")
    b = subprocess.run(["make", "-C", d, "_build/selftest"], capture_output=True, text=True,
                       timeout=300)
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(d, "_build", "selftest")], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "selftest ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
