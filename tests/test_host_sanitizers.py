"""The library's host code under sanitizers (tests/sanitize): scene
construction, the SAH builder (serial and with worker threads, whose trees
must be byte-identical), light tables and the C ABI marshalling including
textures, built with AddressSanitizer + UndefinedBehaviorSanitizer and with
ThreadSanitizer and run over small versions of every procedural scene. CPU
only; skipped where g++ is absent."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_host_code_is_clean_under_sanitizers(kind):
    r = subprocess.run(["make", "-C", HERE, kind], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    # verify_asan_link_order=0: a library the environment preloads may come
    # before the ASan runtime in the link order.
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([os.path.join(HERE, "_build", f"host_check_{kind}")], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-3000:])
    assert "host check passed" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "WARNING: ThreadSanitizer" not in r.stderr
    assert "runtime error" not in r.stderr
