"""Host-only checks of the C ABI's bookkeeping under the sanitizers (CPU).

tests/host/ builds two programs from the library's host-only headers:
host_units (AddressSanitizer + UndefinedBehaviorSanitizer) checks that the
descriptor key ignores Vship_Colorspace_t's padding and the handle table's id
semantics; handle_table_stress (ThreadSanitizer) frees and re-creates handlers
on two threads while four others look them up and use them, the pattern
behind the reference's use-after-free (src/VshipLib.cpp:242-258).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "tests", "host")


def _env(**extra):
    # these programs carry their own (gcc) sanitizer runtimes: drop a
    # preloaded one from a sanitized parent process
    env = {k: v for k, v in os.environ.items() if k not in ("LD_PRELOAD", "ASAN_OPTIONS", "UBSAN_OPTIONS")}
    env.update(extra)
    return env


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-C", HOST], check=True, env=_env())
    return os.path.join(HOST, "_build")


def test_descriptor_key_and_handle_table_under_asan_ubsan(built):
    r = subprocess.run([os.path.join(built, "host_units")], capture_output=True, text=True, timeout=120, env=_env())
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host units: ok" in r.stdout


def test_free_during_compute_under_tsan(built):
    env = _env(TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([os.path.join(built, "handle_table_stress")], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "handle table stress: ok" in r.stdout
    assert "ThreadSanitizer" not in r.stderr
