"""The oracle block pool (nightstream_amd/csrc/tspool.h) under 16 host threads, host-only.

tests/tspool_stress.cpp builds the pool's own code over a malloc backend with emulated streams
(executor threads, so work is in flight while the owner goes on) and drives it the way the
Twist/Shout oracles do: staged uploads, rounds, deferred folds, block growth and teardown, with
every block stamped by its owner.  It must run clean under ThreadSanitizer and AddressSanitizer,
and a planted mistake (the destructor releasing its blocks before draining its stream) must be
caught by the stamps.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "nightstream_amd", "csrc")
SRC = os.path.join(HERE, "tspool_stress.cpp")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def _build(tmp_path, name, flags):
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-pthread", *flags, "-I", CSRC, SRC, "-o", exe], check=True,
                   capture_output=True, timeout=300)
    return exe


def _run(exe, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([exe, "16", "40"], capture_output=True, text=True, timeout=300, env=e)


@pytest.mark.parametrize("san", ["thread", "address"])
def test_pool_clean_under_sanitizer(tmp_path, san):
    exe = _build(tmp_path, f"tspool_{san}", [f"-fsanitize={san}"])
    r = _run(exe, {"ASAN_OPTIONS": "detect_leaks=0", "TSAN_OPTIONS": "halt_on_error=0"})
    assert r.returncode == 0, r.stderr[-3000:]
    assert "Sanitizer" not in r.stderr, r.stderr[-3000:]
    assert r.stdout.startswith("ok")


def test_planted_early_release_is_caught(tmp_path):
    exe = _build(tmp_path, "tspool_plant", ["-DPLANT_RELEASE_BEFORE_DRAIN"])
    r = _run(exe)
    assert r.returncode != 0 and "stamp mismatch" in r.stderr
