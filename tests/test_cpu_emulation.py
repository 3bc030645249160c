"""Bounds check of the whole library on the CPU.

The library's own sources (every kernel and every host launch) are built
by g++ against a CPU stand-in of the HIP runtime with AddressSanitizer
(tests/cpu_build), then driven through the C ABI over the GPU tests' call
sequence: octree builds, extend batches, furnace and scene renders, a
counted and a timed pass. An out-of-bounds access anywhere aborts the run.
"""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "cpu_build")


def _asan_runtime():
    try:
        out = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return out if os.path.isabs(out) and os.path.exists(out) else None


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_library_runs_clean_under_address_sanitizer():
    runtime = _asan_runtime()
    if runtime is None:
        pytest.skip("no AddressSanitizer runtime")
    subprocess.run(["make", "-s", "-C", BUILD, "-j4"], check=True, timeout=600)
    env = dict(os.environ, LD_PRELOAD=runtime, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    env.pop("KDN_EMULATION", None)
    r = subprocess.run([sys.executable, os.path.join(BUILD, "run_emu.py"), "--quick"], env=env, capture_output=True,
                       text=True, timeout=900)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stderr[-6000:]
    assert "[emu] done" in r.stdout
