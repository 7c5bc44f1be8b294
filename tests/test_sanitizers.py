"""The host code (forest, problem set-up checks, surface output, C-ABI
argument validation) under AddressSanitizer and UndefinedBehaviorSanitizer:
the host C++ is rebuilt with both (tmr_amd.build.build_sanitized, g++, no
GPU code instrumented) and the CPU tests that drive it run in a child
process against that build, with the sanitizer runtimes preloaded.  Any
report fails the test."""
import os
import subprocess
import sys

import pytest

from tmr_amd import build

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _runtime(name):
    r = subprocess.run(["gcc", "-print-file-name=%s" % name], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_host_code_under_asan_and_ubsan(tmp_path):
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("g++ sanitizer runtimes not installed")
    try:
        lib = build.build_sanitized(str(tmp_path / "asan"))
    except RuntimeError as e:
        pytest.skip(str(e))
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan + ":" + ubsan, TMRX_LIB=lib,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    # the child processes load the instrumented build
    r = subprocess.run([sys.executable, "-c", "from tmr_amd import _lib; _lib.lib(); "
                        "print(_lib.LIB_PATH)"], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith(lib), r.stdout + r.stderr
    tests = [os.path.join(HERE, t) for t in
             ("test_forest.py", "test_stl.py", "test_validation.py", "test_topo.py",
              "test_nonnested.py", "test_checkpoint.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-m", "not gpu"] + tests, cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=900)
    out = r.stdout + r.stderr
    print(out[-3000:])
    assert "AddressSanitizer" not in out and "runtime error:" not in out
    assert r.returncode == 0
