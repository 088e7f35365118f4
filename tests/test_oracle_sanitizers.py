"""The oracle under the sanitizers (CPU only): oracle/sanitize/oracle_driver.c
runs every oracle code path (BGK/TRT, Zou-He velocity and pressure, no-op
outlets, open boundaries on every side, partially saturated bounce-back,
lift/drag, reductions, the by-value copies of the timing baseline) with
4 OpenMP threads, built with AddressSanitizer + UndefinedBehaviorSanitizer
(gcc) and with ThreadSanitizer (LLVM clang + libomp, run with the Archer
OpenMP tool so the runtime's barriers are known to ThreadSanitizer).  Both
must finish with no report."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

SAN = os.path.join(ROOT, "oracle", "sanitize")
LLVM = "/opt/rocm/llvm"


def _build(tmp_path, target):
    r = subprocess.run(["make", "-s", "-C", SAN, f"OUT={tmp_path}", f"{tmp_path}/{target}"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return os.path.join(str(tmp_path), target)


def _check(r):
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    for bad in ("ERROR: AddressSanitizer", "runtime error", "WARNING: ThreadSanitizer", "LeakSanitizer"):
        assert bad not in out, out[-4000:]
    assert out.count("mass") == 5


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_under_address_and_undefined_sanitizers(tmp_path):
    exe = _build(tmp_path, "oracle_asan")
    _check(subprocess.run([exe, "4"], capture_output=True, text=True, timeout=300))


@pytest.mark.skipif(not os.path.exists(os.path.join(LLVM, "lib", "libarcher.so")),
                    reason="LLVM OpenMP runtime with the Archer tool not available")
def test_oracle_under_thread_sanitizer(tmp_path):
    exe = _build(tmp_path, "oracle_tsan")
    env = dict(os.environ, OMP_TOOL_LIBRARIES=os.path.join(LLVM, "lib", "libarcher.so"),
               TSAN_OPTIONS="halt_on_error=1 ignore_noninstrumented_modules=1", OMP_NUM_THREADS="4")
    _check(subprocess.run([exe, "4"], capture_output=True, text=True, timeout=600, env=env))
