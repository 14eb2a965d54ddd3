"""The host code behind the C ABI -- parameter vectors and their validity,
candidate enumeration, the journal and its mark/rank, the EM search and its
journal replay, the serving trace, the struct layouts -- run by the CPU
tests of those parts against a build whose host objects carry
AddressSanitizer and UndefinedBehaviorSanitizer (`make sanitize`: host side
only, the kernels are the normal objects).  Any sanitizer report fails the
run; the sanitised library is checked to be the one the tests loaded."""
import glob
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gatlas_amd", "csrc")
SAN_LIB = os.path.join(CSRC, "_build_san", "libgatlas_hip.so")
# the CPU tests that call into the library (the CLI tools link the normal build)
FILES = ["tests/test_capi_load.py", "tests/test_params_golden.py", "tests/test_search_replay.py",
         "tests/test_search_sets.py", "tests/test_serve_trace.py"]


def _asan_runtime():
    found = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return found[-1] if found else None


def test_host_code_clean_under_asan_and_ubsan():
    rt = _asan_runtime()
    if rt is None or shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("no hipcc / clang sanitizer runtime in this image")
    subprocess.run(["make", "-s", "-j8", "-C", CSRC, "sanitize"], check=True, timeout=900)
    env = dict(os.environ, LD_PRELOAD=rt, GATLAS_LIB=SAN_LIB,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    probe = subprocess.run([sys.executable, "-c",
                            "from gatlas_amd import _capi; _capi.load(); "
                            "print(any('_build_san' in l for l in open('/proc/self/maps')))"],
                           env=env, cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert probe.returncode == 0 and probe.stdout.strip() == "True", probe.stdout + probe.stderr
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider"] + FILES,
                       env=env, cwd=ROOT, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert " passed" in out
