"""The host tests again, against the library built with AddressSanitizer and
UndefinedBehaviorSanitizer on its host code (pyexner_amd.build.OUT_SANITIZED:
argument checks, launch planning, the ensemble and halo plans, the
boundary-program checks).  A sanitizer report fails the run."""
import glob
import os
import subprocess
import sys

import pytest

from pyexner_amd.build import OUT_SANITIZED

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _asan_runtime():
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


def test_host_tests_under_sanitizers():
    rt = _asan_runtime()
    if rt is None or not os.path.exists(OUT_SANITIZED):
        pytest.skip("no AddressSanitizer runtime or no host-sanitized build (pyexner_amd.build.build makes it)")
    env = dict(os.environ, LD_PRELOAD=rt, PYEXNER_HIP_LIB=OUT_SANITIZED,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    # a child that checks it runs the sanitized build, then the host tests
    probe = ("from pyexner_amd import _hip; _hip.load(); m = open('/proc/self/maps').read(); "
             "assert _hip.lib_path() == %r and 'libclang_rt.asan' in m and _hip.lib_path() in m" % OUT_SANITIZED)
    r = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join("tests", "test_host.py")], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert " passed" in out
