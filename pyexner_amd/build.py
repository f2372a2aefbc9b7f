"""Build libpyexner_hip.so in-tree with hipcc for gfx950.

Two translation units (the forward path and the adjoint) are compiled in
parallel and linked into one shared library.  The library is compiled with
-ffp-contract=off so the double build reproduces the float64 restatement's
operation order, and with packed fp32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "csrc", n) for n in ("pyexner_hip.hip", "adjoint.hip")]
OUT = os.path.join(HERE, "lib", "libpyexner_hip.so")
# The same library with AddressSanitizer and UndefinedBehaviorSanitizer on its
# host code (argument checks, launch planning, the ensemble plan); the device
# code is the ordinary -O3 build.  Test infrastructure: tests/test_host_sanitized.py
# runs the host tests against it.
OUT_SANITIZED = os.path.join(HERE, "lib", "host-sanitized", "libpyexner_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PYEXNER_ARCH", "gfx950")

FLAGS = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]
FLAGS_SANITIZED = ["-Xarch_host", "-O1", "-Xarch_device", "-O3", "-fno-omit-frame-pointer", f"--offload-arch={ARCH}",
                   "-fPIC", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
                   "-packed-fp32-ops"] + SANITIZE + ["-Xarch_host", "-fno-sanitize-recover=undefined"]


def _check(r, cmd):
    if r.returncode != 0:
        sys.stderr.write((r.stdout or "") + (r.stderr or ""))
        raise RuntimeError(f"hipcc failed ({r.returncode}): {' '.join(cmd)}")


def build(verbose=False, sanitized=True):
    """Compile both translation units (and, with sanitized, their
    host-sanitized variants) in parallel and link the libraries."""
    outs = [(OUT, FLAGS, [])]
    if sanitized:
        outs.append((OUT_SANITIZED, FLAGS_SANITIZED, SANITIZE + ["-shared-libasan"]))
    procs, plan = [], []
    for out, flags, _ in outs:
        os.makedirs(os.path.dirname(out), exist_ok=True)
        objs = []
        for src in SRCS:
            obj = os.path.join(os.path.dirname(out), os.path.splitext(os.path.basename(src))[0] + ".o")
            cmd = [HIPCC] + flags + ["-c", "-o", obj, src]
            procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True), cmd))
            objs.append(obj)
        plan.append(objs)
    # wait for every compile before reporting a failure, so that none is left
    # running (and writing its temporary object) behind the error
    results = []
    for p, cmd in procs:
        out, err = p.communicate()
        results.append(subprocess.CompletedProcess(cmd, p.returncode, out, err))
    for r in results:
        _check(r, r.args)
        if verbose:
            sys.stderr.write(r.stderr)
    for (out, _, link), objs in zip(outs, plan):
        tmp = out + ".tmp"
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC"] + link + ["-o", tmp] + objs
        _check(subprocess.run(cmd, capture_output=True, text=True), cmd)
        os.replace(tmp, out)
        for o in objs:
            os.remove(o)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
