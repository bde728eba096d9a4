"""Builds libmops.so (the gfx950 HIP kernels and the C ABI) in this package.

Every ``csrc/*.hip`` file is compiled with hipcc for gfx950 only and linked
into one shared library next to this file. Run ``python -m matops_amd.build``
or call :func:`build`.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmops.so")
OBJDIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 instructions stay off (they are generated from plain float
    # code otherwise)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l)
    return r.returncode, out


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = []
    for f in srcs:
        obj = os.path.join(OBJDIR, f[:-4] + ".o")
        jobs.append((f, [HIPCC, *FLAGS, "-c", os.path.join(CSRC, f), "-o", obj], obj))
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        results = list(ex.map(lambda j: (j[0], *_run(j[1])), jobs))
    failed = False
    for name, rc, out in results:
        if out.strip() and (verbose or rc):
            print(f"[{name}]\n{out}", file=sys.stderr)
        failed |= rc != 0
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = OUT + ".tmp"
    rc, out = _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp,
                    *[j[2] for j in jobs]])
    if rc:
        print(out, file=sys.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
