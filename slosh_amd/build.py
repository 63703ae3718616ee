"""Builds the solver library (slosh_amd/libslosh.so) for gfx950 with hipcc.

Every HIP source is compiled with -ffp-contract=off (the kernels keep the written FP32 evaluation order, which the
parity tests compare bit for bit against the CPU oracle) and with packed-FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libslosh.so")
BUILD_DIR = os.path.join(HERE, "_build")
SOURCES = ["particles.hip", "grid.hip", "level0.hip", "pooled.hip", "solver.hip", "unet.hip", "batch.hip", "ctx.hip"]
ARCH = "gfx950"


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the solver library needs ROCm's HIP compiler")
    return path


def flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    ]


def build(verbose=False, jobs=4):
    os.makedirs(BUILD_DIR, exist_ok=True)
    cc = hipcc()
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(BUILD_DIR, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [cc] + flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = OUT + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    text = out.decode(errors="replace")
    if p.returncode != 0:
        sys.stderr.write(text)
        raise RuntimeError(f"hipcc failed on {src}")
    warn = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if warn:
        sys.stderr.write("\n".join(warn) + "\n")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
