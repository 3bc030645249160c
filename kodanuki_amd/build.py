"""Build libkodanuki_hip.so for gfx950 with hipcc, in the package folder.

Every HIP source is compiled without FP contraction (each operation rounds
once, as the CPU oracle does) and without packed FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libkodanuki_hip.so")

SOURCES = ["abi.cpp", "scene.cpp", "tables.cpp", "image.cpp", "scan.hip", "octree.hip", "trace.hip", "integrator.hip", "serve.hip", "guide.hip", "guide_net.hip",
           "denoise.hip", "hguide.hip", "sort.hip", "stopping.hip", "temporal.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument", "-Wno-unused-result", "-Wno-unused-value",
]


def _compile(src, obj):
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    subprocess.run(cmd, check=True)


def build(verbose=True):
    """Compile every source and link the shared library; return its path."""
    objdir = os.path.join(HERE, "_objs")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(text)
        elif verbose and "warning" in text:
            sys.stderr.write(text)
    if failed:
        raise RuntimeError("hipcc failed on " + ", ".join(failed))
    tmp = OUT + ".tmp"
    subprocess.run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs, check=True)
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build()
