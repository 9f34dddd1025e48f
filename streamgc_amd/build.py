"""Builds libstreamgc_amd.so (HIP kernels for gfx950 + C++17 host code) in-tree.

Usage: python -m streamgc_amd.build    (or streamgc_amd.build.build() from Python)
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libstreamgc_amd.so")
CLI = os.path.join(HERE, "streaming_gc")

HIP_SOURCES = ["sgc_kernels.hip", "sgc_rounds.hip", "sgc_engine.hip"]
CXX_SOURCES = ["sgc_host.cpp", "sgc_gen.cpp", "sgc_server.cpp", "sgc_output.cpp"]
CLI_SOURCES = ["sgc_cli.cpp"]

ARCH = os.environ.get("SGC_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]
# Packed FP32 instructions stay off in every HIP translation unit.
DEVICE = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "argument unused during compilation" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + out)
    if out.strip():
        print(out, file=sys.stderr)


def build(verbose=False):
    hipcc = _hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(objdir, s + ".o")
        _run([hipcc, f"--offload-arch={ARCH}", *COMMON, *DEVICE, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in CXX_SOURCES:
        src = os.path.join(CSRC, s)
        if not os.path.exists(src):
            continue
        o = os.path.join(objdir, s + ".o")
        _run([hipcc, *COMMON, "-fopenmp", "-pthread", "-x", "c++", "-c", src, "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fopenmp", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    cli_src = [os.path.join(CSRC, s) for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if cli_src:
        _run([hipcc, *COMMON, "-x", "c++", *cli_src, "-o", CLI, "-L", HERE, "-lstreamgc_amd",
              "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
