"""Builds libhs_amd.so (gfx950 HIP kernels + C ABI) and the C++ host library
in-tree with hipcc / g++. Usage: python heterogeneous-solvers_amd/build.py"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # keep MFMA accumulators in ordinary VGPRs (no AGPR copies around loops)
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return r


def _compile_hip(src):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    _run([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
    return obj


def build_hip():
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile_hip, srcs))
    out = os.path.join(LIBDIR, "libhs_amd.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


def build_host():
    """C++ drop-in host API (reference class names) over the C ABI."""
    if not os.path.isdir(HOST):
        return None
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    if not srcs:
        return None
    out = os.path.join(LIBDIR, "libhs_host.so")
    tmp = out + ".tmp"
    # the CPU half of the heterogeneous solves runs on OpenMP
    _run(["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall",
          "-I" + os.path.join(ROOT, "include"), "-o", tmp] + srcs + ["-L" + LIBDIR, "-lhs_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    for src, name in ((os.path.join(HOST, "driver", "hs_driver.cpp"), "hs_driver"),
                      (os.path.join(HOST, "tests", "hs_host_tests.cpp"), "hs_host_tests")):
        if os.path.exists(src):
            exe = os.path.join(LIBDIR, name)
            _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-Wall", "-I" + os.path.join(ROOT, "include"), "-o",
                  exe + ".tmp", src, "-L" + LIBDIR, "-lhs_host", "-lhs_amd", "-Wl,-rpath,$ORIGIN"])
            os.replace(exe + ".tmp", exe)
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def main():
    build_hip()
    build_host()
    build_oracle()


if __name__ == "__main__":
    main()
