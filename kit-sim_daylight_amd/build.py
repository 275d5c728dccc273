# SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
"""Builds the MI355X library in-tree: kit-sim_daylight_amd/lib/libdaylight_amd.so (HIP kernels for gfx950 + the C++ host
layer + the C ABI) and the two command-line tools lib/daylight and lib/views.

    python kit-sim_daylight_amd/build.py [-j N]

Objects go to kit-sim_daylight_amd/_build/ (not committed). Every source is compiled by hipcc: .hip files as HIP for
gfx950, .cpp files as host C++17 that include the shared __host__ __device__ headers.
"""
import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# host C++ goes through clang++ directly: hipcc would compile .cpp files as HIP and pre-include its runtime wrapper
# ahead of glibc_guard.h
CXX = os.environ.get("DL_CXX", "/opt/rocm/llvm/bin/clang++")
ARCH = os.environ.get("DL_OFFLOAD_ARCH", "gfx950")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function", "-Wno-unused-result",
          "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "csrc", "include"),
          "-include", os.path.join(HERE, "csrc", "include", "glibc_guard.h")]
HIP_FLAGS = ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]
HOST_FLAGS = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]


def compile_one(src):
    rel = os.path.relpath(src, os.path.join(HERE, "csrc")).replace(os.sep, "_")
    obj = os.path.join(BUILD, rel + ".o")
    flags = COMMON + (HIP_FLAGS if src.endswith(".hip") else HOST_FLAGS)
    cmd = [HIPCC if src.endswith(".hip") else CXX] + flags + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode:
        return None, "compile failed: %s\n%s" % (src, msgs)
    return obj, msgs


def build(jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    tools = {os.path.join(HERE, "csrc", "host", "tools.cpp")}
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "hip", "*.hip")) + glob.glob(os.path.join(HERE, "csrc", "host", "*.cpp")))
    srcs = [s for s in srcs if s not in tools]
    lib_srcs = srcs
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(compile_one, srcs))
    failed = [msgs for obj, msgs in results if obj is None]
    if failed:
        raise RuntimeError("\n".join(failed))
    for obj, msgs in results:
        if msgs.strip() and verbose:
            print(msgs)
    objs = {s: o for s, (o, _) in zip(srcs, results)}
    so = os.path.join(LIB, "libdaylight_amd.so")
    tmp = so + ".tmp"
    link = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + [objs[s] for s in lib_srcs] + ["-lpthread"]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError("link failed\n" + r.stdout + r.stderr)
    os.replace(tmp, so)
    for tool, define in (("daylight", []), ("views", ["-DVIEWS_CLI"])):
        main = os.path.join(HERE, "csrc", "host", "tools.cpp")
        if not os.path.exists(main):
            continue
        exe = os.path.join(LIB, tool)
        cmd = [CXX] + COMMON + HOST_FLAGS + define + [main, "-o", exe, "-L" + LIB, "-ldaylight_amd", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError("tool %s failed\n%s%s" % (tool, r.stdout, r.stderr))
    return so


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.v))
