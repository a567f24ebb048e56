"""Compile libmercury_hip.so in-tree for gfx950.

    python -m mercury_amd.build          # or __graft_entry__.build()

Every HIP/C++ source in mercury_amd/csrc is compiled with hipcc, with
floating-point contraction off (the step's arithmetic sequence is part of its
specification) and packed-FP32 instructions disabled, then linked into
mercury_amd/libmercury_hip.so.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(REPO, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmercury_hip.so")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-Wno-unused-function", "-I" + INCLUDE, "-I" + CSRC]
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-fno-gpu-rdc", "-munsafe-fp-atomics"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def host_cxx():
    for c in (os.environ.get("MH_HOST_CXX"), "/opt/rocm/lib/llvm/bin/clang++", shutil.which("clang++")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("clang++ not found")


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip"):
            out.append((os.path.join(CSRC, f), True))
        elif f.endswith(".cpp"):
            out.append((os.path.join(CSRC, f), False))
    return out


def compile_one(src, device, verbose=False):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if device:
        cmd = [hipcc()] + COMMON + DEVICE + ["-c", src, "-o", obj]
    else:
        # host-only sources: plain C++ (hipcc would compile them as HIP)
        cmd = [host_cxx()] + COMMON + ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                                       "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip() and verbose:
        print(r.stderr, file=sys.stderr)
    return obj


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s[0], s[1], verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "-pthread", "--offload-arch=" + ARCH, "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
