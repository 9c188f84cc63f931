"""Builds libtsg.so (HIP kernels for gfx950 + host C++) and the spgemm_raw_v2 CLI.

Every .hip file is compiled separately with hipcc (in parallel), host .cpp
files with g++, and everything is linked into gpu_spgemm_amd/libtsg.so.
Usage: python -m gpu_spgemm_amd.build   (or gpu_spgemm_amd.build.build())
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libtsg.so")
CLI = os.path.join(PKG, "spgemm_raw_v2")
ARCH = os.environ.get("TSG_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# Packed FP32 instructions are disabled (-packed-fp32-ops) for every HIP file.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast", "-Wno-unused-result", "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + p.stdout + p.stderr)
    return p


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    hips = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpps = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    jobs = []
    for f in hips:
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, f), "-o", os.path.join(OBJ, f + ".o")])
    for f in cpps:
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, f), "-o", os.path.join(OBJ, f + ".o")])
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for p in ex.map(_run, jobs):
            if verbose and (p.stdout or p.stderr):
                print(p.stdout, p.stderr)
    objs = [os.path.join(OBJ, f + ".o") for f in hips + cpps]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "cli", "spgemm_raw_v2.cpp")
    if os.path.exists(cli_src):
        _run(["g++", *CXX_FLAGS, "-o", CLI, cli_src, f"-L{PKG}", "-ltsg", f"-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
