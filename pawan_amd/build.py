"""Build libpawan_gpu.so (HIP kernels + C-ABI + C++ class API) for gfx950.

Every HIP translation unit is compiled with hipcc for gfx950 only; the
objects go to pawan_amd/_build/ and the shared library to
pawan_amd/lib/libpawan_gpu.so (in-tree, so a copy of the tree carries it).
Packed FP32 instruction selection is switched off for all device code
(-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(CSRC, "host")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libpawan_gpu.so")
BINDIR = os.path.join(PKG, "bin")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-I", HOST,
]

HIP_SOURCES = ["pw_pair.hip", "pw_stage.hip", "pw_diag.hip", "pw_capi.hip", "pw_tree.hip",
               "pw_far.hip", "pw_mlp_train.hip", "pw_ff.hip", "pw_vring.hip", "pw_relax.hip",
               "pw_probe.hip"]
HOST_SOURCES = ["pawan_host.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the device-only
    # target feature; that line is noise.
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(err + "\n")
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    if err.strip():
        sys.stderr.write(err + "\n")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(BUILD, s.replace(".hip", ".o"))
        jobs.append([HIPCC, *HIPFLAGS, "-c", src, "-o", obj])
    for s in HOST_SOURCES:
        src = os.path.join(HOST, s)
        if not os.path.exists(src):
            continue
        obj = os.path.join(BUILD, s.replace(".cpp", ".o"))
        jobs.append([HIPCC, *HIPFLAGS, "-x", "hip", "-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    # the class-API driver, linked against the library with an rpath to it
    os.makedirs(BINDIR, exist_ok=True)
    exe = os.path.join(BINDIR, "pawan_run")
    _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", HOST,
          os.path.join(HOST, "pawan_run.cpp"), "-L", LIBDIR, "-lpawan_gpu",
          "-Wl,-rpath,$ORIGIN/../lib", "-o", exe + ".tmp"])
    os.replace(exe + ".tmp", exe)
    if verbose:
        print(f"built {LIB} and {exe}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
