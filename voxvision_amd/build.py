"""Build libvoxgpu.so in-tree (hipcc for gfx950, g++ for the host builder).

    python -m voxvision_amd.build        # or voxvision_amd.build.build()

Flags that matter for parity: -ffp-contract=off everywhere (fused
multiply-adds only where the sources call fmaf), no fast-math, IEEE
division (the HIP default).  Packed fp32 instructions are disabled.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "libvoxgpu.so")
BUILD = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_SOURCES = ["voxgpu.hip", "volume.hip"]
CXX_SOURCES = ["tree_build.cpp", "tree_edit.cpp", "lights.cpp"]
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-fno-fast-math", "-Wall", "-pthread",
             "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    objs = []
    log = ""
    for name in CXX_SOURCES:
        obj = os.path.join(BUILD, name.replace(".cpp", ".o"))
        log += _run([cxx] + CXX_FLAGS + ["-c", os.path.join(CSRC, name),
                                         "-o", obj])
        objs.append(obj)
    tmp = OUT + ".tmp"
    hobjs = []
    for name in HIP_SOURCES:
        hobj = os.path.join(BUILD, name.replace(".hip", ".o"))
        log += _run([HIPCC] + HIP_FLAGS + [
            "-c", os.path.join(CSRC, name), "-o", hobj])
        hobjs.append(hobj)
    log += _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", tmp,
                 "-x", "none"] + hobjs + objs + ["-lpthread"])
    os.replace(tmp, OUT)
    if verbose and log:
        print(log)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    print(OUT)
    sys.exit(0)
