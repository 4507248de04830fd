"""Builds the native libraries in-tree.

  gpufs_amd/libgpufs_amd.so     HIP device API kernels + host runtime (hipcc, gfx950)
  gpufs_amd/libgpufs_datagen.so synthetic workload generator (g++)
  gpufs_amd/examples/compat_copy reference-style driver built against gpufs_compat.h

Run as `python -m gpufs_amd.build` or through __graft_entry__.build().
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("GPUFS_OFFLOAD_ARCH", "gfx950")
HIP_SOURCES = ["runtime.hip", "k_gread.hip", "k_grep.hip", "k_gemv.hip", "k_serve.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    # code objects loadable by the HIP runtime bundled with PyTorch (ROCm 7.0) as well as ROCm 7.2
    "-mcode-object-version=5",
    # packed fp32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_hip(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if out.strip():
                print(out)
    lib = os.path.join(HERE, "libgpufs_amd.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, lib)
    return lib


def build_examples():
    """gpufs_amd/examples/compat_copy: a reference-style driver (gpufs_compat.h)."""
    src = os.path.join(HERE, "examples", "compat_copy.hip")
    exe = os.path.join(HERE, "examples", "compat_copy")
    tmp = exe + ".tmp"
    flags = [f for f in HIP_FLAGS if f != "-fPIC"]
    _run([HIPCC, *flags, "-I", CSRC, src, "-o", tmp, "-L", HERE, "-lgpufs_amd", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, exe)
    return exe


def build_datagen():
    lib = os.path.join(HERE, "libgpufs_datagen.so")
    tmp = lib + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-o", tmp,
          os.path.join(CSRC, "datagen.cpp"), "-lpthread"])
    os.replace(tmp, lib)
    return lib


def build_all():
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        f1 = ex.submit(build_hip)
        f2 = ex.submit(build_datagen)
        out = (f1.result(), f2.result())
    return out + (build_examples(),)


if __name__ == "__main__":
    for p in build_all():
        print("built", p)
    sys.exit(0)
