"""Builds libmfdag.so (HIP kernels + host runtime + C-ABI) for gfx950.

Run as a script or call build().  Every source is compiled on each call; the
objects go to multifrontal-dag_amd/_build and the library to
multifrontal-dag_amd/lib/libmfdag.so.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libmfdag.so")
SOURCES = ["mfd_analysis.cpp", "mfd_layout.cpp", "mfd_mmio.cpp", "mfd_ordering.cpp", "mfd_subtree.cpp", "mfd_host.cpp",
           "mfd_kernels.hip", "mfd_runtime.hip", "mfd_mlp.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for all device code.
FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950", "-Wall",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(ROOT, "include"), "-I", SRC]


# MFMA results in ordinary VGPRs: without it the accumulators of a K loop
# are copied to AGPRs and back around every MFMA group.
HIP_FLAGS = ["-mllvm", "-amdgpu-mfma-vgpr-form"]


# Host numeric code (the hybrid's subtree factorization) vectorized for
# AVX2 + FMA, which every x86-64 host of an MI355X has.
HOST_VECTOR = {"mfd_host.cpp"}


def _compile(src):
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] + HIP_FLAGS if src.endswith(".hip") else []
    if src in HOST_VECTOR:
        lang += ["-mavx2", "-mfma"]
    cmd = [HIPCC] + FLAGS + lang + ["-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "unknown target feature"
    err = "\n".join(l for l in r.stderr.splitlines() if noise not in l and "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stderr}")
    return obj, err


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(_compile, SOURCES))
    for _, err in results:
        if verbose and err.strip():
            print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(tmp, LIB)
    build_cli()
    build_fortran()
    return LIB


CLI_SRC = os.path.join(HERE, "cli", "mfdag_example.cpp")
CLI = os.path.join(HERE, "lib", "mfdag_example")


def build_cli():
    """The command-line driver (reads MatrixMarket, writes the one-line
    record), linked to libmfdag.so through an $ORIGIN run path."""
    libdir = os.path.dirname(LIB)
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), CLI_SRC, "-o", CLI,
                    "-L", libdir, "-lmfdag", "-Wl,-rpath,$ORIGIN"], check=True, capture_output=True, text=True)


FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")
FMOD = os.path.join(HERE, "fortran", "hppllt.f90")
FLIB = os.path.join(HERE, "lib", "libhppllt.so")
FCHECK_SRC = os.path.join(ROOT, "tests", "fortran", "hppllt_check.f90")
FCHECK = os.path.join(HERE, "lib", "hppllt_check")


def build_fortran():
    """The Fortran module hppllt (libhppllt.so, over libmfdag.so) and the
    caller-side check program, both with $ORIGIN run paths so the pair works
    from any copy of the tree."""
    moddir = os.path.join(OUT, "fmod")
    os.makedirs(moddir, exist_ok=True)
    libdir = os.path.dirname(LIB)
    rpath = "-Wl,-rpath,$ORIGIN"
    subprocess.run([FLANG, "-O2", "-fPIC", "-shared", "-module-dir", moddir, FMOD, "-o", FLIB,
                    "-L", libdir, "-lmfdag", rpath], check=True, capture_output=True, text=True)
    subprocess.run([FLANG, "-O2", "-I", moddir, FCHECK_SRC, "-o", FCHECK,
                    "-L", libdir, "-lhppllt", "-lmfdag", rpath], check=True, capture_output=True, text=True)


if __name__ == "__main__":
    print(build(verbose=True))
