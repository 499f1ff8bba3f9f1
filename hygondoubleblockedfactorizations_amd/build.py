"""Builds ``libmifact.so`` (all gfx950 kernels + the C++ drivers) in-tree.

    python -m hygondoubleblockedfactorizations_amd.build

Each source is compiled by hipcc for gfx950 into an object under
``_build/`` (in parallel), then linked into
``hygondoubleblockedfactorizations_amd/lib/libmifact.so``.  Packed FP32 VALU
instructions are disabled for device code (``-packed-fp32-ops``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libmifact.so")

ARCH = os.environ.get("MIFACT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

SOURCES = [
    "handle.cpp",
    "gemm.hip",
    "panel.hip",
    "pivot.hip",
    "potrf.cpp",
    "recursive.cpp",
    "getrf.cpp",
    "refine.hip",
    "batched.hip",
]

# the rocblas_handle compatibility layer: its own library (the reference's
# names, forwarding to libmifact.so), never linked into libmifact.so.  Its
# file name must not end in "rocblas.so": rocBLAS takes the directory of a
# loaded module so named as the place of its Tensile kernels, and a process
# that loaded a library called libmifact_rocblas.so aborted at its first
# rocBLAS GEMM ("Cannot read .../lib/rocblas/library/TensileLibrary.dat").
COMPAT_SOURCE = "compat_rocblas.cpp"
COMPAT_LIB = os.path.join(LIBDIR, "libmifact_compat.so")

COMMON_FLAGS = [
    "-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


LLVM_BIN = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "llvm", "bin")


def _device_kernels(obj: str) -> set:
    """Kernel symbols of the gfx950 code object bundled in `obj`."""
    with tempfile.TemporaryDirectory() as tmp:
        fat = os.path.join(tmp, "fatbin")
        co = os.path.join(tmp, "device.co")
        subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin", obj, fat],
                       check=True, capture_output=True)
        subprocess.run([os.path.join(LLVM_BIN, "clang-offload-bundler"), "--unbundle",
                        "--type=o", f"--input={fat}",
                        f"--targets=hipv4-amdgcn-amd-amdhsa--{ARCH}", f"--output={co}"],
                       check=True, capture_output=True)
        out = subprocess.run([os.path.join(LLVM_BIN, "llvm-readelf"), "-s", co], check=True,
                             capture_output=True, text=True).stdout
    names = set()
    for line in out.splitlines():
        f = line.split()
        if len(f) >= 8 and f[3] == "FUNC" and f[4] == "GLOBAL":
            names.add(f[7])
    return names


def _host_kernels(obj: str) -> set:
    """Kernel names the host half of `obj` registers with the HIP runtime
    (the mangled names in its string table)."""
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "str")
        subprocess.run(["objcopy", "-O", "binary", "--only-section=.rodata.str1.1", obj, out],
                       check=True, capture_output=True)
        data = open(out, "rb").read()
    return {w.decode() for w in data.split(b"\0") if w.startswith(b"_ZN6mifact")}


def verify_object(obj: str) -> None:
    """Every kernel the host code registers must exist in the device code
    object of the same file: a launch of a missing one aborts the process
    at run time ('Cannot find Symbol')."""
    host = _host_kernels(obj)
    if not host:
        return
    dev = _device_kernels(obj)
    missing = sorted(host - dev)
    if missing:
        raise RuntimeError(f"{obj}: host registers kernels absent from its device code "
                           f"(stale or mixed object): {missing[:4]}")


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    # write to a private name, check, then rename: no other writer of the
    # same path can leave a mixed object behind
    tmp_obj = f"{obj}.{os.getpid()}.tmp"
    cmd = [HIPCC, *COMMON_FLAGS, "-x", "hip", "-c", path, "-o", tmp_obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    verify_object(tmp_obj)
    os.replace(tmp_obj, obj)
    # hipcc's host pass does not know the device-only feature: drop that note.
    noise = "is not a recognized feature for this target"
    msgs = [l for l in res.stderr.splitlines() if noise not in l]
    if msgs:
        sys.stderr.write("\n".join(msgs) + "\n")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_compat(verbose)
    return LIB


def build_compat(verbose: bool = True) -> str:
    """libmifact_compat.so: host-only forwarding code, linked to rocBLAS
    (rocblas_get_stream) and libdl."""
    src = os.path.join(CSRC, COMPAT_SOURCE)
    tmp = COMPAT_LIB + ".tmp"
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = [HIPCC, "-O2", "-std=c++20", "-fPIC", "-shared", f"-I{INCLUDE}",
           f"-I{os.path.join(rocm, 'include')}", "-D__HIP_PLATFORM_AMD__", src, "-o", tmp,
           f"-L{os.path.join(rocm, 'lib')}", "-lrocblas", "-ldl",
           f"-Wl,-rpath,{os.path.join(rocm, 'lib')}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compat build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, COMPAT_LIB)
    if verbose:
        print(f"built {COMPAT_LIB}")
    return COMPAT_LIB


if __name__ == "__main__":
    build()
