"""In-tree build of the native libraries.

* cudec_amd/libcudec.so, cudec_amd/libcudec.a -- the product: the HIP batch
  decoders (gfx950) and the C ABI of include/*.h, plus the host runtime
  around them (streaming context, frame decoder, serving loop).
* cudec_amd/libcudec_workload.so -- host-only seeded workload generator and
  the liblz4 / snappy compressors used to build inputs.

Every build recompiles (the build check must exercise the compiler). Objects
go to cudec_amd/_build/, which is not committed.
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
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
ARCH = "gfx950"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# Flags for every HIP translation unit: gfx950 only; packed-FP32 ops are
# disabled (the target-feature value "-packed-fp32-ops" turns them off).
HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-mcode-object-version=5",
    "-Wall", "-Wextra", "-Wno-unused-parameter",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", f"-I{INCLUDE}", f"-I{CSRC}",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", f"-I{INCLUDE}",
              f"-I{CSRC}", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]

HIP_SOURCES = ["abi.hip", "plain.hip", "tuned.hip", "digest.hip", "checksum.hip", "zstd.hip"]
HOST_SOURCES = ["stream.cpp", "frame.cpp", "serve.cpp", "tune.cpp"]


def _run(cmd: list[str]) -> None:
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + p.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    # surface real warnings, but not hipcc's host-pass note about the
    # device-only target feature
    lines = [ln for ln in p.stdout.splitlines()
             if ln.strip() and "packed-fp32-ops" not in ln and "not a recognized feature" not in ln]
    if lines:
        sys.stderr.write("\n".join(lines) + "\n")


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if src.endswith(".hip"):
        _run([HIPCC, *HIP_FLAGS, "-c", path, "-o", obj])
    else:
        _run(["g++", *HOST_FLAGS, "-c", path, "-o", obj])
    return obj


def build_product(jobs: int = 8) -> None:
    os.makedirs(BUILD, exist_ok=True)
    sources = HIP_SOURCES + HOST_SOURCES
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    so = os.path.join(PKG, "libcudec.so")
    tmp = so + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-lpthread", "-ldl"])
    os.replace(tmp, so)
    lib_a = os.path.join(PKG, "libcudec.a")
    if os.path.exists(lib_a):
        os.remove(lib_a)
    _run(["ar", "rcs", lib_a, *objs])


def build_workload() -> None:
    out = os.path.join(PKG, "libcudec_workload.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", "-Werror",
          "-Wno-missing-field-initializers", "-o", out + ".tmp",
          os.path.join(CSRC, "workload.c"), "-ldl", "-lpthread"])
    os.replace(out + ".tmp", out)


def build_all() -> None:
    build_workload()
    build_product()


if __name__ == "__main__":
    build_all()
