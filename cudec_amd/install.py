"""Installs the built library as a CMake package prefix:

    python -m cudec_amd.install --prefix P

copies include/*.h to P/include, cudec_amd/libcudec.a (and libcudec.so) to
P/lib, and the package config to P/lib/cmake/cudec, so that a consumer's
find_package(cudec 0.1.0) finds the static target cudec::cudec. Build the
library first (__graft_entry__.build())."""
from __future__ import annotations

import argparse
import glob
import os
import shutil
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)


def install(prefix: str) -> None:
    lib_a = os.path.join(PKG, "libcudec.a")
    if not os.path.exists(lib_a):
        raise SystemExit(f"{lib_a} is missing: build the library first")
    inc = os.path.join(prefix, "include")
    lib = os.path.join(prefix, "lib")
    cfg = os.path.join(lib, "cmake", "cudec")
    for d in (inc, lib, cfg):
        os.makedirs(d, exist_ok=True)
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        shutil.copy2(h, inc)
    shutil.copy2(lib_a, lib)
    so = os.path.join(PKG, "libcudec.so")
    if os.path.exists(so):
        shutil.copy2(so, lib)
    for f in ("cudec-config.cmake", "cudec-config-version.cmake"):
        shutil.copy2(os.path.join(PKG, "cmake", f), cfg)


def main(argv=None) -> int:
    p = argparse.ArgumentParser(prog="python -m cudec_amd.install")
    p.add_argument("--prefix", required=True)
    install(p.parse_args(argv).prefix)
    return 0


if __name__ == "__main__":
    sys.exit(main())
