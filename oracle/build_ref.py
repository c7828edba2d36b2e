"""Build the reference's own CPU kernels into oracle/_ref/ (test infrastructure).

The reference's CPU backend for this path is nine C++17 sources over ATen
(csrc/ops/{ntt,rns,ckks}/*.cpp schemas and csrc/ops/*/cpu/*.cpp kernels,
with csrc/ops/common/{rns_batch.h,rns_parameters.h,cpu/montgomery.h}).  They
need only the PyTorch headers and libraries this image already has, so they
are compiled here, from where they lie under /root/reference, with g++
directly (the reference's own CMake build is not used), into
oracle/_ref/libfhelium_ref_cpu.so.  Loading it defines the reference's
operator schemas with its CPU kernels; oracle/ref_check.py then runs those
kernels beside the oracle's C restatement on the same inputs.

Nothing from the reference is copied into this repository, and the built
library is git-ignored; it is the checker of the restatement, never part of
the product.  Run: python oracle/build_ref.py  (a no-op without the
reference tree).
"""

from __future__ import annotations

import os
import pathlib
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = pathlib.Path(__file__).resolve().parent
REFERENCE = pathlib.Path(os.environ.get("FHELIUM_REFERENCE", "/root/reference")) / "csrc"
OUT = ROOT / "_ref"
LIB = OUT / "libfhelium_ref_cpu.so"
SOURCES = [
    "ops/ntt/forward_ntt.cpp",
    "ops/ntt/inverse_ntt.cpp",
    "ops/rns/rns_arithmetic.cpp",
    "ops/rns/rns_canonical_arithmetic.cpp",
    "ops/ckks/ckks.cpp",
    "ops/ntt/cpu/indexed_ntt_cpu.cpp",
    "ops/rns/cpu/rns_arithmetic_cpu.cpp",
    "ops/rns/cpu/rns_canonical_arithmetic_cpu.cpp",
    "ops/ckks/cpu/ckks_cpu.cpp",
]


def available() -> bool:
    return all((REFERENCE / s).exists() for s in SOURCES)


def build(verbose: bool = False) -> pathlib.Path | None:
    if not available():
        if verbose:
            print(f"reference sources not found under {REFERENCE}; oracle/_ref not built")
        return None
    import torch
    from torch.utils import cpp_extension

    OUT.mkdir(exist_ok=True)
    abi = int(torch.compiled_with_cxx11_abi())
    flags = ["-std=c++17", "-O2", "-fPIC", "-fopenmp", f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-DTORCH_EXTENSION_NAME=fhelium_ref_cpu",
             *[f"-I{p}" for p in cpp_extension.include_paths()], "-w"]
    objects = []

    def compile_one(src: str) -> pathlib.Path:
        obj = OUT / (src.replace("/", "_") + ".o")
        path = REFERENCE / src
        if obj.exists() and obj.stat().st_mtime > path.stat().st_mtime:
            return obj
        cmd = ["g++", *flags, "-c", str(path), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        objects = list(pool.map(compile_one, SOURCES))
    libs = cpp_extension.library_paths()
    cmd = ["g++", "-shared", "-fopenmp", "-o", str(LIB), *map(str, objects), *[f"-L{p}" for p in libs],
           *[f"-Wl,-rpath,{p}" for p in libs], "-lc10", "-ltorch_cpu", "-ltorch"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
