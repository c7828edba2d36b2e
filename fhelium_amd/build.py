"""Build the gfx950 HIP library ``libfhelium_amd.so`` in-tree with hipcc.

The library is a plain C ABI (declared in ``include/fhelium_amd.h``) with no
PyTorch types in its signatures; the Python package loads it with ctypes.
Next to it the build writes ``_build_manifest.<soabi>.json`` (fhelium_amd/
manifest.py), which the loader checks.  ``CMakeLists.txt`` in this package
builds the same library with CMake's HIP language support.
Run as ``python -m fhelium_amd.build`` or through ``__graft_entry__.build()``.
"""

from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

PACKAGE = pathlib.Path(__file__).resolve().parent
CSRC = PACKAGE / "csrc"
INCLUDE = PACKAGE.parent / "include"
LIB_DIR = PACKAGE / "_lib"
LIB_PATH = LIB_DIR / "libfhelium_amd.so"
OFFLOAD_ARCH = "gfx950"

# -ffp-contract=off: the binary64 modular multiply relies on each product
# being rounded exactly where written, so the compiler must not fuse or
# reassociate.  The packed-FP32 feature is disabled for all device code.
HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={OFFLOAD_ARCH}",
    "-fPIC",
    "-shared",
    "-fvisibility=hidden",
    "-ffp-contract=off",
    # PyTorch-ROCm ships its own (older) HIP runtime, which this library
    # binds to at load time; code object v5 is understood by every runtime
    # of the ROCm 6/7 line.
    "-mcode-object-version=5",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for candidate in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if candidate and pathlib.Path(candidate).exists():
            return candidate
    raise RuntimeError("hipcc not found; the gfx950 library cannot be built")


def sources() -> list[pathlib.Path]:
    return sorted(CSRC.glob("*.hip"))


def build(verbose: bool = False) -> pathlib.Path:
    """Compile every HIP source into one shared library (always rebuilds)."""

    LIB_DIR.mkdir(exist_ok=True)
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{CSRC}", f"-I{INCLUDE}", *map(str, sources()), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with exit code {proc.returncode}")
    if verbose and (proc.stdout or proc.stderr):
        sys.stderr.write(proc.stdout + proc.stderr)
    os.replace(tmp, LIB_PATH)
    from fhelium_amd import _native, manifest

    manifest.write_manifest(LIB_PATH, hip_version=manifest.hipcc_version(hipcc()), offload_arch=OFFLOAD_ARCH,
                            library_abi=_native.ABI_VERSION)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
