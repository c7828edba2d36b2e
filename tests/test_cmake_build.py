"""The CMake build (fhelium_amd/CMakeLists.txt) configures for gfx950 and
builds the same library, with every symbol the header declares and a build
manifest the loader accepts.  CPU only: hipcc cross-compiles without a GPU."""

from __future__ import annotations

import ctypes
import pathlib
import re
import shutil
import subprocess
import sys

import pytest

from fhelium_amd import _native as nat
from fhelium_amd import manifest as M

ROOT = pathlib.Path(__file__).resolve().parents[1]


@pytest.mark.slow
def test_cmake_builds_the_library(tmp_path):
    if shutil.which("cmake") is None:
        pytest.skip("cmake is not installed")
    build = tmp_path / "build"
    out = tmp_path / "lib"
    subprocess.run(["cmake", "-S", str(ROOT / "fhelium_amd"), "-B", str(build), "-DCMAKE_PREFIX_PATH=/opt/rocm",
                    f"-DFHELIUM_AMD_LIB_DIR={out}", f"-DPython3_EXECUTABLE={sys.executable}"], check=True, capture_output=True, text=True, timeout=300)
    subprocess.run(["cmake", "--build", str(build), "-j", "8"], check=True, capture_output=True, text=True,
                   timeout=900)
    lib = out / "libfhelium_amd.so"
    assert lib.exists()
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "fhelium_amd.h").read_text(), flags=re.S)
    declared = set(re.findall(r"\b(fh_[a-z0-9_]+)\s*\(", text))
    handle = ctypes.CDLL(str(lib))
    assert [n for n in sorted(declared) if not hasattr(handle, n)] == []
    body = M.read_manifest(M.manifest_path(out))
    assert M.compatibility_mismatches(body, lib_path=lib, library_abi=nat.ABI_VERSION) == []
    assert body["offload_arch"] == "gfx950"
