"""Build manifest of the gfx950 library: what it was built for, checked at load.

Counterpart of the reference's native ABI manifest
(fhelium/native/_abi.py:13-63, checked in fhelium/native/runtime.py:108-156):
``_lib/_build_manifest.<soabi>.json`` records the schema version, the
dispatch backends the library serves, the Python and Torch identity of the
build, and -- because ``torch.version.cuda`` is None on ROCm -- the HIP
version of the compiler that built the code objects, the offload
architecture, the library's C-ABI version and the SHA-256 of the library
file.  ``_native.load()`` refuses a library whose manifest is missing, stale
(the file changed after the manifest was written) or built for another
Python/Torch identity, by raising ``NativeExtensionError``.
"""

from __future__ import annotations

import hashlib
import json
import pathlib
import re
import subprocess
import sys
import sysconfig
from collections.abc import Mapping
from typing import Any

MANIFEST_SCHEMA_VERSION = 1
SUPPORTED_NATIVE_BACKENDS = frozenset({"cpu", "cuda"})
# A ROCm tensor carries the CUDA dispatch key, so that is the backend the
# library's operators serve.
NATIVE_BACKENDS = ["cuda"]


def soabi() -> str:
    return sysconfig.get_config_var("SOABI") or "unknown"


def manifest_path(lib_dir: pathlib.Path) -> pathlib.Path:
    return lib_dir / f"_build_manifest.{soabi()}.json"


def runtime_identity() -> dict[str, Any]:
    """The Python/Torch fields the manifest must match at load."""

    import torch

    return {
        "python": {
            "implementation": sys.implementation.name,
            "soabi": sysconfig.get_config_var("SOABI") or "",
            "ext_suffix": sysconfig.get_config_var("EXT_SUFFIX") or "",
        },
        "torch": {
            "version": str(torch.__version__),
            "cuda_version": torch.version.cuda,
            "hip_version": getattr(torch.version, "hip", None),
        },
    }


def hipcc_version(hipcc: str) -> str:
    """HIP version reported by the compiler (e.g. '7.2.26015')."""

    try:
        out = subprocess.run([hipcc, "--version"], capture_output=True, text=True, check=True).stdout
    except (OSError, subprocess.CalledProcessError) as exc:
        raise RuntimeError(f"cannot query {hipcc} --version: {exc}") from exc
    m = re.search(r"HIP version:\s*([0-9][0-9.\-a-z]*)", out)
    if not m:
        raise RuntimeError(f"{hipcc} --version printed no HIP version")
    return m.group(1)


def file_sha256(path: pathlib.Path) -> str:
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def write_manifest(lib_path: pathlib.Path, *, hip_version: str, offload_arch: str, library_abi: int) -> pathlib.Path:
    body = {
        "manifest_schema_version": MANIFEST_SCHEMA_VERSION,
        "native_backends": list(NATIVE_BACKENDS),
        **runtime_identity(),
        "hip_version": hip_version,
        "offload_arch": offload_arch,
        "library": lib_path.name,
        "library_abi": library_abi,
        "library_sha256": file_sha256(lib_path),
    }
    path = manifest_path(lib_path.parent)
    tmp = path.with_suffix(".json.tmp")
    tmp.write_text(json.dumps(body, indent=1, sort_keys=True) + "\n")
    tmp.replace(path)
    return path


def read_manifest(path: pathlib.Path) -> dict[str, Any]:
    try:
        value = json.loads(path.read_text(encoding="utf-8"))
    except (OSError, json.JSONDecodeError) as exc:
        raise ValueError(f"cannot read build manifest {path}: {exc}") from exc
    if not isinstance(value, dict):
        raise ValueError(f"build manifest {path} is not a JSON object")
    return value


def compatibility_mismatches(manifest: Mapping[str, Any], *, lib_path: pathlib.Path, library_abi: int,
                             runtime: Mapping[str, Any] | None = None) -> list[str]:
    """Every way the manifest disagrees with this process and this file."""

    runtime = runtime_identity() if runtime is None else runtime
    problems: list[str] = []

    def compare(label: str, built: Any, now: Any) -> None:
        if built != now:
            problems.append(f"{label}: built={built!r}, runtime={now!r}")

    compare("manifest schema", manifest.get("manifest_schema_version"), MANIFEST_SCHEMA_VERSION)
    backends = manifest.get("native_backends")
    if (not isinstance(backends, list) or not backends or any(not isinstance(b, str) for b in backends)
            or len(set(backends)) != len(backends) or set(backends) - SUPPORTED_NATIVE_BACKENDS):
        problems.append(f"native_backends is invalid: {backends!r}")
    if not manifest.get("hip_version"):
        problems.append("hip_version is missing")
    compare("library ABI", manifest.get("library_abi"), library_abi)
    compare("library file", manifest.get("library"), lib_path.name)
    if lib_path.exists():
        compare("library SHA-256", manifest.get("library_sha256"), file_sha256(lib_path))
    for section, fields in (("python", ("implementation", "soabi", "ext_suffix")),
                            ("torch", ("version", "cuda_version", "hip_version"))):
        built = manifest.get(section)
        now = runtime.get(section, {})
        if not isinstance(built, Mapping):
            problems.append(f"{section} identity is missing from the manifest")
            continue
        for field in fields:
            compare(f"{section} {field}", built.get(field), now.get(field))
    return problems


def main(argv: list[str]) -> int:
    """``python -m fhelium_amd.manifest <library>``: write the manifest for a
    library built outside ``fhelium_amd.build`` (the CMake build)."""

    from fhelium_amd import _native, build

    if len(argv) < 1:
        print("usage: python -m fhelium_amd.manifest <libfhelium_amd.so>", file=sys.stderr)
        return 2
    lib = pathlib.Path(argv[0]).resolve()
    print(write_manifest(lib, hip_version=hipcc_version(build.hipcc()), offload_arch=build.OFFLOAD_ARCH,
                         library_abi=_native.ABI_VERSION))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
