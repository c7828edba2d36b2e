"""Test configuration: the ``gpu`` marker and repository imports.

GPU tests exercise the gfx950 library through its C ABI and compare against
the CPU oracle; they are skipped when no ROCm device is visible.  CPU tests
cover the oracle against golden vectors, host logic and the library's
exported symbol set.
"""

from __future__ import annotations

import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: requires a ROCm (gfx950) device and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU checks")


def _gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:  # pragma: no cover
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no ROCm device visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
