"""The oracle against the reference's own CPU kernels.

oracle/build_ref.py compiles the reference's nine CPU sources
(csrc/ops/{ntt,rns,ckks}/*.cpp and their cpu/ kernels) against the image's
PyTorch into oracle/_ref/libfhelium_ref_cpu.so; oracle/ref_check.py loads it
in a child process (it defines the same operator names our package registers
for ROCm tensors) and compares each of its CPU ops with the oracle, bit for
bit, on seeded inputs.  The planted run adds one unit to one element of every
oracle result and must fail every case, so no comparison is vacuous.
Skipped where the library has not been built (no reference tree).
"""

from __future__ import annotations

import json
import pathlib
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
LIB = ROOT / "oracle" / "_ref" / "libfhelium_ref_cpu.so"

pytestmark = pytest.mark.skipif(not LIB.exists(), reason="oracle/_ref not built (needs the reference sources)")


def _run(*args):
    proc = subprocess.run([sys.executable, str(ROOT / "oracle" / "ref_check.py"), *args],
                          capture_output=True, text=True, timeout=300, cwd=ROOT)
    return proc.returncode, json.loads(proc.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("log_n", [10, 12])
def test_reference_cpu_kernels_equal_oracle(log_n):
    rc, out = _run("--log-n", str(log_n))
    assert out["cases"] >= 120
    assert rc == 0 and out["mismatches"] == [], out["mismatches"]
    ops = {r["case"].split("[")[0] for r in out["results"]}
    # every CPU kernel the reference registers for this path
    assert len(ops) == 43, sorted(ops)


def test_planted_oracle_mistake_fails_every_case():
    rc, out = _run("--plant")
    assert rc == 1
    assert len(out["mismatches"]) == out["cases"]


def test_reference_kernel_engine_equals_oracle_engine():
    """HMult and rotation composed on the reference's CPU kernels
    (oracle/ref_engine.py) equal the oracle's composition (oracle/engine_ref.py)
    bit for bit at logN=12, levels 0, 7 and 30."""
    proc = subprocess.run([sys.executable, str(ROOT / "oracle" / "ref_baseline.py"), "--check"],
                          capture_output=True, text=True, timeout=300, cwd=ROOT)
    out = json.loads(proc.stdout.strip().splitlines()[-1])
    assert proc.returncode == 0 and out["equal"], out
