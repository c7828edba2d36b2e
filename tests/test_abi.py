"""The C-ABI boundary on the CPU: the library loads without a GPU, exports
every entry point include/fhelium_amd.h declares, and the package's ctypes
table binds exactly those; the reference operator schemas are registered
for the CUDA dispatch key only (no host fallback)."""

from __future__ import annotations

import ctypes
import pathlib
import re

import pytest
import torch

from fhelium_amd import _native as nat

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "fhelium_amd.h"


def _declared() -> set[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(fh_[a-z0-9_]+)\s*\(", text))


@pytest.fixture(scope="module")
def library():
    if not nat.LIB_PATH.exists():
        from fhelium_amd import build
        build.build()
    return nat.load()


def test_header_declares_entry_points():
    names = _declared()
    assert "fh_keyswitch_fused" in names and "fh_ntt_forward" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol(library):
    lib = ctypes.CDLL(str(nat.LIB_PATH))
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header(library):
    declared = _declared() - {"fh_last_error", "fh_abi_version", "fh_device_info"}
    assert set(nat.exported_symbols()) == declared


def _declared_arity() -> dict[str, int]:
    """Parameter count of every declared entry point ((void) counts 0)."""

    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    out = {}
    for name, params in re.findall(r"\b(fh_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
        params = params.strip()
        out[name] = 0 if params in ("", "void") else len(params.split(","))
    return out


def test_ctypes_arity_matches_header():
    """Every ctypes signature has as many arguments as the C declaration
    (a missing or extra argument would shift every later one)."""

    arity = _declared_arity()
    wrong = {name: (len(types), arity.get(name)) for name, types in nat.SIGNATURES.items()
             if arity.get(name) != len(types)}
    assert not wrong, wrong


def test_abi_version(library):
    assert library.fh_abi_version() == nat.ABI_VERSION


def test_reference_schemas_registered():
    from fhelium_amd import torch_ops
    ops = torch_ops.registered_operators()
    assert len(ops) == len(set(ops)) == 58
    s = torch.ops.fhelium_ntt_ops.forward_ntt_montgomery_compact_keyswitch_accumulate_.default._schema
    assert str(s) == (
        "fhelium_ntt_ops::forward_ntt_montgomery_compact_keyswitch_accumulate_(Tensor(a!) "
        "coefficient_digit_qp, Tensor forward_twiddles, Tensor rns_params, Tensor key_digit_qp, "
        "Tensor(b!) accumulator0_qp, Tensor(c!) accumulator1_qp, int key_row_start) -> ()")
    s = torch.ops.fhelium_ckks_ops.keyswitch_moddown_qp_to_q.default._schema
    assert str(s) == (
        "fhelium_ckks_ops::keyswitch_moddown_qp_to_q(Tensor q_residues, Tensor p_residues, "
        "Tensor moddown_p_drop_inverses_montgomery, Tensor rns_params) -> Tensor")


def test_host_tensors_have_no_fallback():
    from fhelium_amd import torch_ops  # noqa: F401
    x = torch.zeros((1, 2, 8), dtype=torch.int64)
    params = torch.zeros((8, 2), dtype=torch.int64)
    with pytest.raises(NotImplementedError):
        torch.ops.fhelium_rns_ops.add_lazy(x, x, params)
    with pytest.raises(NotImplementedError):
        torch.ops.fhelium_ntt_ops.inverse_ntt_to_standard_compact_grouped_smem_(
            x, torch.zeros((2, 8), dtype=torch.int64), params, 0)


def test_build_manifest_records_hip_and_detects_mismatch(library):
    from fhelium_amd import manifest as M
    path = M.manifest_path(nat.LIB_PATH.parent)
    body = M.read_manifest(path)
    assert body["native_backends"] == ["cuda"] and body["offload_arch"] == "gfx950"
    assert body["hip_version"] and body["torch"]["version"] == torch.__version__
    assert body["library_abi"] == nat.ABI_VERSION
    assert M.compatibility_mismatches(body, lib_path=nat.LIB_PATH, library_abi=nat.ABI_VERSION) == []
    stale = dict(body, library_sha256="0" * 64)
    assert any("SHA-256" in p for p in M.compatibility_mismatches(stale, lib_path=nat.LIB_PATH,
                                                                   library_abi=nat.ABI_VERSION))
    other_torch = dict(body, torch=dict(body["torch"], version="0.0.0"))
    assert any("torch version" in p for p in M.compatibility_mismatches(other_torch, lib_path=nat.LIB_PATH,
                                                                         library_abi=nat.ABI_VERSION))
    bad = dict(body, native_backends=["cuda", "tpu"])
    assert any("native_backends" in p for p in M.compatibility_mismatches(bad, lib_path=nat.LIB_PATH,
                                                                           library_abi=nat.ABI_VERSION))
