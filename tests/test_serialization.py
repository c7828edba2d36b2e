"""Value files in the reference's safetensors format
(fhelium/serialization/safetensors.py, value.py): round trips of every value
type, the manifest layout, and the reference tests' policy and tamper cases
(tests/test_serialization.py:166-257 of the reference)."""

from __future__ import annotations

import json
import re
import stat

import pytest
import torch
from safetensors import safe_open
from safetensors.torch import load_file, save_file

from fhelium_amd.core import (Ciphertext, ConjugationKey, KeySwitchKey, Plaintext, RelinearizationKey,
                              RotationKey, SecretKey)
from fhelium_amd.serialization import inspect_value, load_value, save_value
from tests import helpers as H

CFG = H.config(12)
N = 8  # payload width: the format does not depend on it


def _ct(level=3, basis="Q"):
    rows = CFG.num_q_primes - level + (CFG.num_p_primes if basis == "QP" else 0)
    data = torch.randint(0, 1 << 40, (2, 3, rows, N), generator=torch.Generator().manual_seed(level))
    return Ciphertext(data, level, 2.0 ** 40 * 1.25, "coefficient", "standard", basis)


def _key(cls, **kw):
    data = torch.randint(0, 1 << 40, (4, 2, CFG.total_num_primes, N), generator=torch.Generator().manual_seed(7))
    return cls(data, **kw)


def _metadata(path):
    with safe_open(str(path), framework="pt") as h:
        return dict(h.metadata())


VALUES = {
    "ciphertext": lambda: _ct(),
    "ciphertext-qp": lambda: _ct(2, "QP"),
    "ntt-plaintext": lambda: Plaintext(torch.arange(33 * N, dtype=torch.int64).reshape(33, N), 2,
                                       2.0 ** 40, "ntt", "montgomery"),
    "rotation-key": lambda: _key(RotationKey, rotation_step=-5, galois_element=0),
    "relinearization-key": lambda: _key(RelinearizationKey),
    "key-switch-key": lambda: _key(KeySwitchKey),
    "conjugation-key": lambda: _key(ConjugationKey),
}


@pytest.mark.parametrize("name", list(VALUES))
def test_round_trip(tmp_path, name):
    value = VALUES[name]()
    path = tmp_path / "value.bin"
    meta = save_value(value, path, context=CFG)
    assert meta.value_type == type(value).__name__ and meta.context_id == CFG.context_id
    assert meta.nbytes == value.data.numel() * 8
    assert inspect_value(path) == meta
    back = load_value(path, context=CFG, expected_type=type(value))
    assert type(back) is type(value) and torch.equal(back.data, value.data)
    for attr in ("level", "scale", "polynomial_domain", "residue_representation", "modulus_basis",
                 "rotation_step"):
        if hasattr(value, attr):
            assert getattr(back, attr) == getattr(value, attr)
    if isinstance(value, RotationKey):
        # Galois element re-derived from the step: 3^(-5) mod 2N's canonical form
        assert back.galois_element * pow(3, 5, 2 * CFG.N) % (2 * CFG.N) == 1


def test_manifest_layout(tmp_path):
    path = tmp_path / "ct.safetensors"
    save_value(_ct(), path, context=CFG)
    raw = _metadata(path)
    assert raw["fhelium.format"] == "fhelium-value" and raw["fhelium.schema_version"] == "1"
    manifest = json.loads(raw["fhelium.manifest"])
    assert raw["fhelium.manifest"] == json.dumps(manifest, sort_keys=True, separators=(",", ":"))
    assert manifest["file_schema_version"] == 1 and manifest["value_schema_version"] == 2
    assert manifest["tensor_metadata"] == {"data": {"shape": [2, 3, 32, N], "dtype": "torch.int64",
                                                    "encoding": "direct", "payloads": ["data"]}}
    assert manifest["value_metadata"]["prime_ids"] == list(range(3, 35))
    assert re.fullmatch("[0-9a-f]{64}", manifest["context_id"])
    assert H.config(16).context_id != CFG.context_id


def test_file_lifecycle_policy(tmp_path):
    value = _ct()
    path = tmp_path / "tenant" / "ct.safetensors"
    with pytest.raises(FileNotFoundError, match="parent directory"):
        save_value(value, path, context=CFG)
    path.parent.mkdir()
    save_value(value, path, context=CFG)
    with pytest.raises(FileExistsError):
        save_value(value, path, context=CFG)
    with pytest.raises(ValueError, match="context mismatch"):
        load_value(path, context=H.config(16))
    with pytest.raises(TypeError, match="expected Plaintext"):
        load_value(path, context=CFG, expected_type=Plaintext)
    replacement = Ciphertext(torch.ones_like(value.data), value.level, value.scale)
    save_value(replacement, path, context=CFG, overwrite=True)
    assert torch.equal(load_value(path, context=CFG).data, replacement.data)
    assert stat.S_IMODE(path.stat().st_mode) == 0o600
    assert [p.name for p in path.parent.iterdir()] == ["ct.safetensors"]  # no staged file left
    with pytest.raises(IsADirectoryError):
        save_value(value, path.parent, context=CFG, overwrite=True)


def test_failed_write_leaves_destination_alone(tmp_path, monkeypatch):
    """A write that fails part-way removes its staged file and leaves an
    existing destination as it was."""

    from fhelium_amd import serialization as S

    path = tmp_path / "ct.safetensors"
    save_value(_ct(), path, context=CFG)
    before = path.read_bytes()

    def broken(*args, **kwargs):
        raise OSError("disk full")

    monkeypatch.setattr(S, "save_file", broken)
    with pytest.raises(OSError, match="disk full"):
        save_value(Ciphertext(torch.ones_like(_ct().data), 3, 2.0 ** 40), path, context=CFG, overwrite=True)
    assert path.read_bytes() == before and [p.name for p in tmp_path.iterdir()] == ["ct.safetensors"]


def test_secret_key_needs_opt_in(tmp_path):
    sk = SecretKey(torch.arange(CFG.total_num_primes * N, dtype=torch.int64).reshape(-1, N), None)
    path = tmp_path / "secret.safetensors"
    with pytest.raises(PermissionError, match="disabled by default"):
        save_value(sk, path, context=CFG)
    assert save_value(sk, path, context=CFG, allow_secret=True).value_type == "SecretKey"
    assert torch.equal(load_value(path, context=CFG, expected_type=SecretKey).data, sk.data)


def _tamper(path, edit):
    tensors = load_file(path)
    meta = _metadata(path)
    manifest = json.loads(meta["fhelium.manifest"])
    edit(manifest)
    meta["fhelium.manifest"] = json.dumps(manifest)
    save_file(tensors, path, metadata=meta)


def test_tampered_files_are_rejected(tmp_path):
    path = tmp_path / "ct.safetensors"
    save_value(_ct(), path, context=CFG)
    _tamper(path, lambda m: m.__setitem__("value_schema_version", True))
    with pytest.raises(ValueError, match="schema version"):
        inspect_value(path)
    save_value(_ct(), path, context=CFG, overwrite=True)
    _tamper(path, lambda m: m["tensor_metadata"]["data"]["shape"].__setitem__(-1, N + 1))
    with pytest.raises(ValueError, match="payload header"):
        inspect_value(path)
    save_value(_ct(), path, context=CFG, overwrite=True)
    _tamper(path, lambda m: m["value_metadata"].__setitem__("level", 4))
    with pytest.raises(ValueError, match="prime_ids"):
        load_value(path, context=CFG)
