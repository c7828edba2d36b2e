"""Keys and ciphertexts written to value files by one engine and read by a
fresh engine give bit-identical HMult and rotation results; a secret key
read back recovers its signed coefficients."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd.core import Ciphertext, RelinearizationKey, RotationKey, SecretKey
from fhelium_amd.engine import CkksEngine
from fhelium_amd.serialization import load_value, save_value
from tests import helpers as H

pytestmark = pytest.mark.gpu


def test_engine_state_round_trip(tmp_path):
    a = CkksEngine(H.config(12), rng_seed=51)
    a.generate_secret_key()
    a.generate_relinearization_key()
    a.rotation_key(7)
    x = np.random.default_rng(3).uniform(-1, 1, (2, a.config.slots))
    ct = a.encrypt(x, level=2)
    for name, value, kw in (("sk", a.secret_key, {"allow_secret": True}), ("rlk", a.relinearization_key, {}),
                            ("rot", a.rotation_keys[7], {}), ("ct", ct, {})):
        save_value(value, tmp_path / f"{name}.safetensors", context=a, **kw)
    b = CkksEngine(H.config(12), rng_seed=99, allow_sk_gen=False)
    b.set_secret_key(load_value(tmp_path / "sk.safetensors", context=b, device="cuda", expected_type=SecretKey))
    assert torch.equal(b.secret_key.coefficients.cpu(), a.secret_key.coefficients.cpu())
    b.set_relinearization_key(load_value(tmp_path / "rlk.safetensors", context=b, expected_type=RelinearizationKey))
    b.set_rotation_key(load_value(tmp_path / "rot.safetensors", context=b, expected_type=RotationKey))
    ct_b = load_value(tmp_path / "ct.safetensors", context=b, device="cuda", expected_type=Ciphertext)
    want = a.rotate_by_step(a.hmult(ct, ct), 7)
    got = b.rotate_by_step(b.hmult(ct_b, ct_b), 7)
    assert torch.equal(got.data, want.data) and got.scale == want.scale and got.level == want.level
    assert np.abs(b.decrypt(got).real - np.roll(x * x, 7, axis=-1)).max() < 3e-5
