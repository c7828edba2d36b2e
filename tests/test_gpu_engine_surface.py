"""The rest of the reference engine's value surface on the hot path
(fhelium/engine/ckks_engine.py:987-1129, 1761-1869): residue representation
transitions of coefficient-domain plaintexts checked against r R = m (mod q)
with R = 2^62 on sampled residues, batch sums equal to entry-by-entry sums
bit for bit (odd counts included, along either batch axis), and the
rescale scale helpers equal to what rescale_to_next_level sets, and the
ciphertext and key validators."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd.core import Ciphertext, MaximumLevelError, Plaintext
from fhelium_amd.engine import CkksEngine
from tests import helpers as H

pytestmark = pytest.mark.gpu

R = 1 << 62  # Montgomery radix of every row


@pytest.fixture(scope="module")
def eng():
    e = CkksEngine(H.config(12), rng_seed=53)
    e.generate_secret_key()
    return e


@pytest.mark.parametrize("level", [0, 11])
def test_residue_transitions(eng, level):
    x = np.random.default_rng(level).uniform(-1, 1, (2, eng.config.slots))
    pt = eng.encode(x, level=level)
    assert (pt.polynomial_domain, pt.residue_representation) == ("coefficient", "montgomery")
    std = eng.montgomery_residues_to_standard_residues(pt)
    assert std.residue_representation == "standard" and pt.residue_representation == "montgomery"
    moduli = eng.config.moduli[eng.runtime.row_start(level): eng.config.num_q_primes]
    m, r = H.to_np(pt.data), H.to_np(std.data)
    rng = np.random.default_rng(1)
    for row in range(len(moduli)):
        q = int(moduli[row])
        for k in rng.integers(0, eng.config.N, 16):
            for b in range(2):
                assert 0 <= int(r[b, row, k]) < q
                assert (int(r[b, row, k]) * R - int(m[b, row, k])) % q == 0
    back = eng.standard_residues_to_montgomery_residues(std)
    qcol = torch.tensor(moduli, dtype=torch.int64, device=pt.data.device).view(-1, 1)
    assert torch.equal(back.data.remainder(qcol), pt.data.remainder(qcol))
    work = Plaintext(std.data.clone(), std.level, std.scale, residue_representation="standard")
    storage = work.data.data_ptr()
    assert eng.standard_residues_to_montgomery_residues_(work) is work
    assert work.residue_representation == "montgomery" and work.data.data_ptr() == storage
    assert torch.equal(work.data, back.data)
    assert eng.montgomery_residues_to_standard_residues_(work) is work and torch.equal(work.data, std.data)
    with pytest.raises(ValueError, match="standard residues"):
        eng.standard_residues_to_montgomery_residues(pt)
    with pytest.raises(ValueError, match="coefficient-domain"):
        eng.montgomery_residues_to_standard_residues(eng.encode(x, level=level, polynomial_domain="ntt"))
    with pytest.raises(TypeError):
        eng.montgomery_residues_to_standard_residues(eng.encrypt(x, level=level))


def test_batch_sums(eng):
    x = np.random.default_rng(3).uniform(-1, 1, (5, 3, eng.config.slots)) / 8
    ct = eng.encrypt(x, level=2)
    entries = [[Ciphertext(ct.data[:, i, j].contiguous(), ct.level, ct.scale) for j in range(3)] for i in range(5)]
    over0 = eng.sum_ciphertext_batch(ct)
    assert over0.batch_shape == (3,) and over0.level == ct.level and over0.scale == ct.scale
    for j in range(3):
        want = eng.sum_ciphertexts([entries[i][j] for i in range(5)])
        assert torch.equal(over0.data[:, j], want.data)
    over1 = eng.sum_ciphertext_batch(ct, dim=-1)
    assert over1.batch_shape == (5,)
    for i in range(5):
        assert torch.equal(over1.data[:, i], eng.sum_ciphertexts(entries[i]).data)
    assert np.abs(eng.decrypt(over0).real - x.sum(axis=0)).max() < 1e-5
    single = Ciphertext(ct.data[:, :1].contiguous(), ct.level, ct.scale)
    s1 = eng.sum_ciphertext_batch(single)
    assert torch.equal(s1.data, ct.data[:, 0]) and s1.data.data_ptr() != ct.data.data_ptr()
    with pytest.raises(IndexError):
        eng.sum_ciphertext_batch(ct, dim=2)
    with pytest.raises(ValueError):
        eng.sum_ciphertext_batch(entries[0][0])
    with pytest.raises(ValueError):
        eng.sum_ciphertexts([])


def test_scale_helpers(eng):
    ct = eng.encrypt(np.zeros(eng.config.slots), level=4)
    resc = eng.rescale_to_next_level(ct)
    assert eng.rescale_to_next_output_scale(ct.scale, level=4) == resc.scale
    assert eng.rescale_to_next_drop_prime(level=4) == eng.config.moduli[eng.runtime.row_start(4)]
    assert eng.final_public_level == eng.public_level_count - 1 and eng.num_slots == eng.config.slots
    with pytest.raises(MaximumLevelError):
        eng.rescale_to_next_drop_prime(level=eng.final_public_level)
    with pytest.raises(TypeError):
        eng.rescale_to_next_drop_prime(level=1.0)
    with pytest.raises(ValueError):
        eng.rescale_to_next_output_scale("2", level=0)


def test_validators(eng):
    """validate_ciphertext / validate_key_switch_key (ckks_engine.py:2560-2630)."""

    from fhelium_amd.core import KeySwitchKey, RotationKey

    ct = eng.encrypt(np.zeros(eng.config.slots), level=3)
    eng.validate_ciphertext(ct)
    bad = ct.clone()
    bad.level = 2  # one limb short for level 2
    with pytest.raises(ValueError, match="limbs"):
        eng.validate_ciphertext(bad)
    bad = ct.clone()
    bad.scale = -1.0
    with pytest.raises(ValueError):
        eng.validate_ciphertext(bad)
    with pytest.raises(ValueError):
        eng.validate_ciphertext(Ciphertext(ct.data.cpu(), ct.level, ct.scale))
    with pytest.raises(TypeError):
        eng.validate_ciphertext(ct.data)
    key = eng.rotation_key(5)
    eng.validate_key_switch_key(key)
    with pytest.raises(ValueError, match="Galois"):
        eng.validate_key_switch_key(RotationKey(key.data, rotation_step=5, galois_element=3))
    with pytest.raises(ValueError, match="canonical"):
        eng.validate_key_switch_key(RotationKey(key.data, rotation_step=5 + eng.config.slots, galois_element=3))
    with pytest.raises(ValueError, match="key material"):
        eng.validate_key_switch_key(KeySwitchKey(key.data[1:]))
    with pytest.raises(TypeError):
        eng.validate_key_switch_key(ct)
