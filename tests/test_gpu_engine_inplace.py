"""The reference engine's in-place and key-switch forms of the hot path
(fhelium/engine/ckks_engine.py:1131-1246, 1677-1706, 1752, 2068, 2170,
2333-2346): rescale with nearest and floor rounding against the oracle,
in-place rescale and rotation equal to their functional forms (in-place
rescale writes into the original allocation), switch_key equal to the key
switch inside a rotation, and the in-place additions and plaintext product."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd import kernels as K
from fhelium_amd.core import Ciphertext, CkksStateError
from fhelium_amd.engine import CkksEngine
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = CkksEngine(H.config(12), rng_seed=41)
    e.generate_secret_key()
    e.generate_relinearization_key()
    return e


def _ct(eng, level, batch=2, seed=0):
    x = np.random.default_rng(seed).uniform(-1, 1, (batch, eng.config.slots))
    return eng.encrypt(x, level=level)


def _oracle_rescale(eng, ct, nearest):
    t = H.host_tables(12)
    ids = t.layout.prime_ids(ct.level)
    inv = t.rescale_inverses[ct.level][: len(ids) - 1].numpy()
    params = eng.runtime.rns_params.cpu().numpy()
    sub = np.ascontiguousarray(params[:, ids[1]: ids[-1] + 1])
    half = eng.config.moduli[ids[0]] // 2 if nearest else 0
    outs = []
    for comp in ct.data.cpu().numpy():
        flat = comp.reshape(-1, comp.shape[-2], comp.shape[-1])
        r = O.rescale(flat[:, 1:], flat[:, 0], inv, sub, half, nearest)
        outs.append(r.reshape(*comp.shape[:-2], comp.shape[-2] - 1, comp.shape[-1]))
    return np.stack(outs)


@pytest.mark.parametrize("level", [0, 20])
@pytest.mark.parametrize("rounding", ["nearest", "floor"])
def test_rescale_roundings_match_oracle_and_in_place(eng, level, rounding):
    ct = _ct(eng, level, seed=level)
    got = eng.rescale_to_next_level(ct, rounding=rounding)
    want = _oracle_rescale(eng, ct, rounding == "nearest")
    np.testing.assert_array_equal(H.to_np(got.data), want)
    q0 = eng.config.moduli[eng.runtime.row_start(level)]
    assert got.level == level + 1 and got.scale == ct.scale / q0
    work = ct.clone()
    base = work.data.data_ptr()
    row_bytes = work.data.size(-1) * 8
    same = eng.rescale_to_next_level_(work, rounding=rounding)
    assert same is work and work.level == level + 1 and work.scale == got.scale
    assert work.data.data_ptr() == base + row_bytes  # rows 1.. of the original allocation
    assert torch.equal(work.data, got.data)


def test_rescale_rejects_bad_rounding(eng):
    with pytest.raises(ValueError, match="rounding"):
        eng.rescale_to_next_level(_ct(eng, 0), rounding="ceil")


def test_rotate_by_step_in_place(eng):
    ct = _ct(eng, 3, seed=5)
    want = eng.rotate_by_step(ct, 40)
    work = ct.clone()
    assert eng.rotate_by_step_(work, 40) is work
    assert torch.equal(work.data, want.data) and work.level == ct.level and work.scale == ct.scale


@pytest.mark.parametrize("level", [0, 25])
def test_switch_key_is_the_rotation_key_switch(eng, level):
    """rotate = coefficient automorphism, then switch_key with the rotation key."""

    ct = _ct(eng, level, seed=9 + level)
    key = eng.rotation_key(3)
    rt = eng.runtime
    idx, sgn = rt.coefficient_galois(key.galois_element)
    twice = rt.twice_modulus[rt.row_start(level): eng.config.num_q_primes]
    automorphed = K.apply_coefficient_galois_automorphism(ct.data, idx, sgn, twice)
    switched = eng.switch_key(Ciphertext(automorphed, ct.level, ct.scale), key)
    want = eng.rotate_with_key(ct, key)
    assert torch.equal(switched.data, want.data)
    assert switched.level == ct.level and switched.scale == ct.scale


def test_switch_key_rejects_bad_states(eng):
    ct = _ct(eng, 0)
    with pytest.raises(CkksStateError):
        eng.switch_key(eng.coefficient_domain_to_ntt_domain(ct), eng.relinearization_key)
    with pytest.raises(TypeError):
        eng.switch_key(ct, ct.data)


def test_in_place_additions_and_plaintext_product(eng):
    a, b = _ct(eng, 4, seed=1), _ct(eng, 4, seed=2)
    want = eng.add(a, b)
    work = a.clone()
    assert eng.add_(work, b) is work and torch.equal(work.data, want.data)
    x = np.random.default_rng(3).uniform(-1, 1, (2, eng.config.slots))
    pt = eng.encode(x, level=4)
    want = eng.add_plaintext(a, pt)
    work = a.clone()
    assert eng.add_plaintext_(work, pt) is work and torch.equal(work.data, want.data)
    pt_ntt = eng.encode(x, level=4, polynomial_domain="ntt")
    na = eng.coefficient_domain_to_ntt_domain(a)
    want = eng.multiply_plaintext(na, pt_ntt)
    work = na.clone()
    assert eng.multiply_plaintext_(work, pt_ntt) is work and torch.equal(work.data, want.data)
