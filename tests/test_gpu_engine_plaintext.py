"""Plaintext operations, ciphertext add and the hoisted-rotation
matrix-vector product on gfx950, checked by decryption against binary64
within the reference's tolerances (tests/test_ckks_operation_correctness.py:48-51),
plus bit-exact identities the kernels must satisfy."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd.core import CkksStateError, ScaleMismatchError
from fhelium_amd.engine import CkksEngine
from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL = 2e-5


@pytest.fixture(scope="module")
def eng():
    e = CkksEngine(H.config(12), rng_seed=11)
    e.generate_secret_key()
    e.generate_relinearization_key()
    return e


def _vals(eng, seed, batch=()):
    return np.random.default_rng(seed).uniform(-1, 1, size=batch + (eng.config.slots,))


@pytest.mark.parametrize("level", [0, 20])
def test_add_plaintext_and_add(eng, level):
    x, y = _vals(eng, 1, (2,)), _vals(eng, 2, (2,))
    ct = eng.encrypt(x, level=level)
    pt = eng.encode(y, level=level)
    out = eng.add_plaintext(ct, pt)
    assert np.abs(eng.decrypt(out).real - (x + y)).max() < TOL
    assert torch.equal(out.c1, ct.c1)
    s = eng.add(ct, eng.encrypt(y, level=level))
    assert np.abs(eng.decrypt(s).real - (x + y)).max() < TOL
    # in place mutates c0 only
    before = ct.c1.clone()
    eng.add_plaintext(ct, pt, inplace=True)
    assert torch.equal(ct.c1, before)
    assert np.abs(eng.decrypt(ct).real - (x + y)).max() < TOL


def test_multiply_plaintext_then_rescale(eng):
    x, y = _vals(eng, 3, (3,)), _vals(eng, 4)
    level = 5
    ct = eng.coefficient_domain_to_ntt_domain(eng.encrypt(x, level=level))
    pt = eng.encode(y, level=level, polynomial_domain="ntt")  # unbatched: broadcasts
    prod = eng.multiply_plaintext(ct, pt)
    assert prod.scale == ct.scale * pt.scale
    back = eng.ntt_domain_to_coefficient_domain(prod)
    res = eng.rescale_to_next_level(back)
    assert np.abs(eng.decrypt(res).real - x * y[None]).max() < 3e-5


@pytest.mark.parametrize("level", [2, 17])
def test_constant_products_equal_multiply_plaintext(eng, level):
    """multiply_constants_and_add equals, bit for bit, multiply_plaintext by
    constant plaintexts in the NTT domain (each at the encoding scale that
    puts its product at one scale) summed with add, for terms at three
    levels read at the highest; the combination runs in the coefficient
    domain with no NTT."""
    vals = [0.37, -1.25, 0.0625]
    levels = [level, level + 1, level + 2]
    cts = [eng.encrypt(_vals(eng, 20 + k, (2,)), level=lv) for k, lv in enumerate(levels)]
    top = max(levels)
    q = eng.config.moduli[eng.runtime.row_start(top)]
    product_scale = eng.config.default_scale * q
    got = eng.multiply_constants_and_add(cts, vals, product_scale=product_scale)
    assert got.level == top and got.scale == product_scale
    want = None
    for v, ct in zip(vals, cts):
        at = eng.mod_switch_to_level(ct, top)
        pt = eng.encode(np.full(eng.config.slots, v), level=top, scale=product_scale / ct.scale,
                        polynomial_domain="ntt")
        prod = eng.ntt_domain_to_coefficient_domain(
            eng.multiply_plaintext(eng.coefficient_domain_to_ntt_domain(at), pt))
        prod = eng.reinterpret_at_scale(prod, product_scale, max_relative_change=1e-12)
        want = prod if want is None else eng.add(want, prod)
    assert torch.equal(got.data, want.data)
    res = eng.rescale_to_next_level(got)
    xs = [_vals(eng, 20 + k, (2,)) for k in range(3)]
    expect = sum(v * x for v, x in zip(vals, xs))
    assert np.abs(eng.decrypt(res).real - expect).max() < 3e-5


def test_plaintext_state_errors(eng):
    x = _vals(eng, 5)
    ct = eng.encrypt(x, level=2)
    with pytest.raises(ScaleMismatchError):
        eng.add_plaintext(ct, eng.encode(x, level=2, scale=2.0**30))
    with pytest.raises(CkksStateError):
        eng.add_plaintext(ct, eng.encode(x, level=2, polynomial_domain="ntt"))
    with pytest.raises(ValueError):
        eng.add_plaintext(ct, eng.encode(x, level=3))
    with pytest.raises(CkksStateError):
        eng.multiply_plaintext(ct, eng.encode(x, level=2, polynomial_domain="ntt"))


def test_ntt_round_trip_exact(eng):
    ct = eng.encrypt(_vals(eng, 6, (2,)), level=4)
    back = eng.ntt_domain_to_coefficient_domain(eng.coefficient_domain_to_ntt_domain(ct))
    assert torch.equal(back.data, ct.data)


def test_hoisted_matvec(eng):
    """Diagonal-method matrix-vector product: y = sum_i d_i * rot(x, i) with
    hoisted rotations for steps 1..15 and one rescale (the keyswitch-heavy
    matrix-vector workload at a reduced ring and diagonal count)."""
    steps = list(range(16))
    keys = [eng.rotation_key(s) if s else None for s in steps]
    x = _vals(eng, 7)
    diags = [_vals(eng, 100 + i) for i in steps]
    level = 1
    ct = eng.encrypt(x, level=level)
    nonzero = [k for k in keys if k is not None]
    rotated = eng.rotate_many_with_keys(ct, nonzero)
    acc = None
    for i, rct in zip(steps, [ct] + rotated):
        pt = eng.encode(diags[i], level=level, polynomial_domain="ntt")
        term = eng.multiply_plaintext(eng.coefficient_domain_to_ntt_domain(rct), pt)
        acc = term if acc is None else eng.add(acc, term)
    res = eng.rescale_to_next_level(eng.ntt_domain_to_coefficient_domain(acc))
    want = sum(d * np.roll(x, s) for s, d in zip(steps, diags))
    err = np.abs(eng.decrypt(res).real - want).max()
    assert err < 16 * TOL, err
