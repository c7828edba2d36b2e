"""Construction checks of Ciphertext and Plaintext on the CPU, as the
reference's value model applies them (tests/test_batch_value_model.py:156-170,
fhelium/core/ciphertext.py:67-127, core/scale.py:11-29)."""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from fhelium_amd.core import Ciphertext, InvalidScaleError, Plaintext


def _ct(**kw):
    args = dict(data=torch.zeros(2, 3, 8, dtype=torch.int64), level=0, scale=float(2 ** 40))
    args.update(kw)
    return Ciphertext(**args)


def test_zero_sized_batch_dimensions_are_rejected():
    with pytest.raises(ValueError, match="batch dimensions must be nonzero"):
        _ct(data=torch.empty(2, 0, 3, 8, dtype=torch.int64))
    with pytest.raises(ValueError, match="batch dimensions must be nonzero"):
        Plaintext(torch.empty(0, 3, 8, dtype=torch.int64), 0, float(2 ** 40))


def test_valid_values_construct():
    c = _ct(data=torch.zeros(3, 2, 4, 3, 8, dtype=torch.int64), polynomial_domain="ntt",
            residue_representation="montgomery", modulus_basis="QP")
    assert c.component_count == 3 and c.batch_shape == (2, 4)
    p = Plaintext(torch.zeros(5, 3, 8, dtype=torch.int64), np.int64(2), np.float64(2.0 ** 40))
    assert p.level == 2 and type(p.level) is int and type(p.scale) is float


@pytest.mark.parametrize("bad, error", [
    (dict(data=torch.zeros(4, 3, 8, dtype=torch.int64)), ValueError),   # component axis
    (dict(data=torch.zeros(3, 8, dtype=torch.int64)), ValueError),      # too few axes
    (dict(data=torch.zeros(2, 0, 8, dtype=torch.int64)), ValueError),   # empty limb axis
    (dict(data=torch.zeros(2, 3, 0, dtype=torch.int64)), ValueError),   # empty coefficient axis
    (dict(data=torch.zeros(2, 3, 8)), TypeError),                       # floating point
    (dict(data=torch.zeros(2, 3, 8, dtype=torch.bool)), TypeError),
    (dict(data=np.zeros((2, 3, 8), dtype=np.int64)), TypeError),        # not a tensor
    (dict(level=-1), ValueError),
    (dict(level=True), ValueError),
    (dict(level=1.5), ValueError),
    (dict(polynomial_domain="slots"), ValueError),
    (dict(modulus_basis="P"), ValueError),
    (dict(residue_representation="lazy"), ValueError),
    (dict(polynomial_domain="ntt"), ValueError),                        # NTT values are Montgomery
    (dict(residue_representation="montgomery"), ValueError),           # coefficient values are standard
])
def test_ciphertext_rejects(bad, error):
    with pytest.raises(error):
        _ct(**bad)


@pytest.mark.parametrize("scale", [0.0, -1.0, math.inf, math.nan, True, "2", None])
def test_scales_must_be_positive_and_finite(scale):
    with pytest.raises(InvalidScaleError):
        _ct(scale=scale)
    with pytest.raises(InvalidScaleError):
        Plaintext(torch.zeros(3, 8, dtype=torch.int64), 0, scale)
