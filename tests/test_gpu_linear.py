"""Diagonal-method matrix-vector product with batched per-entry-key
rotations (fhelium_amd/linear.py) against binary64, within K times the
reference's per-operation tolerance (tests/test_ckks_operation_correctness.py:48-51)."""

from __future__ import annotations

import numpy as np
import pytest

from fhelium_amd.engine import CkksEngine
from fhelium_amd.linear import diagonal_matvec, prepare_diagonals
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,chunk,level", [(16, 32, 0), (21, 8, 5)])
def test_diagonal_matvec(k, chunk, level):
    eng = CkksEngine(H.config(12), rng_seed=81)
    eng.generate_secret_key()
    rng = np.random.default_rng(k)
    x = rng.uniform(-1, 1, eng.config.slots)
    diags = rng.uniform(-1, 1, (k, eng.config.slots))
    ct = eng.encrypt(x, level=level)
    pts = prepare_diagonals(eng, diags, level)
    out = diagonal_matvec(eng, ct, pts, chunk=chunk)
    want = sum(diags[i] * np.roll(x, i) for i in range(k))
    assert out.level == level + 1
    assert np.abs(eng.decrypt(out).real - want).max() < k * 2e-5
