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


@pytest.mark.parametrize("k,chunk", [(16, 128), (21, 8)])
def test_diagonal_matvec_equals_reference_composition(k, chunk):
    """The hoisted rotations plus the one-pass inner product give the residues
    of the reference's composition (packed_matrix_vector.py:120-170):
    rotate_many_with_keys (hoisted), multiply_plaintext per diagonal, add,
    inverse NTT, rescale -- bit for bit."""
    import torch
    from fhelium_amd.core import Plaintext
    eng = CkksEngine(H.config(12), rng_seed=82)
    eng.generate_secret_key()
    rng = np.random.default_rng(100 + k)
    level = 2
    ct = eng.encrypt(rng.uniform(-1, 1, eng.config.slots), level=level)
    pts = prepare_diagonals(eng, rng.uniform(-1, 1, (k, eng.config.slots)), level)
    out = diagonal_matvec(eng, ct, pts, chunk=chunk)
    rots = [ct] + eng.rotate_many_with_keys(ct, [eng.rotation_key(s) for s in range(1, k)])
    acc = None
    for i, r in enumerate(rots):
        term = eng.multiply_plaintext(eng.coefficient_domain_to_ntt_domain(r),
                                      Plaintext(pts.data[i], level, pts.scale, "ntt"))
        acc = term if acc is None else eng.add(acc, term)
    want = eng.rescale_to_next_level(eng.ntt_domain_to_coefficient_domain(acc))
    assert torch.equal(out.data, want.data)
    assert out.scale == want.scale and out.level == want.level
