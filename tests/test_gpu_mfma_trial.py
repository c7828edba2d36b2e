"""The byte-sliced INT8 MFMA basis extension (fh_rns_basis_extend_mfma)
against the scalar kernel and the oracle: bit-identical outputs for 1..4
digits, real digit tables of the L34 preset, logN 12 and 16."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd import kernels as K
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("log_n", [12, 16])
def test_mfma_basis_extension_matches_scalar(log_n):
    cfg = H.config(log_n)
    t = H.host_tables(log_n)
    params = t.rns_params.cuda()
    rng = np.random.default_rng(log_n)
    seen = set()
    for ids, dt in t.digit_tables.items():
        d = len(ids)
        if d in seen or d > 4:
            continue
        seen.add(d)
        print(f"mfma basis extension logN={log_n} digits={d}", flush=True)
        mods = np.array([cfg.moduli[i] for i in ids], dtype=np.int64)
        digits = H.random_residues(rng, 2, mods, cfg.N)
        dev = H.to_dev(digits)
        coef = dt.basis_extension.cuda()
        rows = params.size(1)
        scalar = K.mixed_radix_basis_extend_to_montgomery(dev, coef, params, rows)
        mfma = K.mixed_radix_basis_extend_to_montgomery(dev, coef, params, rows, matrix_cores=True)
        assert torch.equal(scalar, mfma), f"{d} digits"
        if log_n == 12:
            want = O.basis_extend(digits, dt.basis_extension.numpy(), t.rns_params.numpy(), rows)
            np.testing.assert_array_equal(H.to_np(mfma), want)
    assert seen >= {1, 2, 4}
