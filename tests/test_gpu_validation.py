"""Operand checks of the fused key switch and the fused tensor product: a
wrong dtype, device, batch or row count raises ValueError on the host before
any kernel is launched (the reference's TORCH_CHECKs before launch,
keyswitch_cuda.cu:159-168)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd.engine import CkksEngine
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = CkksEngine(H.config(12), rng_seed=5)
    e.generate_secret_key()
    e.generate_relinearization_key()
    return e


def _rows(eng, level, batch, seed):
    cfg = eng.config
    m = H.moduli_array(cfg, level, cfg.num_q_primes - level)
    return H.to_dev(H.random_residues(np.random.default_rng(seed), batch, m, cfg.N))


def test_keyswitch_rejects_bad_operands(eng):
    ks = eng.key_switcher
    key = eng.relinearization_key
    level = 4
    comp = _rows(eng, level, 3, 1)
    base = _rows(eng, level, 3, 2)
    good = ks.key_switch_fused(comp, key, level, base0=base)
    torch.cuda.synchronize()
    bad = {
        "short base batch": dict(component=comp, base0=base[:2]),
        "int32 base": dict(component=comp, base0=base.to(torch.int32)),
        "host base": dict(component=comp, base0=base.cpu()),
        "int32 component": dict(component=comp.to(torch.int32)),
        "wrong row count": dict(component=comp[:, 1:]),
        "host base1": dict(component=comp, base1=base.cpu()),
    }
    for name, kw in bad.items():
        with pytest.raises(ValueError):
            ks.key_switch_fused(kw.pop("component"), key, level, **kw)
    # nothing ran: the good result is unchanged by the rejected calls
    again = ks.key_switch_fused(comp, key, level, base0=base)
    assert torch.equal(good, again)


def test_tensor_product_rejects_bad_dtype(eng):
    ks = eng.key_switcher
    level = 2
    a = torch.stack([_rows(eng, level, 1, 3), _rows(eng, level, 1, 4)])
    with pytest.raises(ValueError):
        ks.tensor_product_coefficient(a, a.to(torch.int32), level)
