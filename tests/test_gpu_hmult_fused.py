"""The fused HMult front (fh_ckks_hmult_tensor_coefficient) against the
composition it replaces: coefficient_domain_to_ntt_domain of both inputs,
multiply, and the inverse NTT that starts relinearize
(ckks_engine.py:876-903, 1530-1606, 1608-1640), bit for bit; and the whole
fused hmult against rescale(relinearize(multiply(...)))."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd.engine import CkksEngine
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[12, 16])
def eng(request):
    e = CkksEngine(H.config(request.param), rng_seed=61)
    e.generate_secret_key()
    e.generate_relinearization_key()
    return e


def composed(eng, a, b):
    return eng.multiply(eng.coefficient_domain_to_ntt_domain(a), eng.coefficient_domain_to_ntt_domain(b))


@pytest.mark.parametrize("level,batch", [(0, (3,)), (17, ()), (32, (2,))])
def test_tensor_product_matches_composition(eng, level, batch):
    rng = np.random.default_rng(level)
    x = rng.uniform(-1, 1, batch + (eng.config.slots,))
    y = rng.uniform(-1, 1, batch + (eng.config.slots,))
    a, b = eng.encrypt(x, level=level), eng.encrypt(y, level=level)
    want = eng.runtime.ntt.inverse_to_standard(composed(eng, a, b).data, eng.runtime.row_start(level))
    got = eng.key_switcher.tensor_product_coefficient(a.data, b.data, level)
    assert torch.equal(got, want)
    fused = eng.hmult(a, b)
    ref = eng.rescale_to_next_level(eng.relinearize(composed(eng, a, b)))
    assert torch.equal(fused.data, ref.data)
    assert fused.scale == ref.scale and fused.level == ref.level == level + 1
    assert np.abs(eng.decrypt(fused).real - x * y).max() < 2e-5


def test_tensor_product_on_strided_inputs(eng):
    x = np.random.default_rng(5).uniform(-1, 1, (4, eng.config.slots))
    ct = eng.encrypt(x, level=5)
    a = ct.data[:, ::2]  # batch-strided views
    b = ct.data[:, 1::2]
    want = eng.runtime.ntt.inverse_to_standard(
        eng.multiply(eng.coefficient_domain_to_ntt_domain(type(ct)(a.contiguous(), 5, ct.scale)),
                     eng.coefficient_domain_to_ntt_domain(type(ct)(b.contiguous(), 5, ct.scale))).data,
        eng.runtime.row_start(5))
    assert torch.equal(eng.key_switcher.tensor_product_coefficient(a, b, 5), want)


@pytest.mark.parametrize("level", [0, 9, 31])
def test_keyswitch_with_fused_rescale(eng, level):
    """key_switch_fused(..., rescale=True) equals the un-rescaled key switch
    followed by rescale_to_next_level, for a rotation-shaped call (Galois
    gather, base0 only) and a bare call (no bases)."""

    from fhelium_amd.core import Ciphertext

    x = np.random.default_rng(level + 100).uniform(-1, 1, (2, eng.config.slots))
    ct = eng.encrypt(x, level=level)
    key = eng.rotation_key(3)
    ks = eng.key_switcher
    g = [key.galois_element] * 2
    plain = ks.key_switch_fused(ct.data[1], key, level, base0=ct.data[0], galois_elements=g)
    fused = ks.key_switch_fused(ct.data[1], key, level, base0=ct.data[0], galois_elements=g, rescale=True)
    want = eng.rescale_to_next_level(Ciphertext(plain, level, ct.scale))
    assert torch.equal(fused, want.data)
    bare = ks.key_switch_fused(ct.data[1], key, level)
    bare_r = ks.key_switch_fused(ct.data[1], key, level, rescale=True)
    assert torch.equal(bare_r, eng.rescale_to_next_level(Ciphertext(bare, level, ct.scale)).data)


@pytest.mark.parametrize("level,rescale", [(0, True), (12, False), (30, True)])
def test_seeded_keyswitch_matches_bases(eng, level, rescale):
    """d0, d1 left in the NTT domain and added through the accumulators
    (floor((X + P d) / P) = floor(X / P) + d) give the same key switch, bit
    for bit, as adding the coefficient-domain d0, d1 after ModDown."""
    rng = np.random.default_rng(100 + level)
    x = rng.uniform(-1, 1, (3, eng.config.slots))
    y = rng.uniform(-1, 1, (3, eng.config.slots))
    a, b = eng.encrypt(x, level=level), eng.encrypt(y, level=level)
    ks = eng.key_switcher
    key = eng.relinearization_key
    d = ks.tensor_product_coefficient(a.data, b.data, level)
    seeds, d2 = ks.tensor_product_seeded(a.data, b.data, level)
    assert torch.equal(d2, d[2])
    want = ks.key_switch_fused(d[2], key, level, base0=d[0], base1=d[1], rescale=rescale)
    got = ks.key_switch_fused(d2, key, level, rescale=rescale, seeds=seeds)
    assert torch.equal(got, want)
    # plain residues are not seeds, and seeds carry their level
    with pytest.raises(TypeError):
        ks.key_switch_fused(d2, key, level, rescale=rescale, seeds=d[:2])
