"""fh_keyswitch_fused_static (workspaces in the library's static device
scratch, no allocation) against fh_keyswitch_fused (caller workspaces):
bit-identical outputs on two calls in a row, the second smaller, for a
relinearization with rescale and seeds, a per-entry-key rotation and a
plain key switch; a call beyond the scratch size is rejected before any
launch."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd.engine import CkksEngine
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng14():
    eng = CkksEngine(H.config(14), rng_seed=99)
    eng.generate_secret_key()
    eng.generate_relinearization_key()
    for s in (1, -5, 300):
        eng.rotation_key(s)
    return eng


def _ct(eng, level, batch, seed):
    x = np.random.default_rng(seed).uniform(-1, 1, (batch, eng.config.slots))
    return eng.encrypt(x, level=level)


def test_static_matches_allocating_two_calls(eng14):
    eng = eng14
    ks = eng.key_switcher
    # first call: level 0, one ciphertext (the largest that fits at N = 2^14);
    # second call: level 9, a smaller workspace
    for level, batch in ((0, 1), (9, 1)):
        print(f"static key switch logN=14 level={level} batch={batch}", flush=True)
        a, b = _ct(eng, level, batch, 10 + level), _ct(eng, level, batch, 20 + level)
        seeds, d2 = ks.tensor_product_seeded(a.data, b.data, level)
        want = ks.key_switch_fused(d2, eng.relinearization_key, level, rescale=True, seeds=seeds)
        got = ks.key_switch_fused(d2, eng.relinearization_key, level, rescale=True, seeds=seeds,
                                  static_workspace=True)
        torch.cuda.synchronize()
        assert torch.equal(got, want), f"relinearization level {level}"
        c = _ct(eng, level, batch, 30 + level)
        keys = [eng.rotation_key(300)] * batch
        g = [k.galois_element for k in keys]
        want = ks.key_switch_fused(c.data[1], keys, level, base0=c.data[0], galois_elements=g)
        got = ks.key_switch_fused(c.data[1], keys, level, base0=c.data[0], galois_elements=g,
                                  static_workspace=True)
        torch.cuda.synchronize()
        assert torch.equal(got, want), f"rotation level {level}"
        want = ks.key_switch_fused(c.data[1], eng.relinearization_key, level)
        got = ks.key_switch_fused(c.data[1], eng.relinearization_key, level, static_workspace=True)
        torch.cuda.synchronize()
        assert torch.equal(got, want), f"plain key switch level {level}"


def test_static_batch_at_high_level(eng14):
    eng = eng14
    ks = eng.key_switcher
    level, batch = 28, 4
    print(f"static key switch logN=14 level={level} batch={batch} per-entry keys", flush=True)
    c = _ct(eng, level, batch, 41)
    keys = [eng.rotation_key(s) for s in (1, -5, 300, 1)]
    g = [k.galois_element for k in keys]
    want = ks.key_switch_fused(c.data[1], keys, level, base0=c.data[0], galois_elements=g)
    got = ks.key_switch_fused(c.data[1], keys, level, base0=c.data[0], galois_elements=g, static_workspace=True)
    torch.cuda.synchronize()
    assert torch.equal(got, want)


def test_static_rejects_oversize(eng14):
    eng = eng14
    c = _ct(eng, 0, 2, 51)  # two level-0 ciphertexts need about 126 MiB
    with pytest.raises(RuntimeError, match="static scratch holds"):
        eng.key_switcher.key_switch_fused(c.data[1], eng.relinearization_key, 0, static_workspace=True)
