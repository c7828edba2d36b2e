"""HIP-graph replay of a batched HMult + rotation equals eager execution bit
for bit, on several replays with different inputs and eager work between
them (fhelium/execution/cuda_graph.py:58-560 is the reference's replay)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd.engine import CkksEngine
from fhelium_amd.graph import CapturedStep
from tests import helpers as H

pytestmark = pytest.mark.gpu


def test_captured_hmult_rotation_replays_bit_exact():
    eng = CkksEngine(H.config(12), rng_seed=31)
    eng.generate_secret_key()
    eng.generate_relinearization_key()
    steps = [1, 5, 9]
    for s in steps:
        eng.rotation_key(s)
    level = 3
    rng = np.random.default_rng(3)

    def pair():
        x = rng.uniform(-1, 1, size=(3, eng.config.slots))
        y = rng.uniform(-1, 1, size=(3, eng.config.slots))
        return eng.encrypt(x, level=level), eng.encrypt(y, level=level), x, y

    def step(a, b):
        return eng.rotate_by_steps(eng.hmult(a, b), steps)

    a0, b0, _, _ = pair()
    captured = CapturedStep(step, a0, b0)
    for replay in range(3):
        a, b, x, y = pair()
        got = captured(a, b).data.clone()
        # eager work between replays, then the eager result of the same inputs
        want = step(a, b)
        torch.cuda.synchronize()
        assert torch.equal(got, want.data), f"replay {replay} differs from eager"
        dec = eng.decrypt(want).real
        for i, s in enumerate(steps):
            assert np.abs(dec[i] - np.roll(x[i] * y[i], s)).max() < 3e-5


def test_replay_rejects_state_mismatch():
    eng = CkksEngine(H.config(12), rng_seed=32)
    eng.generate_secret_key()
    x = np.zeros((2, eng.config.slots))
    a = eng.encrypt(x, level=2)
    captured = CapturedStep(lambda c: eng.coefficient_domain_to_ntt_domain(c), a)
    with pytest.raises(ValueError):
        captured(eng.encrypt(x, level=3))
