"""Independent batched calls on several streams (fhelium_amd.schedule):
the results equal the one-stream run bit for bit, for two, three and four
streams, on level groups of different sizes (logN=12, L34)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd.engine import CkksEngine
from fhelium_amd.schedule import run_on_streams
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    eng = CkksEngine(H.config(12), rng_seed=29)
    eng.generate_secret_key()
    eng.generate_relinearization_key()
    rng = np.random.default_rng(3)
    groups = []
    for level, count in ((0, 6), (9, 3), (20, 2), (31, 1), (32, 4)):
        x = rng.uniform(-1, 1, (count, eng.config.slots))
        y = rng.uniform(-1, 1, (count, eng.config.slots))
        steps = [int(s) for s in rng.integers(1, 200, size=count)]
        groups.append((eng.encrypt(x, level=level), eng.encrypt(y, level=level), steps))
    return eng, groups


@pytest.mark.parametrize("streams", [2, 3, 4])
def test_streams_equal_one_stream(setup, streams):
    eng, groups = setup

    def tasks():
        return [lambda a=a, b=b, s=s: eng.rotate_by_steps(eng.hmult(a, b), s) for a, b, s in groups]

    want = run_on_streams(tasks(), 1)
    got = run_on_streams(tasks(), streams)
    torch.cuda.synchronize()
    for w, g in zip(want, got):
        assert torch.equal(w.data, g.data)
        assert w.level == g.level and w.scale == g.scale
    # the caller's stream is ordered after every task: work issued on it
    # afterwards sees the results
    total = sum(int(g.data[..., :1].sum().item()) for g in got)
    assert total == sum(int(w.data[..., :1].sum().item()) for w in want)
