"""The library's device runtime on gfx950: the vector copy kernel, the arena,
events and stream-captured graphs of library launches (replayed on new
inputs, each replay equal to the eager result bit for bit)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd import device as D
from fhelium_amd import kernels as K
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbytes", [16, 8 * 1001, (1 << 20) + 3, 64 << 20])
def test_stream_copy(nbytes):
    print(f"stream copy {nbytes} bytes", flush=True)
    src = torch.randint(0, 256, (nbytes,), dtype=torch.uint8, device="cuda")
    dst = torch.zeros_like(src)
    D.stream_copy(dst, src)
    torch.cuda.synchronize()
    assert torch.equal(dst, src)
    with pytest.raises(ValueError):
        D.stream_copy(dst[:-1], src)


def test_arena_alignment_capacity_and_reset():
    with D.Arena(1 << 20) as ar:
        a = ar.alloc((100,))
        b = ar.alloc((10,), alignment=4096)
        assert a.data_ptr() % 256 == 0 and b.data_ptr() % 4096 == 0
        assert b.data_ptr() >= a.data_ptr() + 800
        a.fill_(7)
        b.fill_(-1)
        torch.cuda.synchronize()
        assert int(a.sum()) == 700 and int(b.sum()) == -10
        with pytest.raises(RuntimeError, match="arena"):
            ar.alloc((1 << 20,))
        st = ar.stats()
        assert st["capacity"] == 1 << 20 and st["used"] == st["high_water"] > 0
        first = a.data_ptr()
        ar.reset()
        assert ar.alloc((4,)).data_ptr() == first
        assert ar.stats()["high_water"] == st["high_water"]


def test_events_time_a_copy():
    src = torch.ones(1 << 24, dtype=torch.int64, device="cuda")
    dst = torch.empty_like(src)
    e0, e1 = D.Event(), D.Event()
    e0.record()
    D.stream_copy(dst, src)
    e1.record()
    assert e0.elapsed_ms(e1) > 0.0


def test_native_graph_replays_library_launches():
    cfg = H.config(12)
    rows = 6
    params = H.host_tables(12).rns_params[:, :rows].contiguous().cuda()
    moduli = H.moduli_array(cfg, 0, rows)
    rng = np.random.default_rng(5)
    lhs = H.to_dev(H.random_residues(rng, 4, moduli, cfg.N, lazy=True).reshape(2, 2, rows, cfg.N))
    rhs = H.to_dev(H.random_residues(rng, 4, moduli, cfg.N, lazy=True).reshape(2, 2, rows, cfg.N))
    out = torch.empty((3, 2, rows, cfg.N), dtype=torch.int64, device="cuda")
    copy = torch.empty_like(out)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    g = D.NativeGraph()
    with torch.cuda.stream(stream):
        g.begin(stream)
        K.tensor_product(out, lhs, rhs, params)
        K.add_canonical_(out[1], out[0], params)
        D.stream_copy(copy, out)
        g.end()
    for rep in range(3):
        new_l = H.to_dev(H.random_residues(rng, 4, moduli, cfg.N, lazy=True).reshape(2, 2, rows, cfg.N))
        new_r = H.to_dev(H.random_residues(rng, 4, moduli, cfg.N, lazy=True).reshape(2, 2, rows, cfg.N))
        want = torch.empty_like(out)
        K.tensor_product(want, new_l, new_r, params)
        K.add_canonical_(want[1], want[0], params)
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            lhs.copy_(new_l)
            rhs.copy_(new_r)
            g.replay(stream)
        stream.synchronize()
        assert torch.equal(copy, want), f"replay {rep}"
        assert torch.equal(out, want), f"replay {rep}"
