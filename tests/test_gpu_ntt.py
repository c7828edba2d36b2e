"""gfx950 NTT / inverse NTT vs the CPU oracle, bit for bit.

The oracle restates the reference's indexed radix-2 CPU transform
(csrc/ops/ntt/cpu/indexed_ntt_cpu.cpp).  In exact mode the GPU kernels must
reproduce its lazy [0, 2q) forward outputs exactly (not merely mod q); in
fast mode the forward output must be exactly the canonical representative
of the oracle's value.  Inverse outputs are canonical in the reference and
must match bit for bit in both modes.  Rows cover the 40/41-bit primes
(binary64 path) and the 60-bit base/P primes (integer path).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd import ntt as gntt
from fhelium_amd import rns
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _backend(log_n: int, mode: str = "exact"):
    cfg = H.config(log_n)
    t = H.host_tables(log_n)
    fwd, inv = rns.compact_twiddles_montgomery(cfg.moduli, log_n)
    params = t.rns_params
    tables = gntt.NttTables(params.cuda(), fwd.cuda(), inv.cuda(), log_n)
    return cfg, params.numpy(), fwd.numpy(), inv.numpy(), gntt.Gfx950NttBackend(tables, mode)


def _canonical(x: np.ndarray, moduli: np.ndarray) -> np.ndarray:
    return x % moduli[None, :, None]


@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("log_n", [4, 9, 11, 12, 13, 14, 15, 16])
@pytest.mark.parametrize("to_mont", [False, True])
def test_forward_bit_exact(log_n, to_mont, mode):
    cfg, params, fwd, _, be = _backend(log_n, mode)
    rng = np.random.default_rng(1000 + log_n)
    n = cfg.N
    # all 39 QP rows: 35 binary64 rows + 1 base + 4 P (integer path)
    x = H.random_residues(rng, 2, H.moduli_array(cfg), n, lazy=not to_mont)
    dev = H.to_dev(x)
    if to_mont:
        be.forward_to_montgomery_(dev, 0)
    else:
        be.forward_montgomery_(dev, 0)
    want = O.ntt_forward(x, fwd, params, to_montgomery=to_mont)
    if mode == "fast":
        want = _canonical(want, H.moduli_array(cfg))
    np.testing.assert_array_equal(H.to_np(dev), want)


@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("log_n", [4, 9, 11, 12, 14, 16])
@pytest.mark.parametrize("out_form", [0, 1, 2, 3])
def test_inverse_bit_exact(log_n, out_form, mode):
    cfg, params, _, inv, be = _backend(log_n, mode)
    rng = np.random.default_rng(2000 + log_n * 4 + out_form)
    x = H.random_residues(rng, 2, H.moduli_array(cfg), cfg.N, lazy=True)
    dev = H.to_dev(x)
    {0: be.inverse_montgomery_, 1: be.inverse_to_standard_lazy_, 2: be.inverse_to_standard_,
     3: be.inverse_to_centered_}[out_form](dev, 0)
    want = O.ntt_inverse(x, inv, params, out_form)
    np.testing.assert_array_equal(H.to_np(dev), want)


@pytest.mark.parametrize("schedule", [None, (1, 2), (2, 3)])
def test_row_start_and_strided_batch(schedule):
    """A level-5 QP interval (rows 5..38) inside a strided batch view, as one
    launch or cut into chunks of entries on several streams (the autotuner's
    schedules): identical results."""
    from fhelium_amd.autotune import ONE_LAUNCH, NttSchedule
    log_n = 13
    cfg, params, fwd, inv, be = _backend(log_n)
    sched = ONE_LAUNCH if schedule is None else NttSchedule(*schedule)
    rng = np.random.default_rng(7)
    start = 5
    mods = H.moduli_array(cfg, start)
    x = H.random_residues(rng, 3, mods, cfg.N)
    big = torch.zeros((3, 2, len(mods), cfg.N), dtype=torch.int64, device="cuda")
    big[:, 1] = H.to_dev(x)
    view = big[:, 1]
    be._apply(view, start, True, 1, sched)
    sub = slice(start, None)
    want = O.ntt_forward(x, np.ascontiguousarray(fwd[sub]), np.ascontiguousarray(params[:, sub]), True)
    np.testing.assert_array_equal(H.to_np(view), want)
    assert int(big[:, 0].abs().sum()) == 0
    be._apply(view, start, False, 2, sched)
    back = O.ntt_inverse(want, np.ascontiguousarray(inv[sub]), np.ascontiguousarray(params[:, sub]), 2)
    np.testing.assert_array_equal(H.to_np(view), back)
    np.testing.assert_array_equal(back, x)  # to-Montgomery then standard inverse is the identity


def test_autotuned_schedule_is_persisted_and_exact(tmp_path):
    """tune_ntt times the candidate schedules of a shape, stores the fastest
    in the JSON cache, and the backend then uses it; results are the same
    as one launch."""
    from fhelium_amd import autotune
    from fhelium_amd.runtime import RnsRuntime
    cfg = H.config(14)
    rt = RnsRuntime(cfg, "cuda", "fast")
    be = rt.ntt
    be.tuning = autotune.TuningCache(tmp_path / "tuning.json")
    rec = autotune.tune_ntt(be, 8, 35, 0, forward=True, reps=3)
    assert rec["choice"] in rec["candidates_ms"] and len(rec["candidates_ms"]) >= 3
    again = autotune.TuningCache(tmp_path / "tuning.json")
    key = autotune.ntt_key(torch.device("cuda", 0), 14, 35, 8, True, be._mode)
    assert again.get(key)["schedule"] == rec["schedule"]
    assert be.schedule_for(8, 35, True, torch.device("cuda", 0)) == autotune.schedule_from_record(rec)
    q = torch.tensor(cfg.moduli[:35], dtype=torch.int64, device="cuda").view(1, -1, 1)
    x = torch.randint(0, 1 << 62, (8, 35, cfg.N), dtype=torch.int64, device="cuda") % q
    tuned = be.forward_to_montgomery(x, 0)
    be.tuning = None
    assert torch.equal(tuned, be.forward_to_montgomery(x, 0))


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_round_trip_full_size(mode):
    """[4, 2, 35, 2^16] slice of the ring-kernel workload: inverse(forward(x)) == x."""
    cfg, params, _, _, be = _backend(16, mode)
    g = torch.Generator(device="cuda").manual_seed(11)
    q = torch.tensor(cfg.moduli[:35], dtype=torch.int64, device="cuda").view(1, 1, 35, 1)
    x = (torch.randint(0, 2**62, (4, 2, 35, cfg.N), generator=g, device="cuda") % q)
    y = x.clone()
    be.forward_to_montgomery_(y, 0)
    assert bool((y < (2 * q if mode == "exact" else q)).all())
    be.inverse_to_standard_(y, 0)
    assert torch.equal(x, y)


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_ring_kernels_full_workload(mode):
    """The standalone ring-kernel workload at full size: [64, 2, 35, 2^16]
    uniform residues, forward NTT (to Montgomery), pointwise Montgomery
    multiply by a second such tensor, inverse NTT to standard.  Two sampled
    polynomials (all 35 rows) are checked against the oracle bit for bit;
    every polynomial is checked for canonical range, and the round trip
    inverse(forward(x)) == x is checked on the whole first operand."""
    from fhelium_amd import kernels as K
    cfg, params, fwd, inv, be = _backend(16, mode)
    g = torch.Generator(device="cuda").manual_seed(20260807)
    q = torch.tensor(cfg.moduli[:35], dtype=torch.int64, device="cuda").view(1, 1, 35, 1)
    a = torch.randint(0, 2**62, (64, 2, 35, cfg.N), generator=g, device="cuda") % q
    b = torch.randint(0, 2**62, (64, 2, 35, cfg.N), generator=g, device="cuda") % q
    sample = [(0, 0), (63, 1)]
    a_s = [H.to_np(a[i, c]) for i, c in sample]
    b_s = [H.to_np(b[i, c]) for i, c in sample]
    fa = a.clone()
    be.forward_to_montgomery_(fa, 0)
    be.forward_to_montgomery_(b, 0)
    prod = K.montgomery_mul(fa, b, torch.from_numpy(params[:, :35]).cuda())
    del b
    be.inverse_to_standard_(prod, 0)
    be.inverse_to_standard_(fa, 0)
    assert torch.equal(fa, a)
    assert bool((prod >= 0).all()) and bool((prod < q).all())
    p35, f35, i35 = params[:, :35].copy(), fwd[:35], inv[:35]
    for (i, c), xa, xb in zip(sample, a_s, b_s):
        na = O.ntt_forward(xa[None], f35, p35, to_montgomery=True)
        nb = O.ntt_forward(xb[None], f35, p35, to_montgomery=True)
        want = O.ntt_inverse(O.binary("mul", na, nb, p35), i35, p35, 2)[0]
        np.testing.assert_array_equal(H.to_np(prod[i, c]), want)


@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("log_n", [9, 13, 16])
def test_inverse_out_of_place(log_n, mode):
    """inverse_to_standard(x) reads a strided level-3 view and leaves it
    untouched; the result equals the in-place inverse bit for bit."""
    cfg, params, _, inv, be = _backend(log_n, mode)
    rng = np.random.default_rng(3000 + log_n)
    q = H.moduli_array(cfg, 3, 32)
    x = H.random_residues(rng, 3, q, cfg.N, lazy=True)
    big = torch.zeros((3, 40, cfg.N), dtype=torch.int64, device="cuda")
    big[:, 5:37] = H.to_dev(x)
    view = big[:, 5:37]
    before = big.clone()
    out = be.inverse_to_standard(view, 3)
    assert torch.equal(big, before)
    want = O.ntt_inverse(x, inv[3:35], np.ascontiguousarray(params[:, 3:35]), 2)
    np.testing.assert_array_equal(H.to_np(out), want)


@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("log_n", [9, 13, 16])
def test_forward_out_of_place(log_n, mode):
    """forward_to_montgomery(x) reads a strided level-3 view and leaves it
    untouched; the result equals the oracle (exact) or its canonical form."""
    cfg, params, fwd, _, be = _backend(log_n, mode)
    rng = np.random.default_rng(4000 + log_n)
    q = H.moduli_array(cfg, 3, 32)
    x = H.random_residues(rng, 3, q, cfg.N)
    big = torch.zeros((3, 40, cfg.N), dtype=torch.int64, device="cuda")
    big[:, 5:37] = H.to_dev(x)
    before = big.clone()
    out = be.forward_to_montgomery(big[:, 5:37], 3)
    assert torch.equal(big, before)
    want = O.ntt_forward(x, fwd[3:35], np.ascontiguousarray(params[:, 3:35]), to_montgomery=True)
    if mode == "fast":
        want = want % q[None, :, None]
    np.testing.assert_array_equal(H.to_np(out), want)
