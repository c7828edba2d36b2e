"""The reference operator schemas (torch.ops.fhelium_*_ops) on gfx950 vs the
oracle: the drop-in boundary the reference's Python layers call.

Compact ops get reference-format tables (compact Montgomery twiddles and
rns_params sliced to the operand's rows, as
fhelium/engine/ntt/interface.py:118-176 slices them); indexed ops get the
reference's radix-2 DIF/DIT schedules (fhelium/engine/ntt/plans/
indexed_radix2.py:58-126) with stage-expanded twiddles (twiddles.py:87-106).
Forward outputs are compared bit for bit in exact mode (the default of the
torch ops) and modulo q in fast mode; inverse outputs bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd import rns
from fhelium_amd import torch_ops
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

NTT = torch.ops.fhelium_ntt_ops
RNS = torch.ops.fhelium_rns_ops
CKKS = torch.ops.fhelium_ckks_ops


def _tables(log_n: int, row_start: int, rows: int):
    cfg = H.config(log_n)
    t = H.host_tables(log_n)
    fwd, inv = rns.compact_twiddles_montgomery(cfg.moduli, log_n)
    sl = slice(row_start, row_start + rows)
    params = t.rns_params[:, sl].contiguous()
    return cfg, params, fwd[sl].contiguous(), inv[sl].contiguous()


def _schedules(log_n: int):
    """Forward DIF and inverse DIT (even, odd, twiddle index) schedules."""
    n = 1 << log_n
    f_even, f_odd, f_tw, i_even, i_odd, i_tw = ([] for _ in range(6))
    for s in range(log_n):
        half = n >> (s + 1)
        e = [2 * g * half + k for g in range(1 << s) for k in range(half)]
        f_even.append(e)
        f_odd.append([x + half for x in e])
        f_tw.append([(1 << s) + g for g in range(1 << s) for _ in range(half)])
        half = 1 << s
        groups = n // (2 * half)
        e = [2 * g * half + k for g in range(groups) for k in range(half)]
        i_even.append(e)
        i_odd.append([x + half for x in e])
        i_tw.append([groups + g for g in range(groups) for _ in range(half)])
    as32 = lambda v: torch.tensor(v, dtype=torch.int32)  # noqa: E731
    return as32(f_even), as32(f_odd), as32(f_tw), as32(i_even), as32(i_odd), as32(i_tw)


def _expand(compact: torch.Tensor, idx: torch.Tensor) -> torch.Tensor:
    stages, bf = idx.shape
    return compact.index_select(1, idx.reshape(-1).long()).reshape(compact.size(0), stages, bf)


@pytest.mark.parametrize("log_n,row_start", [(12, 0), (16, 5)])
@pytest.mark.parametrize("to_mont", [False, True])
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_compact_forward(log_n, row_start, to_mont, mode):
    rows = 39 - row_start
    cfg, params, fwd, _ = _tables(log_n, row_start, rows)
    q = H.moduli_array(cfg, row_start, rows)
    x = H.random_residues(np.random.default_rng(log_n + row_start), 2, q, cfg.N, lazy=not to_mont)
    dev = H.to_dev(x)
    torch_ops.set_ntt_mode(mode)
    try:
        op = NTT.forward_ntt_to_montgomery_compact_grouped_smem_ if to_mont \
            else NTT.forward_ntt_montgomery_compact_grouped_smem_
        op(dev, fwd.cuda(), params.cuda(), 4)
    finally:
        torch_ops.set_ntt_mode("exact")
    want = O.ntt_forward(x, fwd.numpy(), params.numpy(), to_montgomery=to_mont)
    got = H.to_np(dev)
    if mode == "fast":
        np.testing.assert_array_equal(got, want % q[None, :, None])
    else:
        np.testing.assert_array_equal(got, want)


def test_compact_forward_functional_does_not_alias():
    cfg, params, fwd, _ = _tables(12, 0, 39)
    x = H.to_dev(H.random_residues(np.random.default_rng(3), 1, H.moduli_array(cfg), cfg.N))
    before = x.clone()
    out = NTT.forward_ntt_to_montgomery_compact_grouped_smem(x, fwd.cuda(), params.cuda(), 3)
    assert out.data_ptr() != x.data_ptr()
    assert torch.equal(x, before)
    want = O.ntt_forward(H.to_np(before), fwd.numpy(), params.numpy(), to_montgomery=True)
    np.testing.assert_array_equal(H.to_np(out), want)


@pytest.mark.parametrize("form", [0, 1, 2, 3])
@pytest.mark.parametrize("log_n,row_start", [(12, 3), (16, 0)])
def test_compact_inverse(form, log_n, row_start):
    rows = 39 - row_start
    cfg, params, _, inv = _tables(log_n, row_start, rows)
    q = H.moduli_array(cfg, row_start, rows)
    x = H.random_residues(np.random.default_rng(7 * form + log_n), 2, q, cfg.N, lazy=True)
    dev = H.to_dev(x)
    name = ["inverse_ntt_montgomery_compact_grouped_smem_", "inverse_ntt_to_standard_lazy_compact_grouped_smem_",
            "inverse_ntt_to_standard_compact_grouped_smem_", "inverse_ntt_to_centered_compact_grouped_smem_"][form]
    getattr(NTT, name)(dev, inv.cuda(), params.cuda(), 3)
    np.testing.assert_array_equal(H.to_np(dev), O.ntt_inverse(x, inv.numpy(), params.numpy(), form))


@pytest.mark.parametrize("log_n", [5, 10])
def test_indexed_forward_and_inverse(log_n):
    cfg, params, fwd, inv = _tables(log_n, 0, 39)
    q = H.moduli_array(cfg)
    fe, fo, ft, ie, io, it = _schedules(log_n)
    ftw, itw = _expand(fwd, ft), _expand(inv, it)
    x = H.random_residues(np.random.default_rng(log_n), 2, q, cfg.N)
    dev = H.to_dev(x)
    NTT.forward_ntt_to_montgomery_indexed_(dev, fe.cuda(), fo.cuda(), ftw.cuda(), params.cuda())
    want = O.ntt_forward(x, fwd.numpy(), params.numpy(), to_montgomery=True)
    np.testing.assert_array_equal(H.to_np(dev), want)
    for form, name in enumerate(["inverse_ntt_montgomery_indexed_", "inverse_ntt_to_standard_lazy_indexed_",
                                 "inverse_ntt_to_standard_indexed_", "inverse_ntt_to_centered_indexed_"]):
        d = H.to_dev(want)
        getattr(NTT, name)(d, ie.cuda(), io.cuda(), itw.cuda(), params.cuda())
        np.testing.assert_array_equal(H.to_np(d), O.ntt_inverse(want, inv.numpy(), params.numpy(), form))


def test_compact_keyswitch_accumulate_fused():
    log_n, row_start = 12, 4
    rows = 39 - row_start
    cfg, params, fwd, _ = _tables(log_n, row_start, rows)
    q = H.moduli_array(cfg, row_start, rows)
    rng = np.random.default_rng(11)
    digit = H.random_residues(rng, 3, q, cfg.N, lazy=True)
    acc0 = H.random_residues(rng, 3, q, cfg.N, lazy=True)
    acc1 = H.random_residues(rng, 3, q, cfg.N, lazy=True)
    key = H.random_residues(rng, 2, H.moduli_array(cfg), cfg.N, lazy=True)  # [2, 39, N]
    d, a0, a1 = H.to_dev(digit), H.to_dev(acc0), H.to_dev(acc1)
    NTT.forward_ntt_montgomery_compact_keyswitch_accumulate_(d, fwd.cuda(), params.cuda(), H.to_dev(key),
                                                             a0, a1, row_start)
    nd = O.ntt_forward(digit, fwd.numpy(), params.numpy(), to_montgomery=False)
    w0, w1 = O.keyswitch_mac(acc0, acc1, nd, key, params.numpy(), row_start)
    np.testing.assert_array_equal(H.to_np(d), nd)
    np.testing.assert_array_equal(H.to_np(a0) % q[None, :, None], w0 % q[None, :, None])
    np.testing.assert_array_equal(H.to_np(a1) % q[None, :, None], w1 % q[None, :, None])


def test_rns_and_ckks_ops_through_dispatcher():
    cfg = H.config(12)
    t = H.host_tables(12)
    q = H.moduli_array(cfg, 0, 35)
    params = t.rns_params[:, :35].contiguous()
    rng = np.random.default_rng(5)
    a = H.random_residues(rng, 2, q, cfg.N, lazy=True)
    b = H.random_residues(rng, 2, q, cfg.N, lazy=True)
    pd = params.cuda()
    for name, op in (("montgomery_mul", "mul"), ("add_lazy", "add_lazy"), ("sub_lazy", "sub_lazy"),
                     ("add_canonical", "add_canonical"), ("sub_canonical", "sub_canonical")):
        got = getattr(RNS, name)(H.to_dev(a), H.to_dev(b), pd)
        np.testing.assert_array_equal(H.to_np(got), O.binary(op, a, b, params.numpy()), err_msg=name)
    x = H.to_dev(a)
    RNS.canonicalize_residues_(x, pd)
    np.testing.assert_array_equal(H.to_np(x), O.unary("canonicalize", a, params.numpy()))
    idx, sgn = rns.coefficient_galois_tables(cfg.N, 5)
    c = H.random_residues(rng, 1, q, cfg.N)
    got = CKKS.apply_coefficient_galois_automorphism(H.to_dev(c), idx.cuda(), sgn.cuda(), pd[0].contiguous())
    want = O.galois_coefficient(c, idx.numpy(), sgn.numpy(), params.numpy()[0])
    np.testing.assert_array_equal(H.to_np(got), want)


@pytest.mark.parametrize("log_n,radix,row_start", [(12, 4, 0), (12, 8, 3), (12, 16, 0), (16, 16, 0), (16, 4, 30)])
def test_power_of_two_radix_family(log_n, radix, row_start):
    """The reference's radix-4/8/16 NTT schemas (forward_ntt.cpp:56-66,
    inverse_ntt.cpp:58-72) with its outer-twist and radix-root tables:
    forward outputs equal the oracle's NTT modulo q (as the reference's own
    cross-family tests compare them, test_ntt_backend.py:350-455), standard
    and centered inverse outputs bit for bit, lazy inverse outputs modulo q,
    and the functional form leaves its input alone."""
    rows = 39 - row_start
    cfg, params, fwd, inv = _tables(log_n, row_start, rows)
    fo, io, fr, ir = rns.power_of_two_radix_twiddles(cfg.moduli[row_start: row_start + rows], log_n, radix)
    q = H.moduli_array(cfg, row_start, rows)[None, :, None]
    rng = np.random.default_rng(log_n * 100 + radix)
    x = H.random_residues(rng, 2, q.ravel(), cfg.N, lazy=True)
    p, fo_d, fr_d, io_d, ir_d = params.cuda(), fo.cuda(), fr.cuda(), io.cuda(), ir.cuda()
    print(f"radix family logN={log_n} radix={radix} rows={rows}", flush=True)
    dev = H.to_dev(x)
    NTT.forward_ntt_montgomery_power_of_two_radix_compact_(dev, fo_d, fr_d, p)
    want = O.ntt_forward(x, fwd.numpy(), params.numpy(), to_montgomery=False)
    np.testing.assert_array_equal(H.to_np(dev) % q, want % q)
    xs = x % q
    dev = H.to_dev(xs)
    NTT.forward_ntt_to_montgomery_power_of_two_radix_compact_(dev, fo_d, fr_d, p)
    want = O.ntt_forward(xs, fwd.numpy(), params.numpy(), to_montgomery=True)
    np.testing.assert_array_equal(H.to_np(dev) % q, want % q)
    src = H.to_dev(xs)
    out = NTT.forward_ntt_to_montgomery_power_of_two_radix_compact(src, fo_d, fr_d, p)
    assert torch.equal(src, H.to_dev(xs)) and out.data_ptr() != src.data_ptr()
    np.testing.assert_array_equal(H.to_np(out) % q, want % q)
    for form, suffix in enumerate(["montgomery", "to_standard_lazy", "to_standard", "to_centered"]):
        dev = H.to_dev(x)
        getattr(NTT, f"inverse_ntt_{suffix}_power_of_two_radix_compact_")(dev, io_d, ir_d, p)
        want = O.ntt_inverse(x, inv.numpy(), params.numpy(), form)
        if form >= 2:
            np.testing.assert_array_equal(H.to_np(dev), want, err_msg=suffix)
        else:
            np.testing.assert_array_equal(H.to_np(dev) % q, want % q, err_msg=suffix)
    # round trip: standard -> NTT -> standard is the identity on canonical residues
    dev = H.to_dev(xs)
    NTT.forward_ntt_to_montgomery_power_of_two_radix_compact_(dev, fo_d, fr_d, p)
    NTT.inverse_ntt_to_standard_power_of_two_radix_compact_(dev, io_d, ir_d, p)
    np.testing.assert_array_equal(H.to_np(dev), xs)


def test_power_of_two_radix_validation():
    cfg, params, _, _ = _tables(16, 0, 39)
    fo, io, fr, ir = rns.power_of_two_radix_twiddles(cfg.moduli, 16, 16)
    x = torch.zeros((1, 39, cfg.N), dtype=torch.int64, device="cuda")
    p = params.cuda()
    with pytest.raises(RuntimeError, match="radix 4, 8, or 16"):
        NTT.forward_ntt_montgomery_power_of_two_radix_compact_(x, fo.cuda(), fr[:, :5].contiguous().cuda(), p)
    with pytest.raises(RuntimeError, match="coefficient - 1"):
        NTT.forward_ntt_montgomery_power_of_two_radix_compact_(x, fo[:, :-1].contiguous().cuda(), fr.cuda(), p)
    with pytest.raises(RuntimeError, match="different roots"):
        NTT.forward_ntt_montgomery_power_of_two_radix_compact_(x, fo.cuda(), ir.cuda(), p)
    _, _, fr8, _ = rns.power_of_two_radix_twiddles(cfg.moduli[:2], 12, 8)
    with pytest.raises(RuntimeError, match="divisible by 3"):
        NTT.forward_ntt_montgomery_power_of_two_radix_compact_(
            x, fo.cuda(), torch.zeros((39, 8), dtype=torch.int64, device="cuda"), p)
    with pytest.raises(RuntimeError, match="grouped_stage_count"):
        NTT.forward_ntt_montgomery_compact_grouped_smem_(x, fo.cuda(), p, 5)
