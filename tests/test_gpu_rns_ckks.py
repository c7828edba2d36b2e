"""gfx950 RNS / CKKS kernels vs the CPU oracle, bit for bit.

Every element-wise and key-switch kernel behind the reference's
fhelium_rns_ops / fhelium_ckks_ops schemas is run through the C ABI on seeded
inputs at logN=12 over the L34 chain (39 QP rows: 35 40/41-bit primes, one
60-bit base, four 60-bit P primes) and compared with the oracle's restatement
of the reference CPU loops.  Integer work: equality is exact.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd import kernels as K
from fhelium_amd import rns
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

LOG_N = 12


@pytest.fixture(scope="module")
def setup():
    cfg = H.config(LOG_N)
    t = H.host_tables(LOG_N)
    return cfg, t, t.rns_params.numpy()


def _rng(seed):
    return np.random.default_rng(seed)


@pytest.mark.parametrize("op", ["mul", "add_lazy", "sub_lazy", "add_canonical", "sub_canonical"])
@pytest.mark.parametrize("rhs_batch", [3, 1])
def test_binary(setup, op, rhs_batch):
    cfg, t, params = setup
    rng = _rng(hash(op) % 1000 + rhs_batch)
    m = H.moduli_array(cfg)
    a = H.random_residues(rng, 3, m, cfg.N, lazy=True)
    b = H.random_residues(rng, rhs_batch, m, cfg.N, lazy=True)
    if op.endswith("canonical"):
        a, b = a % m[None, :, None], b % m[None, :, None]
    fn = {"mul": K.montgomery_mul, "add_lazy": K.add_lazy, "sub_lazy": K.sub_lazy,
          "add_canonical": K.add_canonical, "sub_canonical": K.sub_canonical}[op]
    got = fn(H.to_dev(a), H.to_dev(b), t.rns_params.cuda())
    np.testing.assert_array_equal(H.to_np(got), O.binary(op, a, b, params))


def test_canonical_inplace(setup):
    cfg, t, params = setup
    rng = _rng(5)
    m = H.moduli_array(cfg)
    a = H.random_residues(rng, 2, m, cfg.N)
    b = H.random_residues(rng, 2, m, cfg.N)
    da = H.to_dev(a)
    K.add_canonical_(da, H.to_dev(b), t.rns_params.cuda())
    np.testing.assert_array_equal(H.to_np(da), O.binary("add_canonical", a, b, params))
    da = H.to_dev(a)
    K.sub_canonical_(da, H.to_dev(b), t.rns_params.cuda())
    np.testing.assert_array_equal(H.to_np(da), O.binary("sub_canonical", a, b, params))


@pytest.mark.parametrize("layout", ["mul_cyclic", "mul_contiguous"])
def test_compressed_mul(setup, layout):
    cfg, t, params = setup
    rng = _rng(17)
    m = H.moduli_array(cfg)
    a = H.random_residues(rng, 2, m, cfg.N, lazy=True)
    b = H.random_residues(rng, 1, m, 64, lazy=True)
    fn = K.montgomery_mul_cyclic_compressed if layout == "mul_cyclic" else K.montgomery_mul_contiguous_compressed
    got = fn(H.to_dev(a), H.to_dev(b), t.rns_params.cuda())
    np.testing.assert_array_equal(H.to_np(got), O.binary(layout, a, b, params))


@pytest.mark.parametrize("op", ["to_montgomery", "from_montgomery", "canonicalize", "center", "shift_positive"])
def test_unary(setup, op):
    cfg, t, params = setup
    rng = _rng(hash(op) % 997)
    m = H.moduli_array(cfg)
    x = H.random_residues(rng, 2, m, cfg.N, lazy=op in ("from_montgomery", "canonicalize"))
    if op == "shift_positive":
        x = x - (m[None, :, None] // 2)
    dev = H.to_dev(x)
    getattr(K, {"to_montgomery": "to_montgomery_", "from_montgomery": "from_montgomery_",
                "canonicalize": "canonicalize_residues_", "center": "center_residues_",
                "shift_positive": "shift_residues_positive_"}[op])(dev, t.rns_params.cuda())
    np.testing.assert_array_equal(H.to_np(dev), O.unary(op, x, params))


def test_row_scalars(setup):
    cfg, t, params = setup
    rng = _rng(23)
    m = H.moduli_array(cfg)
    x = H.random_residues(rng, 2, m, cfg.N, lazy=True)
    s = np.array([rng.integers(0, int(q)) for q in m], dtype=np.int64)
    dev = H.to_dev(x)
    K.montgomery_mul_row_scalars_(dev, H.to_dev(s), t.rns_params.cuda())
    np.testing.assert_array_equal(H.to_np(dev), O.row_scalars(x, s, params, False))
    got = K.montgomery_mul_row_scalars_canonical(H.to_dev(x), H.to_dev(s), t.rns_params.cuda())
    np.testing.assert_array_equal(H.to_np(got), O.row_scalars(x, s, params, True))


def test_add_lazy_twice_and_lift(setup):
    cfg, t, params = setup
    rng = _rng(29)
    m = H.moduli_array(cfg)
    twice = 2 * m
    a = H.random_residues(rng, 2, m, cfg.N, lazy=True)
    b = H.random_residues(rng, 2, m, cfg.N, lazy=True)
    got = K.add_lazy_with_twice_modulus(H.to_dev(a), H.to_dev(b), H.to_dev(twice))
    np.testing.assert_array_equal(H.to_np(got), O.add_lazy_twice(a, b, twice))
    c = rng.integers(-(1 << 30), 1 << 30, size=(2, cfg.N), dtype=np.int64)
    got = K.lift_centered_coefficients(H.to_dev(c), H.to_dev(twice))
    assert tuple(got.shape) == (2, len(m), cfg.N)
    np.testing.assert_array_equal(H.to_np(got), O.lift(c, twice))


@pytest.mark.parametrize("level", [0, 1, 3, 17, 33])
def test_mixed_radix_decompose_and_extend(setup, level):
    cfg, t, params = setup
    rng = _rng(100 + level)
    for spec in t.layout.digit_specs(level):
        d = t.digit_tables[spec.prime_ids]
        m = np.array([cfg.moduli[i] for i in spec.prime_ids], dtype=np.int64)
        x = H.random_residues(rng, 2, m, cfg.N, lazy=True)
        got = K.mixed_radix_decompose(H.to_dev(x), d.normalizers.cuda(), d.propagation.cuda(),
                                      d.modulus_lo.cuda(), d.modulus_hi.cuda(), d.neg_inv_lo.cuda(),
                                      d.neg_inv_hi.cuda())
        want = O.mixed_radix_decompose(x, d.normalizers.numpy(), d.propagation.numpy(),
                                       d.modulus_lo.numpy(), d.modulus_hi.numpy(),
                                       d.neg_inv_lo.numpy(), d.neg_inv_hi.numpy())
        np.testing.assert_array_equal(H.to_np(got), want)
        # ModUp onto the active QP rows of this level
        rows = t.layout.prime_ids(level, include_p=True)
        coef = np.ascontiguousarray(d.basis_extension.numpy()[:, level:])
        sub_params = np.ascontiguousarray(params[:, level:])
        ext = K.mixed_radix_basis_extend_to_montgomery(
            H.to_dev(want), H.to_dev(coef) if coef.size else torch.empty(0, len(rows), dtype=torch.int64, device="cuda"),
            H.to_dev(sub_params), len(rows))
        np.testing.assert_array_equal(H.to_np(ext), O.basis_extend(want, coef, sub_params, len(rows)))


@pytest.mark.parametrize("nearest", [True, False])
@pytest.mark.parametrize("level", [0, 20, 32])
def test_rescale(setup, nearest, level):
    cfg, t, params = setup
    rng = _rng(300 + level + nearest)
    ids = t.layout.prime_ids(level)
    m = np.array([cfg.moduli[i] for i in ids], dtype=np.int64)
    x = H.random_residues(rng, 2, m, cfg.N)
    inv = t.rescale_inverses[level][: len(ids) - 1].numpy()
    sub = np.ascontiguousarray(params[:, ids[1]: ids[-1] + 1])
    half = int(m[0]) // 2
    fn = K.rescale_drop_leading_prime_nearest if nearest else K.rescale_drop_leading_prime_truncate
    args = (H.to_dev(np.ascontiguousarray(x[:, 1:])), H.to_dev(inv),
            H.to_dev(np.ascontiguousarray(x[:, 0])), H.to_dev(sub))
    got = fn(*args, half) if nearest else fn(*args)
    want = O.rescale(x[:, 1:], x[:, 0], inv, sub, half, nearest)
    np.testing.assert_array_equal(H.to_np(got), want)
    # in-place form
    dev = H.to_dev(np.ascontiguousarray(x[:, 1:]))
    if nearest:
        K.rescale_drop_leading_prime_nearest_(dev, args[1], args[2], args[3], half)
    else:
        K.rescale_drop_leading_prime_truncate_(dev, args[1], args[2], args[3])
    np.testing.assert_array_equal(H.to_np(dev), want)


@pytest.mark.parametrize("step", [1, 5, -3, 1000])
def test_galois(setup, step):
    cfg, t, params = setup
    rng = _rng(400 + abs(step))
    m = H.moduli_array(cfg)
    g = rns.rotation_galois_element(cfg.N, step)
    idx, sgn = rns.coefficient_galois_tables(cfg.N, g)
    x = H.random_residues(rng, 2, m, cfg.N)
    got = K.apply_coefficient_galois_automorphism(H.to_dev(x), idx.cuda(), sgn.cuda(), H.to_dev(2 * m))
    np.testing.assert_array_equal(H.to_np(got), O.galois_coefficient(x, idx.numpy(), sgn.numpy(), 2 * m))
    src = rns.ntt_galois_indices(cfg.logN, g)
    got = K.apply_ntt_galois_automorphism(H.to_dev(x), src.cuda())
    np.testing.assert_array_equal(H.to_np(got), O.galois_ntt(x, src.numpy()))


@pytest.mark.parametrize("level", [0, 9, 33])
def test_keyswitch_mac_and_moddown(setup, level):
    cfg, t, params = setup
    rng = _rng(500 + level)
    rows = t.layout.prime_ids(level, include_p=True)
    m_all = H.moduli_array(cfg)
    m = m_all[list(rows)]
    sub = np.ascontiguousarray(params[:, level:])
    acc0 = H.random_residues(rng, 2, m, cfg.N, lazy=True)
    acc1 = H.random_residues(rng, 2, m, cfg.N, lazy=True)
    digit = H.random_residues(rng, 2, m, cfg.N, lazy=True)
    key = np.stack([H.random_residues(rng, 1, m_all, cfg.N, lazy=True)[0] for _ in range(2)])
    d0, d1 = H.to_dev(acc0), H.to_dev(acc1)
    K.keyswitch_accumulate_digit_products_(d0, d1, H.to_dev(digit), H.to_dev(key), H.to_dev(sub), level)
    w0, w1 = O.keyswitch_mac(acc0, acc1, digit, key, sub, level)
    np.testing.assert_array_equal(H.to_np(d0), w0)
    np.testing.assert_array_equal(H.to_np(d1), w1)
    # ModDown of canonical QP rows
    p = cfg.num_p_primes
    qp = H.random_residues(rng, 2, m, cfg.N)
    inv = t.moddown_inverses[level].numpy()
    got = K.keyswitch_moddown_qp_to_q(H.to_dev(np.ascontiguousarray(qp[:, :-p])),
                                      H.to_dev(np.ascontiguousarray(qp[:, -p:])), H.to_dev(inv),
                                      H.to_dev(sub))
    want = O.moddown(qp[:, :-p], qp[:, -p:], inv, sub)
    np.testing.assert_array_equal(H.to_np(got), want)


@pytest.mark.parametrize("layout", ["dense", "cyclic", "contiguous", "strided"])
def test_plaintext_add(setup, layout):
    cfg, t, params = setup
    rng = _rng(600 + len(layout))
    m = H.moduli_array(cfg)
    ct = H.random_residues(rng, 2, m, cfg.N)
    support = cfg.N if layout == "dense" else 256
    pt = H.random_residues(rng, 1, m, support)
    imp = np.array([[rng.integers(0, int(q)) for q in m]], dtype=np.int64) if layout == "strided" else None
    dev_ct = H.to_dev(ct)
    if layout == "dense":
        got = K.add_prepared_plaintext_component(dev_ct, H.to_dev(pt), t.rns_params.cuda())
    elif layout == "cyclic":
        got = K.add_cyclic_compressed_plaintext_component(dev_ct, H.to_dev(pt), t.rns_params.cuda())
    elif layout == "contiguous":
        got = K.add_contiguous_compressed_plaintext_component(dev_ct, H.to_dev(pt), t.rns_params.cuda())
    else:
        got = K.add_strided_plaintext_component(dev_ct, H.to_dev(pt), H.to_dev(imp), t.rns_params.cuda())
    np.testing.assert_array_equal(H.to_np(got), O.add_plaintext(layout, ct, pt, params, imp))


@pytest.mark.parametrize("terms", [1, 3, 8])
def test_scalar_combination(setup, terms):
    """fh_ckks_scalar_combination against the oracle's Montgomery row-scalar
    products (montgomery_mul_row_scalars_canonical) summed with
    add_canonical; terms read through row-offset and strided batch views."""
    cfg, t, params = setup
    rng = _rng(700 + terms)
    m = H.moduli_array(cfg)
    rows, start = 30, 9  # output rows 9..38
    sub = np.ascontiguousarray(params[:, start: start + rows])
    xs, views = [], []
    for k in range(terms):
        extra = k % 3  # some terms sit at a lower level: read from a row offset
        full = H.random_residues(rng, 2 * 3, m[start - extra: start + rows], cfg.N).reshape(2, 3, rows + extra, cfg.N)
        dev = H.to_dev(full)
        view = dev[:, ::2][..., extra:, :] if k % 2 else dev[:, 1:3][..., extra:, :]
        views.append(view)
        xs.append(H.to_np(view))
    s = np.array([[rng.integers(0, int(q)) for q in m[start: start + rows]] for _ in range(terms)], dtype=np.int64)
    got = K.scalar_combination(views, H.to_dev(s), H.to_dev(sub))
    assert tuple(got.shape) == (2, 2, rows, cfg.N)
    want = None
    for x, sc in zip(xs, s):
        term = O.row_scalars(x.reshape(4, rows, cfg.N), sc, sub, True)
        want = term if want is None else O.binary("add_canonical", want, term, sub)
    np.testing.assert_array_equal(H.to_np(got).reshape(4, rows, cfg.N), want)


def test_scalar_combination_rejects_bad_operands(setup):
    cfg, t, params = setup
    x = torch.zeros((2, 1, 3, cfg.N), dtype=torch.int64, device="cuda")
    s = torch.zeros((1, 3), dtype=torch.int64, device="cuda")
    p = H.to_dev(np.ascontiguousarray(params[:, :3]))
    with pytest.raises(RuntimeError):
        K.scalar_combination([x] * 9, torch.zeros((9, 3), dtype=torch.int64, device="cuda"), p)
    with pytest.raises(RuntimeError):
        K.scalar_combination([x, x[:, :, :2]], s, p)
    with pytest.raises(RuntimeError):
        K.scalar_combination([x], torch.zeros((2, 3), dtype=torch.int64, device="cuda"), p)
