"""Pin the CPU oracle (CPU only).

Every oracle function is checked against tests/golden/ckks_golden_v1.json,
whose expected values come from Python-integer closed forms
(tests/golden/make_golden.py), plus closed-form properties on fresh seeded
inputs (round trips, ranges).  A plausible mistake -- a dropped term, a wrong
sign or index, a transposed operand, the wrong rounding -- changes the
expected values and fails one of these.
"""

from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "ckks_golden_v1.json").read_text())
A = lambda v: np.array(v, dtype=np.int64)  # noqa: E731


def test_mont_mul_known_answers():
    for case in GOLDEN["mont_mul"]:
        got = [O.mont_mul_scalar(a, b, case["q"], case["k"]) for a, b in zip(case["a"], case["b"])]
        assert got == case["expected"]  # the reference REDC returns canonical values


@pytest.mark.parametrize("case", GOLDEN["ntt"], ids=lambda c: f"logN{c['log_n']}")
def test_ntt_closed_form(case):
    q = A(case["moduli"]).reshape(1, -1, 1)
    params = A(case["params"])
    x = A(case["x"])[None]
    fwd = O.ntt_forward(x, A(case["forward_twiddles_montgomery"]), params, to_montgomery=False)
    assert (fwd >= 0).all() and (fwd < 2 * q).all()          # lazy range
    np.testing.assert_array_equal(fwd % q, A(case["forward_canonical"])[None])
    fwd_std = O.ntt_forward(x, A(case["forward_twiddles_montgomery"]), params, to_montgomery=True)
    np.testing.assert_array_equal(fwd_std % q, A(case["forward_to_montgomery_canonical"])[None])
    # inverse: canonical outputs; Montgomery form round-trips exactly
    inv = O.ntt_inverse(fwd, A(case["inverse_twiddles_montgomery"]), params, 0)
    np.testing.assert_array_equal(inv, x)
    back = O.ntt_inverse(fwd_std, A(case["inverse_twiddles_montgomery"]), params, 2)
    np.testing.assert_array_equal(back, x)
    centered = O.ntt_inverse(fwd_std, A(case["inverse_twiddles_montgomery"]), params, 3)
    want_c = np.where(x <= q // 2, x, x - q)
    np.testing.assert_array_equal(centered, want_c)


@pytest.mark.parametrize("case", GOLDEN["mixed_radix"], ids=lambda c: f"D{len(c['moduli'])}")
def test_mixed_radix_and_basis_extension(case):
    src = A(case["source"])[None]
    digits = O.mixed_radix_decompose(src, A(case["normalizers"]), A(case["propagation"]), A(case["lo"]),
                                     A(case["hi"]), A(case["klo"]), A(case["khi"]))
    np.testing.assert_array_equal(digits[0], A(case["digits"]))
    dest = A(case["destination_moduli"]).reshape(1, -1, 1)
    ext = O.basis_extend(digits, A(case["basis_extension"]), A(case["destination_params"]), dest.shape[1])
    assert (ext >= 0).all() and (ext < 2 * dest).all()
    np.testing.assert_array_equal(ext % dest, A(case["extended_canonical"])[None])


def test_moddown_floor():
    case = GOLDEN["moddown"]
    got = O.moddown(A(case["q"])[None], A(case["p"])[None], A(case["inverses"]), A(case["params"]))
    np.testing.assert_array_equal(got[0], A(case["expected"]))


@pytest.mark.parametrize("nearest", [True, False])
def test_rescale_rounding(nearest):
    case = GOLDEN["rescale"]
    res = A(case["residues"])[None]
    got = O.rescale(res[:, 1:], res[:, 0], A(case["inverse"]), A(case["params_remaining"]), case["half"],
                    nearest)
    np.testing.assert_array_equal(got[0], A(case["nearest" if nearest else "truncate"]))


@pytest.mark.parametrize("case", GOLDEN["galois"], ids=lambda c: f"g{c['g']}")
def test_galois_brute_force(case):
    n = case["n"]
    moduli = A([1099510054913, 1152921504606584833])
    coeffs = A(case["coefficients"])
    x = (coeffs[None, :] % moduli[:, None])[None]
    got = O.galois_coefficient(x, case["source_indices"], case["source_sign"], 2 * moduli)
    want = (A(case["image"])[None, :] % moduli[:, None])[None]
    np.testing.assert_array_equal(got, want)
    # NTT-domain gather is a permutation: applying it to arange recovers it
    perm = O.galois_ntt(np.arange(n, dtype=np.int64).reshape(1, 1, n), case["source_indices"])
    np.testing.assert_array_equal(perm[0, 0], A(case["source_indices"]))


def test_elementwise_closed_forms():
    rng = np.random.default_rng(42)
    case = GOLDEN["ntt"][0]
    q = A(case["moduli"])
    params = A(case["params"])
    n = 64
    a = np.stack([rng.integers(0, 2 * int(m), n) for m in q])[None]
    b = np.stack([rng.integers(0, 2 * int(m), n) for m in q])[None]
    qq = q.reshape(1, -1, 1)
    R = 1 << 62
    rinv = np.array([pow(R, -1, int(m)) for m in q], dtype=object).reshape(1, -1, 1)
    mul = O.binary("mul", a, b, params)
    np.testing.assert_array_equal(mul, (a.astype(object) * b * rinv % qq).astype(np.int64))
    for op, fn in (("add_lazy", lambda x, y: x + y), ("sub_lazy", lambda x, y: x - y)):
        out = O.binary(op, a, b, params)
        assert (out >= 0).all() and (out < 2 * qq).all()
        np.testing.assert_array_equal(out % qq, fn(a, b) % qq)
    ac, bc = a % qq, b % qq
    np.testing.assert_array_equal(O.binary("add_canonical", ac, bc, params), (ac + bc) % qq)
    np.testing.assert_array_equal(O.binary("sub_canonical", ac, bc, params), (ac - bc) % qq)
    to_m = O.unary("to_montgomery", ac, params)
    np.testing.assert_array_equal(to_m, (ac.astype(object) * R % qq).astype(np.int64))
    np.testing.assert_array_equal(O.unary("from_montgomery", to_m, params), ac)
    np.testing.assert_array_equal(O.unary("canonicalize", a, params), a % qq)
    cen = O.unary("center", ac, params)
    np.testing.assert_array_equal(cen, np.where(ac <= qq // 2, ac, ac - qq))
    np.testing.assert_array_equal(O.unary("shift_positive", cen, params), cen + qq)
    # plaintext add: c + p R^{-1} (p is Montgomery), canonical
    pt = O.add_plaintext("dense", ac, bc, params)
    np.testing.assert_array_equal(pt, ((ac.astype(object) + bc * rinv) % qq).astype(np.int64))
    # key MAC: acc + x k R^{-1}, lazy
    acc0, acc1 = O.keyswitch_mac(np.zeros_like(a), np.zeros_like(a), a, np.stack([b[0], b[0]]), params, 0)
    np.testing.assert_array_equal(acc0 % qq, mul % qq)


def test_host_tables_match_closed_form_fixtures():
    """The product's table builder reproduces the independently generated
    reference-format tables (parameter rows, twiddles, ModUp tables)."""
    import torch  # noqa: F401

    from fhelium_amd import rns

    case = GOLDEN["ntt"][1]
    rows = rns.montgomery_rows(tuple(case["moduli"]), case["log_n"], 40)
    assert rows == case["params"]
    fwd, inv = rns.compact_twiddles_montgomery(tuple(case["moduli"]), case["log_n"])
    assert fwd.tolist() == case["forward_twiddles_montgomery"]
    assert inv.tolist() == case["inverse_twiddles_montgomery"]
    mr = GOLDEN["mixed_radix"][0]
    t = rns._digit_tables(tuple(mr["destination_moduli"]), (0, 1, 2, 3), tuple(range(len(mr["destination_moduli"]))))
    assert t.normalizers.tolist() == mr["normalizers"]
    assert t.propagation.tolist() == mr["propagation"]
    assert t.basis_extension.tolist() == mr["basis_extension"]


def test_remaining_elementwise_closed_forms():
    """Row scalars, add_lazy_with_twice_modulus, centered lift and the three
    compressed plaintext layouts against big-integer closed forms."""
    rng = np.random.default_rng(43)
    case = GOLDEN["ntt"][0]
    q = A(case["moduli"])
    params = A(case["params"])
    n = 64
    qq = q.reshape(1, -1, 1)
    R = 1 << 62
    rinv = np.array([pow(R, -1, int(m)) for m in q], dtype=object).reshape(1, -1, 1)
    a = np.stack([rng.integers(0, 2 * int(m), n) for m in q])[None]
    ac = a % qq
    s = np.array([rng.integers(0, int(m)) for m in q], dtype=np.int64)
    for canonical in (False, True):
        got = O.row_scalars(a, s, params, canonical)
        want = (a.astype(object) * s.reshape(1, -1, 1) * rinv % qq).astype(np.int64)
        np.testing.assert_array_equal(got, want)
    b = np.stack([rng.integers(0, 2 * int(m), n) for m in q])[None]
    got = O.add_lazy_twice(a, b, 2 * q)
    assert (got >= 0).all() and (got < 2 * qq).all()
    np.testing.assert_array_equal(got % qq, (a + b) % qq)
    small = rng.integers(-(1 << 20), 1 << 20, size=(2, n))
    got = O.lift(small, 2 * q)
    assert (got >= 0).all() and (got < 2 * qq).all()
    np.testing.assert_array_equal(got % qq, small[:, None, :] % qq)
    support = 16
    repeat = n // support
    pt = np.stack([rng.integers(0, int(m), support) for m in q])[None]
    implicit = np.array([rng.integers(0, int(m)) for m in q], dtype=np.int64)
    idx = np.arange(n)
    expanded = {
        "cyclic": pt[:, :, idx % support],
        "contiguous": pt[:, :, idx // repeat],
        "strided": np.where(idx % repeat == 0, pt[:, :, idx // repeat], implicit.reshape(1, -1, 1)),
    }
    for layout, p in expanded.items():
        got = O.add_plaintext(layout, ac, pt, params, implicit if layout == "strided" else None)
        want = ((ac.astype(object) + p * rinv) % qq).astype(np.int64)
        np.testing.assert_array_equal(got, want, err_msg=layout)
