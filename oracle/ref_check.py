"""Run the reference's own CPU kernels beside the oracle (test infrastructure).

Loads oracle/_ref/libfhelium_ref_cpu.so (built by oracle/build_ref.py from
the reference's csrc/ops/*/cpu/*.cpp) into this process, which defines the
reference's torch.ops.fhelium_{ntt,rns,ckks}_ops schemas with the
reference's CPU kernels, and compares every CPU-implemented op with the
oracle's C restatement on the same seeded inputs, bit for bit:

  rns   montgomery_mul (+ cyclic/contiguous compressed), add/sub lazy and
        canonical (+ in-place), row scalars (lazy and canonical), to/from
        Montgomery, canonicalize, center, shift positive,
        add_lazy_with_twice_modulus, lift_centered_coefficients,
        mixed_radix_decompose, mixed_radix_basis_extend_to_montgomery
        (rns_arithmetic_cpu.cpp:702-751, rns_canonical_arithmetic_cpu.cpp:253-259)
  ntt   the indexed forward transforms and the four inverse output forms
        with the radix-2 DIF/DIT schedules and stage-expanded twiddles
        (indexed_ntt_cpu.cpp:464-478)
  ckks  the four plaintext-add layouts (+ in-place), rescale nearest and
        truncate (+ in-place), coefficient and NTT-domain Galois maps,
        keyswitch_accumulate_digit_products_, keyswitch_moddown_qp_to_q
        (ckks_cpu.cpp:710-822)

It must run in its own process: our package registers the same schema
names for ROCm tensors, and the dispatcher allows one definition.  Prints a
JSON summary; exits 1 if any case differs.  ``--plant`` adds one to a single
element of every oracle result first, so every case must then fail: that
shows each comparison can see a one-unit mistake.
"""

from __future__ import annotations

import argparse
import json
import pathlib
import sys

import numpy as np
import torch

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT.parent))

from fhelium_amd import rns  # noqa: E402
from fhelium_amd.config import CkksConfig, Preset  # noqa: E402
from oracle import oracle as O  # noqa: E402

LIB = ROOT / "_ref" / "libfhelium_ref_cpu.so"


def _residues(rng, batch, moduli, n, lazy=False):
    out = np.empty((batch, len(moduli), n), dtype=np.int64)
    for j, q in enumerate(moduli):
        out[:, j, :] = rng.integers(0, int(q) * (2 if lazy else 1), size=(batch, n), dtype=np.int64)
    return out


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a))


def _schedules(log_n):
    """Radix-2 forward DIF and inverse DIT (even, odd, twiddle-index) rows,
    as the reference's indexed plans lay them out
    (fhelium/engine/ntt/plans/indexed_radix2.py:58-126)."""
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


def _expand(compact, idx):
    stages, bf = idx.shape
    return compact.index_select(1, idx.reshape(-1).long()).reshape(compact.size(0), stages, bf)


def cases(log_n: int):
    """Yield (name, reference output, oracle output)."""
    R, N_, C = torch.ops.fhelium_rns_ops, torch.ops.fhelium_ntt_ops, torch.ops.fhelium_ckks_ops
    cfg = CkksConfig.from_preset(Preset.slots32768_scale40_levels34_int64).with_ring_log(log_n)
    t = rns.HostTables(cfg)
    n = cfg.N
    m = np.array(cfg.moduli, dtype=np.int64)
    params = t.rns_params.numpy()
    P = t.rns_params
    rng = np.random.default_rng(2024)

    # element-wise RNS arithmetic
    for op, name in (("mul", "montgomery_mul"), ("add_lazy", "add_lazy"), ("sub_lazy", "sub_lazy"),
                     ("add_canonical", "add_canonical"), ("sub_canonical", "sub_canonical")):
        for rhs_batch in (3, 1):
            a = _residues(rng, 3, m, n, lazy=True)
            b = _residues(rng, rhs_batch, m, n, lazy=True)
            if op.endswith("canonical"):
                a, b = a % m[None, :, None], b % m[None, :, None]
            yield f"{name}[rhs_batch={rhs_batch}]", getattr(R, name)(_t(a), _t(b), P).numpy(), \
                O.binary(op, a, b, params)
            if op.endswith("canonical") and rhs_batch == 1:
                x = _t(a).clone()
                getattr(R, name + "_")(x, _t(b), P)
                yield f"{name}_", x.numpy(), O.binary(op, a, b, params)
    for op, name in (("mul_cyclic", "montgomery_mul_cyclic_compressed"),
                     ("mul_contiguous", "montgomery_mul_contiguous_compressed")):
        a = _residues(rng, 2, m, n, lazy=True)
        b = _residues(rng, 1, m, 64, lazy=True)
        yield name, getattr(R, name)(_t(a), _t(b), P).numpy(), O.binary(op, a, b, params)
    for op, name in (("to_montgomery", "to_montgomery_"), ("from_montgomery", "from_montgomery_"),
                     ("canonicalize", "canonicalize_residues_"), ("center", "center_residues_"),
                     ("shift_positive", "shift_residues_positive_")):
        x = _residues(rng, 2, m, n, lazy=op in ("from_montgomery", "canonicalize"))
        if op == "shift_positive":
            x = x - (m[None, :, None] // 2)
        d = _t(x).clone()
        getattr(R, name)(d, P)
        yield name, d.numpy(), O.unary(op, x, params)
    x = _residues(rng, 2, m, n, lazy=True)
    s = np.array([rng.integers(0, int(q)) for q in m], dtype=np.int64)
    d = _t(x).clone()
    R.montgomery_mul_row_scalars_(d, _t(s), P)
    yield "montgomery_mul_row_scalars_", d.numpy(), O.row_scalars(x, s, params, False)
    yield "montgomery_mul_row_scalars_canonical", R.montgomery_mul_row_scalars_canonical(_t(x), _t(s), P).numpy(), \
        O.row_scalars(x, s, params, True)
    twice = 2 * m
    a, b = _residues(rng, 2, m, n, lazy=True), _residues(rng, 2, m, n, lazy=True)
    yield "add_lazy_with_twice_modulus", R.add_lazy_with_twice_modulus(_t(a), _t(b), _t(twice)).numpy(), \
        O.add_lazy_twice(a, b, twice)
    c = rng.integers(-(1 << 30), 1 << 30, size=(2, n), dtype=np.int64)
    yield "lift_centered_coefficients", R.lift_centered_coefficients(_t(c), _t(twice)).numpy(), O.lift(c, twice)

    # digit decomposition and ModUp at several levels
    for level in (0, 3, 17, 33):
        for k, spec in enumerate(t.layout.digit_specs(level)):
            dt = t.digit_tables[spec.prime_ids]
            dm = np.array([cfg.moduli[i] for i in spec.prime_ids], dtype=np.int64)
            x = _residues(rng, 2, dm, n, lazy=True)
            got = R.mixed_radix_decompose(_t(x), dt.normalizers, dt.propagation, dt.modulus_lo, dt.modulus_hi,
                                          dt.neg_inv_lo, dt.neg_inv_hi).numpy()
            want = O.mixed_radix_decompose(x, dt.normalizers.numpy(), dt.propagation.numpy(),
                                           dt.modulus_lo.numpy(), dt.modulus_hi.numpy(),
                                           dt.neg_inv_lo.numpy(), dt.neg_inv_hi.numpy())
            yield f"mixed_radix_decompose[level={level},digit={k}]", got, want
            rows = len(t.layout.prime_ids(level, include_p=True))
            coef = np.ascontiguousarray(dt.basis_extension.numpy()[:, level:])
            sub = np.ascontiguousarray(params[:, level:])
            coef_t = _t(coef) if coef.size else torch.empty(0, rows, dtype=torch.int64)
            got = R.mixed_radix_basis_extend_to_montgomery(_t(want), coef_t, _t(sub), rows).numpy()
            yield f"mixed_radix_basis_extend_to_montgomery[level={level},digit={k}]", got, \
                O.basis_extend(want, coef, sub, rows)

    # indexed NTT: forward into Montgomery form, then the four inverse forms
    fwd, inv = rns.compact_twiddles_montgomery(cfg.moduli, log_n)
    fe, fo, ft, ie, io, it = _schedules(log_n)
    ftw, itw = _expand(fwd, ft), _expand(inv, it)
    x = _residues(rng, 2, m, n)
    d = _t(x).clone()
    N_.forward_ntt_to_montgomery_indexed_(d, fe, fo, ftw, P)
    want_f = O.ntt_forward(x, fwd.numpy(), params, to_montgomery=True)
    yield "forward_ntt_to_montgomery_indexed_", d.numpy(), want_f
    yield "forward_ntt_to_montgomery_indexed", N_.forward_ntt_to_montgomery_indexed(_t(x), fe, fo, ftw, P).numpy(), \
        want_f
    xl = _residues(rng, 2, m, n, lazy=True)
    d = _t(xl).clone()
    N_.forward_ntt_montgomery_indexed_(d, fe, fo, ftw, P)
    yield "forward_ntt_montgomery_indexed_", d.numpy(), O.ntt_forward(xl, fwd.numpy(), params, to_montgomery=False)
    for form, name in enumerate(["inverse_ntt_montgomery_indexed_", "inverse_ntt_to_standard_lazy_indexed_",
                                 "inverse_ntt_to_standard_indexed_", "inverse_ntt_to_centered_indexed_"]):
        d = _t(want_f).clone()
        getattr(N_, name)(d, ie, io, itw, P)
        yield name, d.numpy(), O.ntt_inverse(want_f, inv.numpy(), params, form)

    # CKKS: plaintext adds
    for layout, name in (("dense", "add_prepared_plaintext_component"),
                         ("cyclic", "add_cyclic_compressed_plaintext_component"),
                         ("contiguous", "add_contiguous_compressed_plaintext_component"),
                         ("strided", "add_strided_plaintext_component")):
        ct = _residues(rng, 2, m, n)
        pt = _residues(rng, 1, m, n if layout == "dense" else 256)
        imp = np.array([[rng.integers(0, int(q)) for q in m]], dtype=np.int64) if layout == "strided" else None
        args = (_t(ct), _t(pt)) + ((_t(imp),) if imp is not None else ()) + (P,)
        want = O.add_plaintext(layout, ct, pt, params, imp)
        yield name, getattr(C, name)(*args).numpy(), want
        d = _t(ct).clone()
        getattr(C, name + "_")(d, *args[1:])
        yield name + "_", d.numpy(), want

    # rescale at three levels, nearest and truncate, functional and in place
    for level in (0, 20, 32):
        ids = t.layout.prime_ids(level)
        qm = m[list(ids)]
        x = _residues(rng, 2, qm, n)
        invs = t.rescale_inverses[level][: len(ids) - 1].numpy()
        sub = np.ascontiguousarray(params[:, ids[1]: ids[-1] + 1])
        half = int(qm[0]) // 2
        rem, drop = np.ascontiguousarray(x[:, 1:]), np.ascontiguousarray(x[:, 0])
        for nearest in (True, False):
            kind = "nearest" if nearest else "truncate"
            want = O.rescale(rem, drop, invs, sub, half, nearest)
            extra = (half,) if nearest else ()
            got = getattr(C, f"rescale_drop_leading_prime_{kind}")(_t(rem), _t(invs), _t(drop), _t(sub), *extra)
            yield f"rescale_drop_leading_prime_{kind}[level={level}]", got.numpy(), want
            d = _t(rem).clone()
            getattr(C, f"rescale_drop_leading_prime_{kind}_")(d, _t(invs), _t(drop), _t(sub), *extra)
            yield f"rescale_drop_leading_prime_{kind}_[level={level}]", d.numpy(), want

    # Galois automorphisms for rotations and conjugation
    for step in (1, 5, -3, 1000, None):
        g = 2 * n - 1 if step is None else rns.rotation_galois_element(n, step)
        idx, sgn = rns.coefficient_galois_tables(n, g)
        x = _residues(rng, 2, m, n)
        got = C.apply_coefficient_galois_automorphism(_t(x), idx, sgn, _t(twice)).numpy()
        yield f"apply_coefficient_galois_automorphism[g={g}]", got, \
            O.galois_coefficient(x, idx.numpy(), sgn.numpy(), twice)
        src = rns.ntt_galois_indices(log_n, g)
        yield f"apply_ntt_galois_automorphism[g={g}]", C.apply_ntt_galois_automorphism(_t(x), src).numpy(), \
            O.galois_ntt(x, src.numpy())

    # key-switch inner product and ModDown
    for level in (0, 9, 33):
        rows = list(t.layout.prime_ids(level, include_p=True))
        qm = m[rows]
        sub = np.ascontiguousarray(params[:, level:])
        acc0, acc1 = _residues(rng, 2, qm, n, lazy=True), _residues(rng, 2, qm, n, lazy=True)
        digit = _residues(rng, 2, qm, n, lazy=True)
        key = np.stack([_residues(rng, 1, m, n, lazy=True)[0] for _ in range(2)])
        d0, d1 = _t(acc0).clone(), _t(acc1).clone()
        C.keyswitch_accumulate_digit_products_(d0, d1, _t(digit), _t(key), _t(sub), level)
        w0, w1 = O.keyswitch_mac(acc0, acc1, digit, key, sub, level)
        yield f"keyswitch_accumulate_digit_products_[level={level},acc0]", d0.numpy(), w0
        yield f"keyswitch_accumulate_digit_products_[level={level},acc1]", d1.numpy(), w1
        p = cfg.num_p_primes
        qp = _residues(rng, 2, qm, n)
        invs = t.moddown_inverses[level].numpy()
        qpart, ppart = np.ascontiguousarray(qp[:, :-p]), np.ascontiguousarray(qp[:, -p:])
        got = C.keyswitch_moddown_qp_to_q(_t(qpart), _t(ppart), _t(invs), _t(sub)).numpy()
        yield f"keyswitch_moddown_qp_to_q[level={level}]", got, O.moddown(qpart, ppart, invs, sub)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--log-n", type=int, default=10)
    ap.add_argument("--plant", action="store_true", help="perturb every oracle result by one unit")
    args = ap.parse_args(argv)
    if not LIB.exists():
        print(json.dumps({"error": f"{LIB} not built (python oracle/build_ref.py)"}))
        return 2
    torch.ops.load_library(str(LIB))
    results = []
    for name, got, want in cases(args.log_n):
        want = np.array(want, dtype=np.int64, copy=True)
        if args.plant:
            want.reshape(-1)[want.size // 2] += 1
        same = got.shape == want.shape and bool(np.array_equal(got, want))
        results.append({"case": name, "equal": same, "shape": list(got.shape)})
    bad = [r["case"] for r in results if not r["equal"]]
    print(json.dumps({"log_n": args.log_n, "plant": args.plant, "cases": len(results),
                      "mismatches": bad, "results": results}))
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
