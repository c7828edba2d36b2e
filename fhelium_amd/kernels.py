"""Tensor-level entry points over the gfx950 C ABI.

Each function validates like the reference's native op (``TORCH_CHECK`` ->
``RuntimeError``), allocates functional outputs with ``torch.empty_like``,
builds strided views without copying, and launches on PyTorch's current
stream of the operand's device.  The reference op each one backs is named in
its docstring; ``fhelium_amd.native.ops`` registers them under the
reference's unchanged schemas.
"""

from __future__ import annotations

import ctypes

import torch

from fhelium_amd import _native as nat

_BINARY = {"mul": 0, "add_lazy": 1, "sub_lazy": 2, "add_canonical": 3, "sub_canonical": 4,
           "mul_cyclic": 5, "mul_contiguous": 6}
_UNARY = {"to_montgomery": 0, "from_montgomery": 1, "canonicalize": 2, "center": 3,
          "shift_positive": 4}
_PLAIN = {"dense": 0, "cyclic": 1, "contiguous": 2, "strided": 3}


def _peers(op: str, ref: torch.Tensor, *others: torch.Tensor) -> None:
    nat.require_device(ref)
    if ref.dtype != torch.int64:
        raise RuntimeError(f"{op} supports int64 residues on gfx950 (got {ref.dtype})")
    for o in others:
        if o.device != ref.device:
            raise RuntimeError(f"{op} requires every tensor on {ref.device}")
        if o.dtype != torch.int64:
            raise RuntimeError(f"{op} requires int64 tables and operands")


def _check_params(op: str, limbs: int, params: torch.Tensor) -> None:
    if params.dim() != 2 or params.size(0) < 8:
        raise RuntimeError(f"{op} requires [parameter, limb] RNS parameters")
    if params.size(1) != limbs:
        raise RuntimeError(f"{op} operand and parameter limb counts differ: {limbs} vs {params.size(1)}")


def _check_row_vector(op: str, v: torch.Tensor, limbs: int, name: str) -> None:
    if v.dim() != 1:
        raise RuntimeError(f"{op} requires {name} to be a rank-one limb vector")
    if v.size(0) != limbs:
        raise RuntimeError(f"{op} {name} count must match limbs: {v.size(0)} vs {limbs}")


def _overlaps(a: torch.Tensor, b: torch.Tensor) -> bool:
    if a.untyped_storage().data_ptr() != b.untyped_storage().data_ptr():
        return False
    a0, a1 = a.data_ptr(), a.data_ptr() + a.numel() * a.element_size()
    b0, b1 = b.data_ptr(), b.data_ptr() + b.numel() * b.element_size()
    return a0 < b1 and b0 < a1


def _no_overlap(op: str, out: torch.Tensor, *read_only: torch.Tensor) -> None:
    for r in read_only:
        if r is not None and _overlaps(out, r):
            raise RuntimeError(f"{op}: output storage overlaps a read-only operand")


def _binary(opname: str, op: str, lhs: torch.Tensor, rhs: torch.Tensor, params: torch.Tensor,
            out: torch.Tensor | None = None) -> torch.Tensor:
    _peers(opname, lhs, rhs, params)
    left = nat.batch_view(lhs, "lhs")
    right = nat.batch_view(rhs, "rhs")
    compressed = op in ("mul_cyclic", "mul_contiguous")
    if left.size(1) != right.size(1):
        raise RuntimeError(f"{opname} operand limb counts differ: {left.size(1)} vs {right.size(1)}")
    if right.size(0) not in (left.size(0), 1):
        raise RuntimeError(f"{opname} operand batch counts differ: {left.size(0)} vs {right.size(0)}")
    n = left.size(2)
    if compressed:
        s = right.size(2)
        if not (0 < s < n and (s & (s - 1)) == 0 and n % s == 0):
            raise RuntimeError(f"{opname} requires a power-of-two compressed extent dividing N")
    elif right.size(2) != n:
        raise RuntimeError(f"{opname} operand coefficient counts differ: {n} vs {right.size(2)}")
    _check_params(opname, left.size(1), params)
    if out is None:
        out = torch.empty_like(lhs)
    else:
        _no_overlap(opname, out, rhs, params)
    o = nat.batch_view(out, "out")
    nat.call("fh_rns_binary", _BINARY[op], nat.v3(o), nat.v3(left),
             nat.v3(right, broadcast_batch=right.size(0) == 1 and left.size(0) != 1),
             nat.t2(params), left.size(0), left.size(1), n, right.size(2),
             nat.stream_ptr(lhs.device))
    return out


# -- fhelium_rns_ops (csrc/ops/rns/rns_arithmetic.cpp:30-72) -----------------

def montgomery_mul(lhs, rhs, rns_params):
    """montgomery_mul: a b R^{-1} mod q (canonical, like the CPU backend)."""
    return _binary("rns_montgomery_mul", "mul", lhs, rhs, rns_params)


def montgomery_mul_cyclic_compressed(lhs, compressed_rhs, rns_params):
    return _binary("rns_montgomery_mul_cyclic_compressed", "mul_cyclic", lhs, compressed_rhs, rns_params)


def montgomery_mul_contiguous_compressed(lhs, compressed_rhs, rns_params):
    return _binary("rns_montgomery_mul_contiguous_compressed", "mul_contiguous", lhs, compressed_rhs, rns_params)


def add_lazy(lhs, rhs, rns_params):
    return _binary("rns_add_lazy", "add_lazy", lhs, rhs, rns_params)


def sub_lazy(lhs, rhs, rns_params):
    return _binary("rns_sub_lazy", "sub_lazy", lhs, rhs, rns_params)


def add_canonical(lhs, rhs, rns_params):
    return _binary("rns_add_canonical", "add_canonical", lhs, rhs, rns_params)


def add_canonical_(lhs, rhs, rns_params):
    _binary("rns_add_canonical_", "add_canonical", lhs, rhs, rns_params, out=lhs)


def sub_canonical(lhs, rhs, rns_params):
    return _binary("rns_sub_canonical", "sub_canonical", lhs, rhs, rns_params)


def sub_canonical_(lhs, rhs, rns_params):
    _binary("rns_sub_canonical_", "sub_canonical", lhs, rhs, rns_params, out=lhs)


def _unary(opname: str, op: str, x: torch.Tensor, params: torch.Tensor) -> None:
    _peers(opname, x, params)
    rows = nat.batch_view(x, "residues")
    _check_params(opname, rows.size(1), params)
    _no_overlap(opname, x, params)
    nat.call("fh_rns_unary_inplace", _UNARY[op], nat.v3(rows), nat.t2(params), rows.size(0),
             rows.size(1), rows.size(2), nat.stream_ptr(x.device))


def to_montgomery_(standard_residues, rns_params):
    _unary("rns_to_montgomery", "to_montgomery", standard_residues, rns_params)


def from_montgomery_(montgomery_residues, rns_params):
    _unary("rns_from_montgomery", "from_montgomery", montgomery_residues, rns_params)


def canonicalize_residues_(lazy_residues, rns_params):
    _unary("rns_canonicalize", "canonicalize", lazy_residues, rns_params)


def center_residues_(canonical_residues, rns_params):
    _unary("rns_center_residues", "center", canonical_residues, rns_params)


def shift_residues_positive_(centered_residues, rns_params):
    _unary("rns_shift_positive", "shift_positive", centered_residues, rns_params)


def _row_scalars(opname, residues, row_scalars, params, out, canonical):
    _peers(opname, residues, row_scalars, params)
    rows = nat.batch_view(residues, "residues")
    _check_params(opname, rows.size(1), params)
    _check_row_vector(opname, row_scalars, rows.size(1), "row_scalars")
    o = nat.batch_view(out, "out")
    nat.call("fh_rns_row_scalars", nat.v3(o), nat.v3(rows), nat.vec(row_scalars), nat.t2(params),
             rows.size(0), rows.size(1), rows.size(2), int(canonical), nat.stream_ptr(residues.device))


def montgomery_mul_row_scalars_(residues, row_scalars, rns_params):
    _no_overlap("rns_montgomery_mul_row_scalars", residues, row_scalars, rns_params)
    _row_scalars("rns_montgomery_mul_row_scalars", residues, row_scalars, rns_params, residues, False)


def montgomery_mul_row_scalars_canonical(residues, row_scalars, rns_params):
    out = torch.empty_like(residues)
    _row_scalars("rns_montgomery_mul_row_scalars_canonical", residues, row_scalars, rns_params, out, True)
    return out


def add_lazy_with_twice_modulus(lhs, rhs, twice_modulus):
    op = "rns_add_lazy_with_twice_modulus"
    _peers(op, lhs, rhs, twice_modulus)
    left = nat.batch_view(lhs, "lhs")
    right = nat.batch_view(rhs, "rhs")
    if left.shape[1:] != right.shape[1:] or right.size(0) not in (left.size(0), 1):
        raise RuntimeError(f"{op} operand shapes differ")
    _check_row_vector(op, twice_modulus, left.size(1), "twice_modulus")
    out = torch.empty_like(lhs)
    o = nat.batch_view(out, "out")
    nat.call("fh_rns_add_lazy_twice_modulus", nat.v3(o), nat.v3(left),
             nat.v3(right, broadcast_batch=right.size(0) == 1 and left.size(0) != 1),
             nat.vec(twice_modulus), left.size(0), left.size(1), left.size(2),
             nat.stream_ptr(lhs.device))
    return out


def lift_centered_coefficients(centered_coefficients, twice_modulus):
    op = "rns_lift_centered_coefficients"
    _peers(op, centered_coefficients, twice_modulus)
    src = nat.coef_batch_view(centered_coefficients, "coefficients")
    if twice_modulus.dim() != 1:
        raise RuntimeError(f"{op} requires a rank-one twice_modulus")
    shape = list(centered_coefficients.shape)
    shape.insert(len(shape) - 1, twice_modulus.size(0))
    out = centered_coefficients.new_empty(shape)
    o = nat.batch_view(out, "out")
    nat.call("fh_rns_lift_centered", nat.v3(o), nat.v3_from_2d(src), nat.vec(twice_modulus),
             src.size(0), twice_modulus.size(0), src.size(1), nat.stream_ptr(out.device))
    return out


def mixed_radix_decompose(source_residues, mixed_radix_normalizers,
                          mixed_radix_propagation_coefficients, modulus_lo, modulus_hi,
                          neg_inv_modulus_lo, neg_inv_modulus_hi):
    op = "mixed_radix_decompose"
    _peers(op, source_residues, modulus_lo, modulus_hi, neg_inv_modulus_lo, neg_inv_modulus_hi)
    src = nat.batch_view(source_residues, "source_residues")
    rows = src.size(1)
    if rows > 8:
        raise RuntimeError(f"{op} supports at most 8 rows")
    if rows > 1:
        _check_row_vector(op, mixed_radix_normalizers, rows - 1, "normalizers")
        p = mixed_radix_propagation_coefficients
        if p.dim() != 2 or p.size(0) != rows - 1 or p.size(1) != rows:
            raise RuntimeError(f"{op} propagation table shape mismatch")
    out = torch.empty_like(source_residues)
    o = nat.batch_view(out, "out")
    nat.call("fh_rns_mixed_radix_decompose", nat.v3(o), nat.v3(src), src.size(0), rows,
             src.size(2), nat.vec(mixed_radix_normalizers if rows > 1 else None),
             nat.t2(mixed_radix_propagation_coefficients if rows > 1 else None),
             nat.vec(modulus_lo), nat.vec(modulus_hi), nat.vec(neg_inv_modulus_lo),
             nat.vec(neg_inv_modulus_hi), nat.stream_ptr(out.device))
    return out


def mixed_radix_basis_extend_to_montgomery(mixed_radix_components, basis_extension_coefficients,
                                           rns_params, destination_row_count, *, matrix_cores: bool = False):
    """``matrix_cores``: run the byte-sliced INT8 MFMA trial kernel
    (fh_rns_basis_extend_mfma, 1..4 digits, N a multiple of 128) instead of
    the scalar one; the output is bit-identical."""
    op = "mixed_radix_basis_extend_to_montgomery"
    if destination_row_count <= 0:
        raise RuntimeError("destination_row_count must be positive")
    _peers(op, mixed_radix_components, basis_extension_coefficients, rns_params)
    digits = nat.batch_view(mixed_radix_components, "components")
    c = basis_extension_coefficients
    if c.dim() != 2 or c.size(0) != digits.size(1) - 1 or (
            digits.size(1) != 1 and c.size(1) != destination_row_count):
        raise RuntimeError(f"{op} coefficient table shape mismatch")
    shape = list(mixed_radix_components.shape)
    shape[-2] = destination_row_count
    out = torch.empty(shape, dtype=torch.int64, device=mixed_radix_components.device)
    o = nat.batch_view(out, "out")
    _check_params(op, destination_row_count, rns_params)
    nat.call("fh_rns_basis_extend_mfma" if matrix_cores else "fh_rns_basis_extend", nat.v3(o), nat.v3(digits),
             digits.size(0), digits.size(1), destination_row_count, digits.size(2),
             nat.t2(c if digits.size(1) > 1 else None), nat.t2(rns_params), nat.stream_ptr(out.device))
    return out


# -- fhelium_ckks_ops (csrc/ops/ckks/ckks.cpp:29-90) --------------------------

def _rescale(opname, remaining_residues, drop_prime_inverse_mont, dropped_residue, rns_params,
             half_drop_prime, nearest, out):
    _peers(opname, remaining_residues, drop_prime_inverse_mont, dropped_residue, rns_params)
    rem = nat.batch_view(remaining_residues, "remaining_residues")
    dropped = nat.coef_batch_view(dropped_residue, "dropped_residue")
    _check_params(opname, rem.size(1), rns_params)
    _check_row_vector(opname, drop_prime_inverse_mont, rem.size(1), "inverse")
    if dropped.size(0) != rem.size(0) or dropped.size(1) != rem.size(2):
        raise RuntimeError(f"{opname} dropped residue shape mismatch")
    _no_overlap(opname, out, drop_prime_inverse_mont, dropped_residue, rns_params)
    o = nat.batch_view(out, "out")
    nat.call("fh_ckks_rescale", nat.v3(o), nat.v3(rem), nat.v3_from_2d(dropped),
             nat.vec(drop_prime_inverse_mont), nat.t2(rns_params), rem.size(0), rem.size(1),
             rem.size(2), int(half_drop_prime), int(nearest), nat.stream_ptr(out.device))


def rescale_drop_leading_prime_nearest(remaining_residues, drop_prime_inverse_mont,
                                       dropped_residue, rns_params, half_drop_prime):
    out = torch.empty_like(remaining_residues)
    _rescale("ckks_rescale_nearest", remaining_residues, drop_prime_inverse_mont, dropped_residue,
             rns_params, half_drop_prime, True, out)
    return out


def rescale_drop_leading_prime_nearest_(remaining_residues, drop_prime_inverse_mont,
                                        dropped_residue, rns_params, half_drop_prime):
    _rescale("ckks_rescale_nearest_", remaining_residues, drop_prime_inverse_mont, dropped_residue,
             rns_params, half_drop_prime, True, remaining_residues)


def rescale_drop_leading_prime_truncate(remaining_residues, drop_prime_inverse_mont,
                                        dropped_residue, rns_params):
    out = torch.empty_like(remaining_residues)
    _rescale("ckks_rescale_truncate", remaining_residues, drop_prime_inverse_mont, dropped_residue,
             rns_params, 0, False, out)
    return out


def rescale_drop_leading_prime_truncate_(remaining_residues, drop_prime_inverse_mont,
                                         dropped_residue, rns_params):
    _rescale("ckks_rescale_truncate_", remaining_residues, drop_prime_inverse_mont, dropped_residue,
             rns_params, 0, False, remaining_residues)


def _check_index_table(op, idx: torch.Tensor, n: int, device) -> None:
    if idx.device != device or idx.dtype != torch.int32 or idx.dim() != 1 or idx.size(0) != n:
        raise RuntimeError(f"{op} gather table shape or dtype mismatch")


def apply_coefficient_galois_automorphism(residues, source_indices, source_sign, twice_modulus):
    op = "apply_coefficient_galois_automorphism"
    _peers(op, residues, twice_modulus)
    x = nat.batch_view(residues, "residues")
    _check_index_table(op, source_indices, x.size(2), residues.device)
    if source_sign.dtype != torch.int8 or source_sign.dim() != 1 or source_sign.size(0) != x.size(2):
        raise RuntimeError(f"{op} gather table shape or dtype mismatch")
    _check_row_vector(op, twice_modulus, x.size(1), "twice_modulus")
    out = torch.empty_like(residues)
    o = nat.batch_view(out, "out")
    nat.call("fh_ckks_galois_coefficient", nat.v3(o), nat.v3(x), source_indices.data_ptr(),
             source_indices.stride(0), source_sign.data_ptr(), source_sign.stride(0),
             nat.vec(twice_modulus), x.size(0), x.size(1), x.size(2), nat.stream_ptr(out.device))
    return out


def apply_ntt_galois_automorphism(residues_ntt, source_indices):
    op = "apply_ntt_galois_automorphism"
    _peers(op, residues_ntt)
    x = nat.batch_view(residues_ntt, "residues_ntt")
    _check_index_table(op, source_indices, x.size(2), residues_ntt.device)
    out = torch.empty_like(residues_ntt)
    o = nat.batch_view(out, "out")
    nat.call("fh_ckks_galois_ntt", nat.v3(o), nat.v3(x), source_indices.data_ptr(),
             source_indices.stride(0), x.size(0), x.size(1), x.size(2), nat.stream_ptr(out.device))
    return out


def keyswitch_moddown_qp_to_q(q_residues, p_residues, moddown_p_drop_inverses_montgomery, rns_params):
    op = "keyswitch_moddown_qp_to_q"
    _peers(op, q_residues, p_residues, moddown_p_drop_inverses_montgomery, rns_params)
    q = nat.batch_view(q_residues, "q_residues")
    p = nat.batch_view(p_residues, "p_residues")
    if q.size(0) != p.size(0) or q.size(2) != p.size(2):
        raise RuntimeError(f"{op} Q/P batch and coefficient shapes differ")
    if rns_params.dim() != 2 or rns_params.size(1) != q.size(1) + p.size(1):
        raise RuntimeError(f"{op} RNS parameter row count mismatch")
    inv = moddown_p_drop_inverses_montgomery
    if inv.dim() != 2 or inv.size(0) != p.size(1) or inv.size(1) < q.size(1) + p.size(1) - 1:
        raise RuntimeError(f"{op} inverse table shape mismatch")
    out = torch.empty_like(q_residues)
    o = nat.batch_view(out, "out")
    nat.call("fh_ckks_moddown", nat.v3(o), nat.v3(q), nat.v3(p), nat.t2(inv), nat.t2(rns_params),
             q.size(0), q.size(1), p.size(1), q.size(2), nat.stream_ptr(out.device))
    return out


def keyswitch_accumulate_digit_products_(accumulator0_qp, accumulator1_qp, extended_digit_ntt_qp,
                                         key_switch_key_digit, rns_params, key_digit_row_start):
    op = "keyswitch_accumulate_digit_products"
    _peers(op, extended_digit_ntt_qp, accumulator0_qp, accumulator1_qp, key_switch_key_digit, rns_params)
    a0 = nat.batch_view(accumulator0_qp, "accumulator0_qp")
    a1 = nat.batch_view(accumulator1_qp, "accumulator1_qp")
    d = nat.batch_view(extended_digit_ntt_qp, "digit_qp")
    if a0.shape != d.shape or a1.shape != d.shape:
        raise RuntimeError(f"{op} accumulator and digit shapes differ")
    _check_params(op, d.size(1), rns_params)
    k = key_switch_key_digit
    if (k.dim() != 3 or k.size(0) != 2 or key_digit_row_start < 0
            or key_digit_row_start + d.size(1) > k.size(1) or k.size(2) != d.size(2)):
        raise RuntimeError(f"{op} key digit shape or row interval mismatch")
    if _overlaps(accumulator0_qp, accumulator1_qp):
        raise RuntimeError(f"{op}: accumulators overlap")
    for acc in (accumulator0_qp, accumulator1_qp):
        _no_overlap(op, acc, extended_digit_ntt_qp, key_switch_key_digit, rns_params)
    nat.call("fh_ckks_keyswitch_accumulate", nat.v3(a0), nat.v3(a1), nat.v3(d), nat.v3(k),
             nat.t2(rns_params), int(key_digit_row_start), d.size(0), d.size(1), d.size(2),
             nat.stream_ptr(d.device))


def keyswitch_accumulate_multi_key_(accumulator0_qp, accumulator1_qp, extended_digit_ntt_qp,
                                    keys: list[torch.Tensor], key_bases: torch.Tensor,
                                    digit_index: int, rns_params, key_digit_row_start):
    """keyswitch_accumulate_digit_products_ with keys[b] for batch entry b.

    ``keys`` are same-shaped [digit, 2, QP, N] tensors (kept alive by the
    caller); ``key_bases`` is the int64 device tensor of their addresses."""

    op = "keyswitch_accumulate_multi_key"
    _peers(op, extended_digit_ntt_qp, accumulator0_qp, accumulator1_qp, rns_params)
    a0 = nat.batch_view(accumulator0_qp, "accumulator0_qp")
    a1 = nat.batch_view(accumulator1_qp, "accumulator1_qp")
    d = nat.batch_view(extended_digit_ntt_qp, "digit_qp")
    if a0.shape != d.shape or a1.shape != d.shape or len(keys) != d.size(0):
        raise RuntimeError(f"{op}: one key per batch entry and matching accumulator shapes required")
    _check_params(op, d.size(1), rns_params)
    k = keys[0]
    for other in keys:
        if other.shape != k.shape or other.stride() != k.stride() or other.device != k.device:
            raise RuntimeError(f"{op}: keys must share shape, strides and device")
    if (k.dim() != 4 or k.size(1) != 2 or not 0 <= digit_index < k.size(0)
            or key_digit_row_start + d.size(1) > k.size(2) or k.size(3) != d.size(2)):
        raise RuntimeError(f"{op} key shape or row interval mismatch")
    if key_bases.dtype != torch.int64 or key_bases.numel() != d.size(0) or key_bases.device != d.device:
        raise RuntimeError(f"{op}: key_bases must be an int64 device vector of length batch")
    nat.call("fh_ckks_keyswitch_accumulate_multi_key", nat.v3(a0), nat.v3(a1), nat.v3(d),
             key_bases.data_ptr(), digit_index, k.stride(0), k.stride(1), k.stride(2),
             nat.t2(rns_params), int(key_digit_row_start), d.size(0), d.size(1), d.size(2),
             nat.stream_ptr(d.device))


def _add_plaintext(opname, layout, ciphertext_component, plaintext, rns_params, out, implicit=None):
    _peers(opname, ciphertext_component, plaintext, rns_params)
    ct = nat.batch_view(ciphertext_component, "ciphertext_component")
    pt = nat.batch_view(plaintext, "prepared_plaintext")
    if pt.size(1) != ct.size(1) or pt.size(0) not in (ct.size(0), 1):
        raise RuntimeError(f"{opname} plaintext shape mismatch")
    n = ct.size(2)
    support = pt.size(2)
    if layout == "dense":
        if support != n:
            raise RuntimeError(f"{opname} operand coefficient counts differ")
    elif not (0 < support < n and (support & (support - 1)) == 0 and n % support == 0):
        raise RuntimeError(f"{opname} compressed support must be a power-of-two divisor of N")
    if layout == "strided":
        if implicit is None or implicit.numel() != pt.size(0) * pt.size(1) or not implicit.is_contiguous():
            raise RuntimeError(f"{opname} implicit plaintext shape mismatch")
    _check_params(opname, ct.size(1), rns_params)
    _no_overlap(opname, out, plaintext, rns_params, implicit)
    o = nat.batch_view(out, "out")
    nat.call("fh_ckks_add_plaintext", _PLAIN[layout], nat.v3(o), nat.v3(ct),
             nat.v3(pt, broadcast_batch=pt.size(0) == 1 and ct.size(0) != 1),
             implicit.data_ptr() if implicit is not None else None, pt.size(0), nat.t2(rns_params),
             ct.size(0), ct.size(1), n, support, nat.stream_ptr(out.device))


def add_prepared_plaintext_component(ciphertext_component, prepared_plaintext, rns_params):
    out = torch.empty_like(ciphertext_component)
    _add_plaintext("ckks_add_prepared_plaintext_component", "dense", ciphertext_component,
                   prepared_plaintext, rns_params, out)
    return out


def add_prepared_plaintext_component_(ciphertext_component, prepared_plaintext, rns_params):
    _add_plaintext("ckks_add_prepared_plaintext_component_", "dense", ciphertext_component,
                   prepared_plaintext, rns_params, ciphertext_component)


def add_cyclic_compressed_plaintext_component(ciphertext_component, compressed_plaintext, rns_params):
    out = torch.empty_like(ciphertext_component)
    _add_plaintext("ckks_add_cyclic_plaintext", "cyclic", ciphertext_component, compressed_plaintext,
                   rns_params, out)
    return out


def add_cyclic_compressed_plaintext_component_(ciphertext_component, compressed_plaintext, rns_params):
    _add_plaintext("ckks_add_cyclic_plaintext_", "cyclic", ciphertext_component, compressed_plaintext,
                   rns_params, ciphertext_component)


def add_contiguous_compressed_plaintext_component(ciphertext_component, compressed_plaintext, rns_params):
    out = torch.empty_like(ciphertext_component)
    _add_plaintext("ckks_add_contiguous_plaintext", "contiguous", ciphertext_component,
                   compressed_plaintext, rns_params, out)
    return out


def add_contiguous_compressed_plaintext_component_(ciphertext_component, compressed_plaintext, rns_params):
    _add_plaintext("ckks_add_contiguous_plaintext_", "contiguous", ciphertext_component,
                   compressed_plaintext, rns_params, ciphertext_component)


def add_strided_plaintext_component(ciphertext_component, strided_plaintext, implicit_plaintext, rns_params):
    out = torch.empty_like(ciphertext_component)
    _add_plaintext("ckks_add_strided_plaintext", "strided", ciphertext_component, strided_plaintext,
                   rns_params, out, implicit_plaintext)
    return out


def add_strided_plaintext_component_(ciphertext_component, strided_plaintext, implicit_plaintext, rns_params):
    _add_plaintext("ckks_add_strided_plaintext_", "strided", ciphertext_component, strided_plaintext,
                   rns_params, ciphertext_component, implicit_plaintext)


# -- fused engine helpers (bit-identical to the reference compositions) ------

def tensor_product(out: torch.Tensor, lhs: torch.Tensor, rhs: torch.Tensor, rns_params: torch.Tensor) -> None:
    """out[0..2] = (a0 b0, add_lazy(a0 b1, a1 b0), a1 b1) for lhs/rhs [2, *b, limb, N]
    (ckks_engine.py:1583-1598 as one pass)."""

    op = "ckks_tensor_product"
    _peers(op, lhs, rhs, out, rns_params)
    if lhs.size(0) != 2 or rhs.size(0) != 2 or out.size(0) != 3 or lhs.shape[1:] != rhs.shape[1:] \
            or out.shape[1:] != lhs.shape[1:]:
        raise RuntimeError(f"{op}: expects two 2-component operands and a 3-component output")
    v = [nat.batch_view(t, "operand") for t in (out[0], out[1], out[2], lhs[0], lhs[1], rhs[0], rhs[1])]
    _check_params(op, v[3].size(1), rns_params)
    nat.call("fh_ckks_tensor_product", *[nat.v3(x) for x in v], nat.t2(rns_params), v[3].size(0),
             v[3].size(1), v[3].size(2), nat.stream_ptr(lhs.device))


def add_two_canonicalize(out: torch.Tensor, d0: torch.Tensor, d1: torch.Tensor, corr0: torch.Tensor,
                         corr1: torch.Tensor, rns_params: torch.Tensor) -> None:
    """out = canonicalize(d + corr) per component (ckks_engine.py:1664-1667).
    For canonical d and corr the canonical add computes exactly that."""

    _binary("rns_add_canonical", "add_canonical", d0, corr0, rns_params, out=out[0])
    _binary("rns_add_canonical", "add_canonical", d1, corr1, rns_params, out=out[1])


def rescale_components(data: torch.Tensor, inverse: torch.Tensor, remaining_params: torch.Tensor,
                       half_drop_prime: int) -> torch.Tensor:
    """Nearest rescale of every component of [C, *b, limb, N] (ckks_rescale.py:339-377)."""

    return rescale_drop_leading_prime_nearest(data[..., 1:, :], inverse, data[..., 0, :],
                                              remaining_params, half_drop_prime)


MAX_COMBINATION_TERMS = 8


def scalar_combination(terms: list[torch.Tensor], scalars: torch.Tensor, rns_params: torch.Tensor) -> torch.Tensor:
    """Terms of equal shape [C, E, limb, N] (any strides, unit coefficient
    stride; canonical residues) and Montgomery scalars [T, limb]
    (m_t R mod q) -> [C, E, limb, N]: sum_t x_t * m_t mod q, one pass
    (fh_ckks_scalar_combination).  A term at a lower level is passed as the
    row slice at the output's level."""

    op = "ckks_scalar_combination"
    if not 1 <= len(terms) <= MAX_COMBINATION_TERMS:
        raise RuntimeError(f"{op}: takes 1..{MAX_COMBINATION_TERMS} terms, got {len(terms)}")
    first = terms[0]
    _peers(op, first, *terms[1:], scalars, rns_params)
    if first.dim() != 4 or any(t.shape != first.shape for t in terms):
        raise RuntimeError(f"{op}: every term must be [C, E, limb, N] of one shape")
    if any(t.stride(-1) != 1 for t in terms):
        raise RuntimeError(f"{op}: terms need a unit coefficient stride")
    c, e, limbs, n = first.shape
    _check_params(op, limbs, rns_params)
    if scalars.dim() != 2 or scalars.shape != (len(terms), limbs) or scalars.stride(1) != 1:
        raise RuntimeError(f"{op}: scalars must be [{len(terms)}, {limbs}] with unit limb stride")
    out = torch.empty((c, e, limbs, n), dtype=torch.int64, device=first.device)
    ptrs = (ctypes.c_void_p * len(terms))(*[t.data_ptr() for t in terms])
    strides = (ctypes.c_int64 * (3 * len(terms)))(*[x for t in terms for x in t.stride()[:3]])
    nat.call("fh_ckks_scalar_combination", out.data_ptr(), out.stride(0), out.stride(1), out.stride(2),
             ctypes.addressof(ptrs), ctypes.addressof(strides), scalars.data_ptr(), scalars.stride(0), len(terms),
             c, e, limbs, n, nat.t2(rns_params), nat.stream_ptr(first.device))
    return out


def plaintext_inner_product(ciphertexts: torch.Tensor, plaintexts: torch.Tensor, rns_params: torch.Tensor) -> torch.Tensor:
    """[C, B, limb, N] NTT-domain ciphertext entries and [B, limb, N]
    NTT-domain plaintexts -> [C, limb, N]: sum_b ct_b * pt_b (Montgomery
    products, canonical sum) in one pass (fh_ckks_plaintext_inner_product)."""

    op = "ckks_plaintext_inner_product"
    _peers(op, ciphertexts, plaintexts, rns_params)
    if ciphertexts.dim() != 4 or plaintexts.dim() != 3 or plaintexts.shape != ciphertexts.shape[1:]:
        raise RuntimeError(f"{op}: expects [C, B, limb, N] ciphertexts and [B, limb, N] plaintexts")
    if ciphertexts.stride(-1) != 1 or plaintexts.stride(-1) != 1:
        raise RuntimeError(f"{op}: operands need a unit coefficient stride")
    c, b, limbs, n = ciphertexts.shape
    _check_params(op, limbs, rns_params)
    out = torch.empty((c, limbs, n), dtype=torch.int64, device=ciphertexts.device)
    nat.call("fh_ckks_plaintext_inner_product", out.data_ptr(), out.stride(0), out.stride(1), ciphertexts.data_ptr(),
             ciphertexts.stride(0), ciphertexts.stride(1), ciphertexts.stride(2), plaintexts.data_ptr(),
             plaintexts.stride(0), plaintexts.stride(1), c, b, limbs, n, nat.t2(rns_params),
             nat.stream_ptr(ciphertexts.device))
    return out
