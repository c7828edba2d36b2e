"""The reference's operator schemas, served by the gfx950 kernels.

The reference's Python layers reach native code only through
``torch.ops.fhelium_{ntt,rns,ckks}_ops.*`` (schemas in
csrc/ops/ntt/forward_ntt.cpp:25-58, csrc/ops/ntt/inverse_ntt.cpp:23-72,
csrc/ops/rns/rns_arithmetic.cpp:30-72,
csrc/ops/rns/rns_canonical_arithmetic.cpp:15-23, csrc/ops/ckks/ckks.cpp:29-90).
Importing this module defines those schemas with the same text and
registers implementations for the ``CUDA`` dispatch key, which is the key a
ROCm device tensor carries in PyTorch.  No CPU kernel is registered: a call
with host tensors fails in the dispatcher instead of running elsewhere.

Every implementation validates its operands and calls the C ABI of
libfhelium_amd.so (via ``fhelium_amd.kernels``) on PyTorch's current stream.

The compact NTT ops take the reference's table format (compact Montgomery
twiddles ``[limb, N]`` aligned with the operand's limbs, plus ``rns_params``);
the backend-specific records those become are derived on the device once per
(table, parameter) pair and cached.  ``grouped_stage_count`` is a launch
tuning knob of the reference's CUDA kernel (validated as 2, 3 or 4, as there)
and has no effect here.  The power-of-two-radix family takes the reference's
outer-twist and radix-root tables; the 2N-th root each row's tables are
built from is read from them, checked against the radix-root powers, and
expanded on the device into the compact form, so every NTT family runs the
same two-pass transform.  Forward outputs equal the reference's modulo q
(its radix-4/8/16 kernels pick their own lazy representatives); standard
inverse outputs are bit-identical.
"""

from __future__ import annotations

import threading

import torch

from fhelium_amd import _native as nat
from fhelium_amd import kernels as K
from fhelium_amd import ntt as gntt

_NTT_SCHEMAS = [
    "forward_ntt_montgomery_indexed_(Tensor(a!) montgomery_residues, Tensor even_indices, "
    "Tensor odd_indices, Tensor forward_twiddles, Tensor rns_params) -> ()",
    "forward_ntt_to_montgomery_indexed_(Tensor(a!) standard_residues, Tensor even_indices, "
    "Tensor odd_indices, Tensor forward_twiddles, Tensor rns_params) -> ()",
    "forward_ntt_to_montgomery_indexed(Tensor standard_residues, Tensor even_indices, "
    "Tensor odd_indices, Tensor forward_twiddles, Tensor rns_params) -> Tensor",
    "forward_ntt_montgomery_compact_grouped_smem_(Tensor(a!) montgomery_residues, "
    "Tensor forward_twiddles, Tensor rns_params, int grouped_stage_count) -> ()",
    "forward_ntt_to_montgomery_compact_grouped_smem_(Tensor(a!) standard_residues, "
    "Tensor forward_twiddles, Tensor rns_params, int grouped_stage_count) -> ()",
    "forward_ntt_to_montgomery_compact_grouped_smem(Tensor standard_residues, "
    "Tensor forward_twiddles, Tensor rns_params, int grouped_stage_count) -> Tensor",
    "forward_ntt_montgomery_compact_keyswitch_accumulate_(Tensor(a!) coefficient_digit_qp, "
    "Tensor forward_twiddles, Tensor rns_params, Tensor key_digit_qp, Tensor(b!) accumulator0_qp, "
    "Tensor(c!) accumulator1_qp, int key_row_start) -> ()",
    "inverse_ntt_montgomery_indexed_(Tensor(a!) montgomery_residues, Tensor even_indices, "
    "Tensor odd_indices, Tensor inverse_twiddles, Tensor rns_params) -> ()",
    "inverse_ntt_to_standard_lazy_indexed_(Tensor(a!) montgomery_residues, Tensor even_indices, "
    "Tensor odd_indices, Tensor inverse_twiddles, Tensor rns_params) -> ()",
    "inverse_ntt_to_standard_indexed_(Tensor(a!) montgomery_residues, Tensor even_indices, "
    "Tensor odd_indices, Tensor inverse_twiddles, Tensor rns_params) -> ()",
    "inverse_ntt_to_centered_indexed_(Tensor(a!) montgomery_residues, Tensor even_indices, "
    "Tensor odd_indices, Tensor inverse_twiddles, Tensor rns_params) -> ()",
    "inverse_ntt_montgomery_compact_grouped_smem_(Tensor(a!) montgomery_residues, "
    "Tensor inverse_twiddles, Tensor rns_params, int grouped_stage_count) -> ()",
    "inverse_ntt_to_standard_lazy_compact_grouped_smem_(Tensor(a!) montgomery_residues, "
    "Tensor inverse_twiddles, Tensor rns_params, int grouped_stage_count) -> ()",
    "inverse_ntt_to_standard_compact_grouped_smem_(Tensor(a!) montgomery_residues, "
    "Tensor inverse_twiddles, Tensor rns_params, int grouped_stage_count) -> ()",
    "inverse_ntt_to_centered_compact_grouped_smem_(Tensor(a!) montgomery_residues, "
    "Tensor inverse_twiddles, Tensor rns_params, int grouped_stage_count) -> ()",
    "forward_ntt_montgomery_power_of_two_radix_compact_(Tensor(a!) montgomery_residues, "
    "Tensor outer_twiddles, Tensor radix_root_powers, Tensor rns_params) -> ()",
    "forward_ntt_to_montgomery_power_of_two_radix_compact_(Tensor(a!) standard_residues, "
    "Tensor outer_twiddles, Tensor radix_root_powers, Tensor rns_params) -> ()",
    "forward_ntt_to_montgomery_power_of_two_radix_compact(Tensor standard_residues, "
    "Tensor outer_twiddles, Tensor radix_root_powers, Tensor rns_params) -> Tensor",
    "inverse_ntt_montgomery_power_of_two_radix_compact_(Tensor(a!) montgomery_residues, "
    "Tensor outer_twiddles, Tensor radix_root_powers, Tensor rns_params) -> ()",
    "inverse_ntt_to_standard_lazy_power_of_two_radix_compact_(Tensor(a!) montgomery_residues, "
    "Tensor outer_twiddles, Tensor radix_root_powers, Tensor rns_params) -> ()",
    "inverse_ntt_to_standard_power_of_two_radix_compact_(Tensor(a!) montgomery_residues, "
    "Tensor outer_twiddles, Tensor radix_root_powers, Tensor rns_params) -> ()",
    "inverse_ntt_to_centered_power_of_two_radix_compact_(Tensor(a!) montgomery_residues, "
    "Tensor outer_twiddles, Tensor radix_root_powers, Tensor rns_params) -> ()",
]

_RNS_SCHEMAS = [
    "montgomery_mul(Tensor lhs, Tensor rhs, Tensor rns_params) -> Tensor",
    "montgomery_mul_cyclic_compressed(Tensor lhs, Tensor compressed_rhs, Tensor rns_params) -> Tensor",
    "montgomery_mul_contiguous_compressed(Tensor lhs, Tensor compressed_rhs, Tensor rns_params) -> Tensor",
    "montgomery_mul_row_scalars_(Tensor(a!) residues, Tensor row_scalars, Tensor rns_params) -> ()",
    "to_montgomery_(Tensor(a!) standard_residues, Tensor rns_params) -> ()",
    "from_montgomery_(Tensor(a!) montgomery_residues, Tensor rns_params) -> ()",
    "add_lazy(Tensor lhs, Tensor rhs, Tensor rns_params) -> Tensor",
    "add_lazy_with_twice_modulus(Tensor lhs, Tensor rhs, Tensor twice_modulus) -> Tensor",
    "sub_lazy(Tensor lhs, Tensor rhs, Tensor rns_params) -> Tensor",
    "canonicalize_residues_(Tensor(a!) lazy_residues, Tensor rns_params) -> ()",
    "center_residues_(Tensor(a!) canonical_residues, Tensor rns_params) -> ()",
    "shift_residues_positive_(Tensor(a!) centered_residues, Tensor rns_params) -> ()",
    "lift_centered_coefficients(Tensor centered_coefficients, Tensor twice_modulus) -> Tensor",
    "mixed_radix_decompose(Tensor source_residues, Tensor mixed_radix_normalizers, "
    "Tensor mixed_radix_propagation_coefficients, Tensor modulus_lo, Tensor modulus_hi, "
    "Tensor neg_inv_modulus_lo, Tensor neg_inv_modulus_hi) -> Tensor",
    "mixed_radix_basis_extend_to_montgomery(Tensor mixed_radix_components, "
    "Tensor basis_extension_coefficients, Tensor rns_params, int destination_row_count) -> Tensor",
    "add_canonical(Tensor lhs, Tensor rhs, Tensor rns_params) -> Tensor",
    "add_canonical_(Tensor(a!) lhs, Tensor rhs, Tensor rns_params) -> ()",
    "sub_canonical(Tensor lhs, Tensor rhs, Tensor rns_params) -> Tensor",
    "sub_canonical_(Tensor(a!) lhs, Tensor rhs, Tensor rns_params) -> ()",
    "montgomery_mul_row_scalars_canonical(Tensor residues, Tensor row_scalars, Tensor rns_params) -> Tensor",
]

_CKKS_SCHEMAS = [
    "add_prepared_plaintext_component(Tensor ciphertext_component, Tensor prepared_plaintext, "
    "Tensor rns_params) -> Tensor",
    "add_prepared_plaintext_component_(Tensor(a!) ciphertext_component, Tensor prepared_plaintext, "
    "Tensor rns_params) -> ()",
    "add_cyclic_compressed_plaintext_component(Tensor ciphertext_component, "
    "Tensor compressed_plaintext, Tensor rns_params) -> Tensor",
    "add_cyclic_compressed_plaintext_component_(Tensor(a!) ciphertext_component, "
    "Tensor compressed_plaintext, Tensor rns_params) -> ()",
    "add_contiguous_compressed_plaintext_component(Tensor ciphertext_component, "
    "Tensor compressed_plaintext, Tensor rns_params) -> Tensor",
    "add_contiguous_compressed_plaintext_component_(Tensor(a!) ciphertext_component, "
    "Tensor compressed_plaintext, Tensor rns_params) -> ()",
    "add_strided_plaintext_component(Tensor ciphertext_component, Tensor strided_plaintext, "
    "Tensor implicit_plaintext, Tensor rns_params) -> Tensor",
    "add_strided_plaintext_component_(Tensor(a!) ciphertext_component, Tensor strided_plaintext, "
    "Tensor implicit_plaintext, Tensor rns_params) -> ()",
    "rescale_drop_leading_prime_nearest(Tensor remaining_residues, Tensor drop_prime_inverse_mont, "
    "Tensor dropped_residue, Tensor rns_params, int half_drop_prime) -> Tensor",
    "rescale_drop_leading_prime_nearest_(Tensor(a!) remaining_residues, Tensor drop_prime_inverse_mont, "
    "Tensor dropped_residue, Tensor rns_params, int half_drop_prime) -> ()",
    "rescale_drop_leading_prime_truncate(Tensor remaining_residues, Tensor drop_prime_inverse_mont, "
    "Tensor dropped_residue, Tensor rns_params) -> Tensor",
    "rescale_drop_leading_prime_truncate_(Tensor(a!) remaining_residues, Tensor drop_prime_inverse_mont, "
    "Tensor dropped_residue, Tensor rns_params) -> ()",
    "apply_coefficient_galois_automorphism(Tensor residues, Tensor source_indices, Tensor source_sign, "
    "Tensor twice_modulus) -> Tensor",
    "apply_ntt_galois_automorphism(Tensor residues_ntt, Tensor source_indices) -> Tensor",
    "keyswitch_moddown_qp_to_q(Tensor q_residues, Tensor p_residues, "
    "Tensor moddown_p_drop_inverses_montgomery, Tensor rns_params) -> Tensor",
    "keyswitch_accumulate_digit_products_(Tensor(a!) accumulator0_qp, Tensor(b!) accumulator1_qp, "
    "Tensor extended_digit_ntt_qp, Tensor key_switch_key_digit, Tensor rns_params, "
    "int key_digit_row_start) -> ()",
]

SCHEMAS = {
    "fhelium_ntt_ops": _NTT_SCHEMAS,
    "fhelium_rns_ops": _RNS_SCHEMAS,
    "fhelium_ckks_ops": _CKKS_SCHEMAS,
}

# rns_params row of N^{-1} R mod q (rns_parameters.h:6-16)
_N_INVERSE_MONTGOMERY_ROW = 7

_lock = threading.Lock()
_libraries: list = []
_ntt_mode = gntt.MODE_EXACT


def set_ntt_mode(mode: str) -> None:
    """``"exact"`` (default): forward outputs bit-identical to the reference
    CPU backend's lazy values; ``"fast"``: canonical forward outputs (equal
    modulo q, as the reference's own CPU-vs-CUDA tests compare them)."""

    global _ntt_mode
    if mode not in ("exact", "fast"):
        raise ValueError("mode must be 'exact' or 'fast'")
    _ntt_mode = gntt.MODE_EXACT if mode == "exact" else gntt.MODE_FAST


# -- compact NTT tables --------------------------------------------------------
class _PreparedRows:
    """Backend records for one (compact Montgomery twiddles, rns_params) pair."""

    def __init__(self, twiddles: torch.Tensor, rns_params: torch.Tensor) -> None:
        rows, n = twiddles.shape
        dev = twiddles.device
        self.sources = (twiddles, rns_params)  # keeps the addresses in the key unique
        self.log_n = n.bit_length() - 1
        self.rows = torch.zeros((rows, gntt.NTT_ROW_WORDS), dtype=torch.int64, device=dev)
        self.w = torch.empty((rows, n), dtype=torch.int64, device=dev)
        self.wp = torch.empty((rows, n), dtype=torch.int64, device=dev)
        nat.call("fh_ntt_prepare_tables", rns_params.data_ptr(), rns_params.stride(0), rns_params.stride(1),
                 rows, self.log_n, twiddles.data_ptr(), twiddles.stride(0), self.rows.data_ptr(),
                 self.w.data_ptr(), self.wp.data_ptr(), nat.stream_ptr(dev))


_prepared: dict[tuple, _PreparedRows] = {}


def _tensor_key(t: torch.Tensor) -> tuple:
    return (t.data_ptr(), tuple(t.shape), tuple(t.stride()), t._version, t.device.index)


def _prepared_rows(op: str, twiddles: torch.Tensor, rns_params: torch.Tensor, limbs: int,
                   n: int) -> _PreparedRows:
    nat.require_device(twiddles, rns_params)
    if twiddles.dtype != torch.int64 or rns_params.dtype != torch.int64:
        raise RuntimeError(f"{op}: the gfx950 backend supports int64 tables")
    if twiddles.dim() != 2 or twiddles.size(0) != limbs or twiddles.size(1) != n:
        raise RuntimeError(f"{op}: twiddles must be [limb, N] matching the operand")
    if twiddles.stride(1) != 1:
        raise RuntimeError(f"{op}: twiddle rows must be contiguous")
    if rns_params.dim() != 2 or rns_params.size(0) < 8 or rns_params.size(1) != limbs:
        raise RuntimeError(f"{op}: rns_params must be [8, limb] matching the operand")
    if n < 2 or n & (n - 1):
        raise RuntimeError(f"{op}: ring dimension must be a power of two")
    key = (_tensor_key(twiddles), _tensor_key(rns_params))
    with _lock:
        hit = _prepared.get(key)
        if hit is None:
            if len(_prepared) > 256:
                _prepared.clear()
            hit = _PreparedRows(twiddles, rns_params)
            _prepared[key] = hit
    return hit


def _compact(op: str, residues: torch.Tensor, twiddles: torch.Tensor, rns_params: torch.Tensor,
             grouped_stage_count: int, forward: bool, form: int) -> None:
    if grouped_stage_count not in (2, 3, 4):
        raise RuntimeError("grouped_stage_count must be 2, 3, or 4")
    view = _residue_view(op, residues)
    prep = _prepared_rows(op, twiddles, rns_params, view.size(1), view.size(2))
    _run(view, prep, residues, forward, form)


def _residue_view(op: str, residues: torch.Tensor) -> torch.Tensor:
    nat.require_device(residues)
    if residues.dtype != torch.int64:
        raise RuntimeError(f"{op}: the gfx950 backend supports int64 residues")
    view = nat.batch_view(residues, "residues")
    if view.stride(2) != 1:
        raise RuntimeError(f"{op}: residues need a unit coefficient stride")
    return view


def _run(view: torch.Tensor, prep: "_PreparedRows", residues: torch.Tensor, forward: bool, form: int) -> None:
    fn = "fh_ntt_forward" if forward else "fh_ntt_inverse"
    nat.call(fn, view.data_ptr(), view.size(0), view.size(1), prep.log_n, view.stride(0), view.stride(1),
             prep.rows.data_ptr(), prep.w.data_ptr(), prep.wp.data_ptr(), 0, form, _ntt_mode,
             nat.stream_ptr(residues.device))


def forward_ntt_montgomery_compact_grouped_smem_(montgomery_residues, forward_twiddles, rns_params,
                                                 grouped_stage_count):
    _compact("forward_ntt_montgomery_compact_grouped_smem_", montgomery_residues, forward_twiddles,
             rns_params, grouped_stage_count, True, 0)


def forward_ntt_to_montgomery_compact_grouped_smem_(standard_residues, forward_twiddles, rns_params,
                                                    grouped_stage_count):
    _compact("forward_ntt_to_montgomery_compact_grouped_smem_", standard_residues, forward_twiddles,
             rns_params, grouped_stage_count, True, 1)


def forward_ntt_to_montgomery_compact_grouped_smem(standard_residues, forward_twiddles, rns_params,
                                                   grouped_stage_count):
    out = standard_residues.clone(memory_format=torch.contiguous_format)
    forward_ntt_to_montgomery_compact_grouped_smem_(out, forward_twiddles, rns_params, grouped_stage_count)
    return out


def _inverse_compact(form: int, name: str):
    def impl(montgomery_residues, inverse_twiddles, rns_params, grouped_stage_count):
        _compact(name, montgomery_residues, inverse_twiddles, rns_params, grouped_stage_count, False, form)

    impl.__name__ = name
    return impl


def forward_ntt_montgomery_compact_keyswitch_accumulate_(coefficient_digit_qp, forward_twiddles, rns_params,
                                                         key_digit_qp, accumulator0_qp, accumulator1_qp,
                                                         key_row_start):
    """Forward NTT of the digit in place, then the key MAC
    (forward_ntt_compact_keyswitch_fused.inc.cuh:9-172 semantics)."""

    forward_ntt_montgomery_compact_grouped_smem_(coefficient_digit_qp, forward_twiddles, rns_params, 3)
    K.keyswitch_accumulate_digit_products_(accumulator0_qp, accumulator1_qp, coefficient_digit_qp,
                                           key_digit_qp, rns_params, key_row_start)


# -- power-of-two-radix NTT ----------------------------------------------------
def _compact_from_root(root_mont: torch.Tensor, rns_params: torch.Tensor, log_n: int) -> torch.Tensor:
    """[limb, N] compact Montgomery twiddles root^{br(k)} R from each row's
    Montgomery root (a [limb] vector), by square-and-multiply on the device."""

    limbs = root_mont.size(0)
    n = 1 << log_n
    dev = root_mont.device
    one = K.montgomery_mul(rns_params[5].reshape(limbs, 1).contiguous(),
                           torch.ones((limbs, 1), dtype=torch.int64, device=dev), rns_params)  # R mod q
    k = torch.arange(n, device=dev)
    br = torch.zeros_like(k)
    for bit in range(log_n):
        br |= ((k >> bit) & 1) << (log_n - 1 - bit)
    acc = one.expand(limbs, n).contiguous()
    power = root_mont.reshape(limbs, 1).contiguous()
    for bit in range(log_n):
        take = ((br >> bit) & 1).bool()
        prod = K.montgomery_mul(acc, power.expand(limbs, n).contiguous(), rns_params)
        acc = torch.where(take, prod, acc)
        power = K.montgomery_mul(power, power, rns_params)
    return acc


_radix_prepared: dict[tuple, tuple] = {}


def _radix_rows(op: str, residues: torch.Tensor, outer: torch.Tensor, roots: torch.Tensor,
                rns_params: torch.Tensor, forward: bool) -> tuple[torch.Tensor, "_PreparedRows"]:
    """Validate as the reference does (csrc/ops/ntt/ntt_validation.h:123-157)
    and return the operand view and the prepared rows of these tables."""

    view = _residue_view(op, residues)
    nat.require_device(outer, roots, rns_params)
    limbs, n = view.size(1), view.size(2)
    if outer.dtype != torch.int64 or roots.dtype != torch.int64 or rns_params.dtype != torch.int64:
        raise RuntimeError(f"{op}: the gfx950 backend supports int64 tables")
    if outer.dim() != 2 or outer.size(0) != limbs or outer.size(1) != n - 1:
        raise RuntimeError("Power-of-two radix outer twiddles must have shape [prime, coefficient - 1]")
    radix = roots.size(1) if roots.dim() == 2 else -1
    if radix not in (4, 8, 16):
        raise RuntimeError("Power-of-two radix root width must select radix 4, 8, or 16")
    if roots.device != residues.device:
        raise RuntimeError("Power-of-two radix root powers must share the operand device")
    if roots.size(0) != limbs:
        raise RuntimeError("Power-of-two radix root powers must have shape [prime, radix]")
    if K._overlaps(residues, roots):
        raise RuntimeError("power-of-two radix root powers overlap the residues")
    log_n = n.bit_length() - 1
    bits = {4: 2, 8: 3, 16: 4}[radix]
    if n & (n - 1) or log_n % bits:
        raise RuntimeError(f"Strict fixed-radix execution requires logN divisible by {bits}; got logN={log_n}")
    if rns_params.dim() != 2 or rns_params.size(0) < 8 or rns_params.size(1) != limbs:
        raise RuntimeError(f"{op}: rns_params must be [8, limb] matching the operand")
    key = (_tensor_key(outer), _tensor_key(roots), _tensor_key(rns_params), forward)
    with _lock:
        hit = _radix_prepared.get(key)
    if hit is None:
        # the root is the first twist of the digit whose groups start at
        # stage logN - 1: position N / radix - 1 of a forward table, 0 of an
        # inverse table
        root = outer[:, n // radix - 1 if forward else 0].contiguous()
        compact = _compact_from_root(root, rns_params, log_n)
        # the table's radix-root powers must be root^(2N/radix) = compact[br(2N/radix)]
        probe = int(f"{2 * n // radix:0{log_n}b}"[::-1], 2)
        if not torch.equal(compact[:, probe], roots[:, 1]):
            raise RuntimeError(f"{op}: outer twiddles and radix root powers describe different roots")
        prep = _PreparedRows(compact, rns_params)
        hit = (compact, prep)
        with _lock:
            if len(_radix_prepared) > 64:
                _radix_prepared.clear()
            _radix_prepared[key] = hit
    return view, hit[1]


def _radix(name: str, forward: bool, form: int):
    def impl(residues, outer_twiddles, radix_root_powers, rns_params):
        view, prep = _radix_rows(name, residues, outer_twiddles, radix_root_powers, rns_params, forward)
        _run(view, prep, residues, forward, form)

    impl.__name__ = name
    return impl


def forward_ntt_to_montgomery_power_of_two_radix_compact(standard_residues, outer_twiddles, radix_root_powers,
                                                         rns_params):
    out = standard_residues.clone(memory_format=torch.contiguous_format)
    _radix("forward_ntt_to_montgomery_power_of_two_radix_compact", True, 1)(out, outer_twiddles,
                                                                             radix_root_powers, rns_params)
    return out


# -- indexed NTT ---------------------------------------------------------------
def _indexed(op: str, residues, even_indices, odd_indices, twiddles, rns_params, forward: bool) -> None:
    nat.require_device(residues, even_indices, odd_indices, twiddles, rns_params)
    view = nat.batch_view(residues, "residues")
    if even_indices.dtype != torch.int32 or odd_indices.dtype != torch.int32:
        raise RuntimeError(f"{op}: index schedules must be int32")
    if even_indices.dim() != 2 or odd_indices.shape != even_indices.shape:
        raise RuntimeError(f"{op}: index schedules must be [stage, butterfly]")
    stages, butterflies = even_indices.shape
    if butterflies * 2 != view.size(2):
        raise RuntimeError(f"{op}: schedule does not match the ring dimension")
    if twiddles.dim() != 3 or twiddles.shape != (view.size(1), stages, butterflies):
        raise RuntimeError(f"{op}: twiddles must be [limb, stage, butterfly]")
    K._check_params(op, view.size(1), rns_params)
    if even_indices.stride(1) != odd_indices.stride(1):
        raise RuntimeError(f"{op}: even/odd schedules need equal butterfly strides")
    s = nat.stream_ptr(residues.device)
    for stage in range(stages):
        e, o, w = even_indices[stage], odd_indices[stage], twiddles[:, stage]
        nat.call("fh_ntt_indexed_stage", nat.v3(view), view.size(0), view.size(1), e.data_ptr(),
                 o.data_ptr(), e.stride(0), w.data_ptr(), w.stride(0), w.stride(1), nat.t2(rns_params),
                 butterflies, int(forward), s)


def forward_ntt_montgomery_indexed_(montgomery_residues, even_indices, odd_indices, forward_twiddles,
                                    rns_params):
    _indexed("forward_ntt_montgomery_indexed_", montgomery_residues, even_indices, odd_indices,
             forward_twiddles, rns_params, True)


def forward_ntt_to_montgomery_indexed_(standard_residues, even_indices, odd_indices, forward_twiddles,
                                       rns_params):
    K.to_montgomery_(standard_residues, rns_params)
    _indexed("forward_ntt_to_montgomery_indexed_", standard_residues, even_indices, odd_indices,
             forward_twiddles, rns_params, True)


def forward_ntt_to_montgomery_indexed(standard_residues, even_indices, odd_indices, forward_twiddles,
                                      rns_params):
    out = standard_residues.clone(memory_format=torch.contiguous_format)
    forward_ntt_to_montgomery_indexed_(out, even_indices, odd_indices, forward_twiddles, rns_params)
    return out


def _inverse_indexed(form: int, name: str):
    """Stages, then x N^{-1} R (REDC) and the output representation, as
    indexed_ntt_cpu.cpp:276-312 does."""

    def impl(montgomery_residues, even_indices, odd_indices, inverse_twiddles, rns_params):
        _indexed(name, montgomery_residues, even_indices, odd_indices, inverse_twiddles, rns_params, False)
        K.montgomery_mul_row_scalars_(montgomery_residues, rns_params[_N_INVERSE_MONTGOMERY_ROW], rns_params)
        if form != gntt.OUT_MONTGOMERY:
            K.from_montgomery_(montgomery_residues, rns_params)
            if form == gntt.OUT_CENTERED:
                K.center_residues_(montgomery_residues, rns_params)

    impl.__name__ = name
    return impl


# -- registration ----------------------------------------------------------------
def _implementations() -> dict[str, dict[str, object]]:
    ntt = {
        "forward_ntt_montgomery_indexed_": forward_ntt_montgomery_indexed_,
        "forward_ntt_to_montgomery_indexed_": forward_ntt_to_montgomery_indexed_,
        "forward_ntt_to_montgomery_indexed": forward_ntt_to_montgomery_indexed,
        "forward_ntt_montgomery_compact_grouped_smem_": forward_ntt_montgomery_compact_grouped_smem_,
        "forward_ntt_to_montgomery_compact_grouped_smem_": forward_ntt_to_montgomery_compact_grouped_smem_,
        "forward_ntt_to_montgomery_compact_grouped_smem": forward_ntt_to_montgomery_compact_grouped_smem,
        "forward_ntt_montgomery_compact_keyswitch_accumulate_":
            forward_ntt_montgomery_compact_keyswitch_accumulate_,
    }
    for form, suffix in ((gntt.OUT_MONTGOMERY, "montgomery"), (gntt.OUT_STANDARD_LAZY, "to_standard_lazy"),
                         (gntt.OUT_STANDARD, "to_standard"), (gntt.OUT_CENTERED, "to_centered")):
        ntt[f"inverse_ntt_{suffix}_indexed_"] = _inverse_indexed(form, f"inverse_ntt_{suffix}_indexed_")
        name = f"inverse_ntt_{suffix}_compact_grouped_smem_"
        ntt[name] = _inverse_compact(form, name)
        name = f"inverse_ntt_{suffix}_power_of_two_radix_compact_"
        ntt[name] = _radix(name, False, form)
    ntt["forward_ntt_montgomery_power_of_two_radix_compact_"] = _radix(
        "forward_ntt_montgomery_power_of_two_radix_compact_", True, 0)
    ntt["forward_ntt_to_montgomery_power_of_two_radix_compact_"] = _radix(
        "forward_ntt_to_montgomery_power_of_two_radix_compact_", True, 1)
    ntt["forward_ntt_to_montgomery_power_of_two_radix_compact"] = \
        forward_ntt_to_montgomery_power_of_two_radix_compact
    rns = {name.split("(")[0]: getattr(K, name.split("(")[0]) for name in _RNS_SCHEMAS}
    ckks = {name.split("(")[0]: getattr(K, name.split("(")[0]) for name in _CKKS_SCHEMAS}
    return {"fhelium_ntt_ops": ntt, "fhelium_rns_ops": rns, "fhelium_ckks_ops": ckks}


def _fake_impl(name: str, schema: str):
    """Shape-only implementation for FakeTensor tracing
    (fhelium/native/wrapper/fake_op.py): in-place ops return nothing,
    functional ops a fresh tensor of the output shape."""

    if schema.endswith("-> ()"):
        return lambda *args, **kwargs: None
    if name == "mixed_radix_basis_extend_to_montgomery":
        def extend(components, coefficients, rns_params, destination_row_count):
            return components.new_empty((*components.shape[:-2], destination_row_count, components.shape[-1]))
        return extend
    if name == "lift_centered_coefficients":
        def lift(centered, twice_modulus):
            return centered.new_empty((*centered.shape[:-1], twice_modulus.shape[0], centered.shape[-1]))
        return lift
    return lambda first, *args, **kwargs: torch.empty_like(first, memory_format=torch.contiguous_format)


def _existing_schema(ns: str, name: str) -> str | None:
    try:
        packet = getattr(getattr(torch.ops, ns), name)
    except (AttributeError, RuntimeError):
        return None
    return str(packet.default._schema)


def register() -> None:
    """Register the CUDA-key implementations (once).

    Schemas that are not defined yet are defined here.  Where a library
    already defined them (the reference's own extension with its CPU
    kernels), the text must match and only the CUDA kernels are added."""

    with _lock:
        if _libraries:
            return
        impls = _implementations()
        for ns, schemas in SCHEMAS.items():
            defs = torch.library.Library(ns, "FRAGMENT")
            kernels = torch.library.Library(ns, "IMPL")
            for schema in schemas:
                name = schema.split("(")[0]
                existing = _existing_schema(ns, name)
                if existing is None:
                    defs.define(schema)
                    # the defining library also describes output shapes
                    torch.library.register_fake(f"{ns}::{name}", _fake_impl(name, schema), lib=defs)
                elif existing != f"{ns}::{schema}":
                    raise RuntimeError(f"{ns}::{name} is already defined with a different schema: {existing}")
                kernels.impl(name, impls[ns][name], "CUDA")
            _libraries.extend((defs, kernels))


def registered_operators() -> list[str]:
    return [f"{ns}::{schema.split('(')[0]}" for ns, schemas in SCHEMAS.items() for schema in schemas]


register()
