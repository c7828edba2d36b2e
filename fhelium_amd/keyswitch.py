"""Hybrid-RNS key switching on gfx950.

Same algorithm and data flow as the reference's ``HybridKeySwitcher``
(fhelium/engine/hybrid_keyswitch.py:61-595): for every active digit of the
level, Garner-decompose the digit's Q rows (mixed radix), basis-extend onto
the active QP rows in Montgomery form, forward-NTT, multiply-accumulate with
the matching key digit into two QP accumulators; then inverse-NTT the
accumulators to canonical standard form and ModDown by P.  Accumulators are
lazy [0, 2q) exactly as in the reference (ckks_cpu.cpp:688-701), so the
corrections are bit-identical to the reference CPU backend.
"""

from __future__ import annotations

import torch

from fhelium_amd import kernels as K
from fhelium_amd.core import KeySwitchKey
from fhelium_amd.runtime import RnsRuntime


class HybridKeySwitcher:
    def __init__(self, runtime: RnsRuntime) -> None:
        self.rt = runtime
        self.layout = runtime.layout
        self.config = runtime.config

    def _mod_up(self, component: torch.Tensor, spec, level: int) -> torch.Tensor:
        """Digit rows -> mixed-radix digits -> [*, QP_level, N] Montgomery."""

        rt = self.rt
        rows = spec.component_row_ids
        source = component[..., rows[0]: rows[-1] + 1, :]
        t = rt.digit_tables[spec.prime_ids]
        if len(rows) > 1:
            digits = K.mixed_radix_decompose(source, t.normalizers, t.propagation, t.modulus_lo,
                                             t.modulus_hi, t.neg_inv_lo, t.neg_inv_hi)
        else:
            digits = source
        qp_params = rt.params(level, include_p=True)
        start = rt.row_start(level)
        return K.mixed_radix_basis_extend_to_montgomery(
            digits, t.basis_extension[:, start:], qp_params, qp_params.size(1))

    def apply_key_switch(self, component: torch.Tensor, key: KeySwitchKey | list[KeySwitchKey],
                         level: int) -> tuple[torch.Tensor, torch.Tensor]:
        """``component``: [*batch, Q_level, N] coefficient/standard (canonical).

        ``key`` is one key shared by the batch, or a list with one key per
        batch entry (batched rotations by per-ciphertext steps)."""

        rt = self.rt
        start = rt.row_start(level)
        qp_params = rt.params(level, include_p=True)
        per_entry = isinstance(key, (list, tuple))
        if per_entry:
            flat = component.reshape(-1, component.size(-2), component.size(-1))
            if len(key) != flat.size(0):
                raise ValueError("one key per batch entry is required")
            keys = [k.data for k in key]
            bases = torch.tensor([k.data_ptr() for k in keys], dtype=torch.int64, device=component.device)
        acc0 = acc1 = None
        for spec in self.layout.digit_specs(level):
            digit_qp = self._mod_up(component, spec, level)
            if acc0 is None:
                acc0 = torch.zeros_like(digit_qp)
                acc1 = torch.zeros_like(digit_qp)
            rt.ntt.forward_montgomery_(digit_qp, start)
            if per_entry:
                K.keyswitch_accumulate_multi_key_(acc0, acc1, digit_qp, keys, bases, spec.key_digit_index,
                                                  qp_params, start)
            else:
                K.keyswitch_accumulate_digit_products_(acc0, acc1, digit_qp,
                                                       key.digit(spec.key_digit_index), qp_params, start)
        return self.mod_down(acc0, acc1, level)

    def prepare_rotation_digits(self, component: torch.Tensor, level: int) -> torch.Tensor:
        """NTT-domain QP digits [digit, *batch, QP, N] for hoisted rotations
        (hybrid_keyswitch.py:373-417)."""

        rt = self.rt
        start = rt.row_start(level)
        specs = self.layout.digit_specs(level)
        out = None
        for spec in specs:
            digit_qp = self._mod_up(component, spec, level)
            rt.ntt.forward_montgomery_(digit_qp, start)
            if out is None:
                out = torch.empty((len(specs), *digit_qp.shape), dtype=digit_qp.dtype,
                                  device=digit_qp.device)
            out[spec.digit_index].copy_(digit_qp)
        return out

    def apply_prepared_rotation(self, prepared: torch.Tensor, key: KeySwitchKey, level: int,
                                galois_element: int) -> tuple[torch.Tensor, torch.Tensor]:
        """Hoisted rotation: NTT-domain gather of each prepared digit, MAC,
        ModDown (hybrid_keyswitch.py:419-504)."""

        rt = self.rt
        start = rt.row_start(level)
        qp_params = rt.params(level, include_p=True)
        src = rt.ntt_galois(galois_element)
        acc0 = torch.zeros_like(prepared[0])
        acc1 = torch.zeros_like(prepared[0])
        for spec, digit in zip(self.layout.digit_specs(level), prepared):
            rotated = K.apply_ntt_galois_automorphism(digit, src)
            K.keyswitch_accumulate_digit_products_(acc0, acc1, rotated, key.digit(spec.key_digit_index),
                                                   qp_params, start)
        return self.mod_down(acc0, acc1, level)

    def mod_down(self, acc0: torch.Tensor, acc1: torch.Tensor,
                 level: int) -> tuple[torch.Tensor, torch.Tensor]:
        """INTT to canonical standard, then floor(x / P) (hybrid_keyswitch.py:530-568)."""

        rt = self.rt
        start = rt.row_start(level)
        p = self.config.num_p_primes
        rt.ntt.inverse_to_standard_(acc0, start)
        rt.ntt.inverse_to_standard_(acc1, start)
        inv = rt.moddown_inverses[level]
        qp_params = rt.params(level, include_p=True)
        c0 = K.keyswitch_moddown_qp_to_q(acc0[..., :-p, :], acc0[..., -p:, :], inv, qp_params)
        c1 = K.keyswitch_moddown_qp_to_q(acc1[..., :-p, :], acc1[..., -p:, :], inv, qp_params)
        return c0, c1
