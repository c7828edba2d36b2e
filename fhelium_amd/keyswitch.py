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
from fhelium_amd import rns
from fhelium_amd.rns import F64_DIGIT_LIMIT
from fhelium_amd.runtime import RnsRuntime


# Primes below this bound use the binary64 kernels (matches the NTT row records).
SMALL_PRIME_LIMIT = rns.BINARY64_PRIME_LIMIT
# The integer rows' lazy arithmetic (forward values below 4q, exact product
# sums with one REDC) is sized for primes below this.
INTEGER_PRIME_LIMIT = 1 << 60


def _check_operand(name: str, t: torch.Tensor, batch: int, rows: int, n: int, device: torch.device) -> None:
    """The kernels read exactly [batch, rows, n] int64 words on ``device``
    through this operand's strides: anything else is rejected before a
    launch."""

    if t.dtype != torch.int64:
        raise ValueError(f"{name} must be int64 residues, got {t.dtype}")
    if t.device != device:
        raise ValueError(f"{name} is on {t.device}, the component on {device}")
    if t.dim() < 2 or tuple(t.shape[-2:]) != (rows, n) or t.numel() != batch * rows * n:
        raise ValueError(f"{name} must hold [{batch}, {rows}, {n}] residues, got {tuple(t.shape)}")


class AccumulatorSeeds:
    """d0, d1 of an HMult front left in the NTT domain for the key switch.

    ``data`` is [2, *batch, Q_level, N] int64 words in a private working
    format, not residues: rows whose prime is below 2^42 hold the binary64
    bit patterns of a b R^2 (signed, not reduced), the other rows a b R below
    2q; within each 256-coefficient tile, coefficient 8t + k sits at position
    t + 32k.  Only ``HybridKeySwitcher.key_switch_fused(seeds=...)`` reads
    it."""

    __slots__ = ("data", "level")

    def __init__(self, data: torch.Tensor, level: int) -> None:
        self.data = data
        self.level = level


class HybridKeySwitcher:
    def __init__(self, runtime: RnsRuntime) -> None:
        self.rt = runtime
        self.layout = runtime.layout
        self.config = runtime.config
        too_large = [q for q in self.config.moduli if q >= INTEGER_PRIME_LIMIT]
        if too_large:
            raise ValueError(f"fused key switch needs every prime below 2^60, got {too_large}")
        self._digit_tables: dict[int, torch.Tensor] = {}
        # Optional list receiving (start_event, end_event, algorithmic_bytes)
        # per fused key switch, recorded on the operand's stream.
        self.recorder: list | None = None

    # -- fused path (fh_keyswitch_fused) -----------------------------------
    def digit_table(self, level: int) -> torch.Tensor:
        """[active digits, 16] int64 descriptors for fh_keyswitch_fused."""

        hit = self._digit_tables.get(level)
        if hit is not None:
            return hit
        words = []
        for spec in self.layout.digit_specs(level):
            t = self.rt.digit_tables[spec.prime_ids]
            words.append(self._digit_words(t, spec.component_row_ids[0], spec.key_digit_index))
        table = torch.tensor(words, dtype=torch.int64, device=self.rt.device)
        self._digit_tables[level] = table
        return table

    def _digit_words(self, t, row_begin: int, key_digit: int,
                     moddown_table: torch.Tensor | None = None) -> list[int]:
        d = len(t.prime_ids)
        multi = d > 1
        # binary64 basis extension needs every digit below 2^45
        f64_ok = int(all(self.config.moduli[i] < F64_DIGIT_LIMIT for i in t.prime_ids))
        # words 14, 15: the P digit's ModDown records, or a Q digit's
        # binary64 Garner record (f64_ok, D > 1)
        if moddown_table is not None:
            f64_words = [moddown_table.data_ptr(), moddown_table.stride(0)]
        elif t.garner_f64.numel():
            f64_words = [t.garner_f64.data_ptr(), 0]
        else:
            f64_words = [0, 0]
        return [
            row_begin, d, key_digit,
            t.normalizers.data_ptr() if multi else 0,
            t.propagation.data_ptr() if multi else 0,
            t.modulus_lo.data_ptr(), t.modulus_hi.data_ptr(), t.neg_inv_lo.data_ptr(),
            t.neg_inv_hi.data_ptr(),
            t.basis_extension.data_ptr() if multi else 0,
            t.basis_extension.stride(0) if multi else 0,
            t.basis_extension_std.data_ptr() if multi else 0,
            f64_ok,
            t.basis_extension_shoup.data_ptr() if multi else 0,
            *f64_words,
        ]

    def p_digit_table(self) -> torch.Tensor:
        """[16] descriptor of the P primes as one mixed-radix digit."""

        hit = self.__dict__.get("_p_digit_table")
        if hit is None:
            words = self._digit_words(self.rt.p_digit, 0, 0, self.rt.moddown_binary64)
            hit = torch.tensor(words, dtype=torch.int64, device=self.rt.device)
            self._p_digit_table = hit
        return hit

    def _address_table(self, tensors: list[torch.Tensor], dev: torch.device) -> torch.Tensor:
        """Device table of tensor addresses, cached: building it is a host-to-
        device copy, which from pageable memory would stall the host."""

        cache = self.__dict__.setdefault("_bases_cache", {})
        tag = tuple(k.data_ptr() for k in tensors)
        hit = cache.get(tag)
        if hit is None or hit[1] != [id(k) for k in tensors]:
            if len(cache) > 4096:
                cache.clear()
            table = torch.tensor(tag, dtype=torch.int64).pin_memory().to(dev, non_blocking=True)
            # keep the tensors alive alongside their addresses
            hit = (table, [id(k) for k in tensors], list(tensors))
            cache[tag] = hit
        return hit[0]

    def key_switch_fused(self, component: torch.Tensor, key: KeySwitchKey | list[KeySwitchKey],
                         level: int, base0: torch.Tensor | None = None,
                         base1: torch.Tensor | None = None,
                         galois_elements: list[int] | None = None, rescale: bool = False,
                         seeds: AccumulatorSeeds | None = None,
                         static_workspace: bool = False) -> torch.Tensor:
        """Whole key switch in eight launches; returns [2, *batch, Q_level, N]
        holding (base0 + corr0, base1 + corr1) canonically added, or the bare
        corrections where a base is None.  With ``rescale`` the result is
        also rescaled in the same ModDown launch (rescale_to_next_level's
        nearest rounding) and has Q_level - 1 rows.

        ``seeds`` ([2, *batch, Q_level, N], the NTT-domain d0, d1 that
        ``tensor_product_coefficient(..., ntt_d01=True)`` leaves) is added to
        the results exactly as base0 = d0, base1 = d1 would be, through the
        accumulators (floor((X + P d) / P) = floor(X / P) + d).

        With ``galois_elements`` (one per batch entry) the component and base0
        are read through entry b's coefficient automorphism, i.e. this is the
        key switch of a rotation applied to the unrotated ciphertext.

        With ``static_workspace`` the launches take their workspaces from the
        library's per-device static scratch (fh_keyswitch_fused_static: 64 MiB,
        one call at a time per device, every launch on the current stream)
        instead of three PyTorch allocations; only the output is allocated
        here."""

        from fhelium_amd import _native as nat

        rt = self.rt
        cfg = self.config
        n = cfg.N
        lq = cfg.num_q_primes - rt.row_start(level)
        lqp = lq + cfg.num_p_primes
        if component.dim() < 2 or tuple(component.shape[-2:]) != (lq, n):
            raise ValueError(f"key switch component must be [*batch, {lq}, {n}] at level {level}")
        batch_shape = component.shape[:-2]
        comp = component.reshape(-1, lq, n)
        b = comp.size(0)
        dev = component.device
        _check_operand("component", component, b, lq, n, dev)
        for name, t in (("base0", base0), ("base1", base1)):
            if t is not None:
                _check_operand(name, t, b, lq, n, dev)
        table = self.digit_table(level)
        dnum = table.size(0)
        if static_workspace:
            ws_args = ()
        else:
            ws_digits = torch.empty((b, lq, n), dtype=torch.int64, device=dev)
            ws_inter = torch.empty((b, dnum, lqp, n), dtype=torch.int64, device=dev)
            ws_acc = torch.empty((2, b, lqp, n), dtype=torch.int64, device=dev)
            ws_args = (ws_digits.data_ptr(), ws_inter.data_ptr(), ws_acc.data_ptr())
        out_rows = lq - 1 if rescale else lq
        if rescale and lq < 2:
            raise ValueError("a fused rescale needs at least two Q rows")
        out = torch.empty((2, b, out_rows, n), dtype=torch.int64, device=dev)
        if isinstance(key, (list, tuple)):
            if len(key) != b:
                raise ValueError("one key per batch entry is required")
            keys = [k.data for k in key]
            ref = keys[0]
            for k in keys:
                if k.shape != ref.shape or k.stride() != ref.stride():
                    raise ValueError("per-entry keys must share shape and strides")
            bases = self._address_table(keys, dev)
            key_ptr, bases_ptr = None, bases.data_ptr()
        else:
            ref = key.data
            key_ptr, bases_ptr = ref.data_ptr(), None
        if ref.dim() != 4 or ref.size(1) != 2 or ref.size(2) != cfg.total_num_primes or ref.size(3) != n:
            raise ValueError("key must be [digit, 2, QP, N]")
        gal_idx_ptr = gal_sgn_ptr = None
        if galois_elements is not None:
            if len(galois_elements) != b:
                raise ValueError("one Galois element per batch entry is required")
            tables = [rt.coefficient_galois(int(g)) for g in galois_elements]
            gal_idx_ptr = self._address_table([t[0] for t in tables], dev).data_ptr()
            gal_sgn_ptr = self._address_table([t[1] for t in tables], dev).data_ptr()
        seed_ptrs = (None, None, None)
        if seeds is not None:
            if not isinstance(seeds, AccumulatorSeeds):
                raise TypeError("seeds must be the AccumulatorSeeds that tensor_product_seeded returns")
            if seeds.level != level:
                raise ValueError(f"seeds were made at level {seeds.level}, not {level}")
            sv = seeds.data.reshape(2, -1, lq, n)
            _check_operand("seeds", sv[0], b, lq, n, dev)
            if sv.size(1) != b or not sv.is_contiguous():
                raise ValueError("seeds must be a contiguous [2, *batch, Q_level, N] tensor")
            seed_ptrs = (sv[0].data_ptr(), sv[1].data_ptr(), rt.seed_scale.data_ptr())
        none = nat.fh_view3(None, 0, 0, 0)
        view = lambda t: nat.v3(t.reshape(-1, lq, n)) if t is not None else none  # noqa: E731
        nt = rt.ntt_tables
        rec = self.recorder
        if rec is not None:
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
        start = rt.row_start(level)
        small = [q < SMALL_PRIME_LIMIT for q in cfg.moduli[start: start + lqp]]
        small_rows = sum(small)
        if any(small[small_rows:]):
            raise ValueError("fused key switch needs the binary64 rows before the integer rows")
        entry = "fh_keyswitch_fused_static" if static_workspace else "fh_keyswitch_fused"
        nat.call(entry, nat.v3(comp), b, lq, lqp, start, cfg.logN, small_rows,
                 table.data_ptr(), dnum, self.p_digit_table().data_ptr(), rt.p_inverse_q.data_ptr(),
                 nat.t2(rt.rns_params), nt.rows.data_ptr(), nt.fwd.data_ptr(), nt.fwd_shoup.data_ptr(),
                 nt.inv.data_ptr(), nt.inv_shoup.data_ptr(), key_ptr, bases_ptr,
                 ref.stride(0), ref.stride(1), ref.stride(2), gal_idx_ptr, gal_sgn_ptr,
                 *ws_args, view(base0), view(base1),
                 nat.v3(out[0]), nat.v3(out[1]), rt.ntt._mode, *self._rescale_args(level, rescale),
                 *seed_ptrs, nat.stream_ptr(dev))
        if rec is not None:
            ev1.record()
            addends = (base0 is not None) + (base1 is not None) + (2 if seeds is not None else 0)
            rec.append((ev0, ev1, self.algorithmic_bytes(b, lq, lqp, dnum, isinstance(key, (list, tuple)),
                                                         addends)))
        # Workspaces and the address tables may be released now: PyTorch's
        # caching allocator only hands their blocks to later work on this
        # same stream, which is ordered after these launches.
        return out.reshape(2, *batch_shape, out_rows, n)

    def _galois_pairs(self, elements: list[int], dev: torch.device) -> torch.Tensor:
        """Device [K, 2] table of (g, g^{-1} mod 2N) per entry, cached."""

        cache = self.__dict__.setdefault("_galois_pair_cache", {})
        tag = tuple(int(g) for g in elements)
        hit = cache.get(tag)
        if hit is None:
            if len(cache) > 256:
                cache.clear()
            two_n = 2 * self.config.N
            pairs = [(g % two_n, pow(g, -1, two_n)) for g in tag]
            hit = torch.tensor(pairs, dtype=torch.int64).pin_memory().to(dev, non_blocking=True)
            cache[tag] = hit
        return hit

    def key_switch_hoisted(self, component: torch.Tensor, base0: torch.Tensor,
                           keys: list[KeySwitchKey], level: int) -> torch.Tensor:
        """Hoisted rotations of one ciphertext (c0 = ``base0``, c1 =
        ``component``, each [Q_level, N] or [1, Q_level, N], canonical
        standard) by several rotation keys, in nine launches
        (fh_keyswitch_hoisted): c1 is decomposed and extended once, and every
        key's MAC reads that one intermediate through the NTT-domain Galois
        map of the key's element, as prepare_rotation_digits +
        _apply_prepared_rotation do (hybrid_keyswitch.py:373-504), with
        sigma_g(c0) added canonically (ckks_engine.py:3767-3820).  Returns
        [2, K, Q_level, N]: entry e is the rotation by keys[e]."""

        from fhelium_amd import _native as nat

        rt = self.rt
        cfg = self.config
        n = cfg.N
        lq = cfg.num_q_primes - rt.row_start(level)
        lqp = lq + cfg.num_p_primes
        dev = component.device
        _check_operand("component", component, 1, lq, n, dev)
        _check_operand("base0", base0, 1, lq, n, dev)
        if not keys:
            raise ValueError("at least one rotation key is required")
        k = len(keys)
        data = [key.data for key in keys]
        ref = data[0]
        for t in data:
            if t.shape != ref.shape or t.stride() != ref.stride():
                raise ValueError("rotation keys must share shape and strides")
        if ref.dim() != 4 or ref.size(1) != 2 or ref.size(2) != cfg.total_num_primes or ref.size(3) != n:
            raise ValueError("key must be [digit, 2, QP, N]")
        elements = [int(key.galois_element) for key in keys]
        comp = component.reshape(1, lq, n)
        c0 = base0.reshape(1, lq, n)
        table = self.digit_table(level)
        dnum = table.size(0)
        ws_digits = torch.empty((k, lq, n), dtype=torch.int64, device=dev)
        ws_inter = torch.empty((1, dnum, lqp, n), dtype=torch.int64, device=dev)
        ws_acc = torch.empty((2, k, lqp, n), dtype=torch.int64, device=dev)
        out = torch.empty((2, k, lq, n), dtype=torch.int64, device=dev)
        bases = self._address_table(data, dev)
        tables = [rt.coefficient_galois(g) for g in elements]
        gal_idx = self._address_table([t[0] for t in tables], dev)
        gal_sgn = self._address_table([t[1] for t in tables], dev)
        pairs = self._galois_pairs(elements, dev)
        start = rt.row_start(level)
        small = [q < SMALL_PRIME_LIMIT for q in cfg.moduli[start: start + lqp]]
        small_rows = sum(small)
        if any(small[small_rows:]):
            raise ValueError("fused key switch needs the binary64 rows before the integer rows")
        nt = rt.ntt_tables
        rec = self.recorder
        if rec is not None:
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
        nat.call("fh_keyswitch_hoisted", nat.v3(comp), k, lq, lqp, start, cfg.logN, small_rows,
                 table.data_ptr(), dnum, self.p_digit_table().data_ptr(), rt.p_inverse_q.data_ptr(),
                 nat.t2(rt.rns_params), nt.rows.data_ptr(), nt.fwd.data_ptr(), nt.fwd_shoup.data_ptr(),
                 nt.inv.data_ptr(), nt.inv_shoup.data_ptr(), bases.data_ptr(), ref.stride(0), ref.stride(1),
                 ref.stride(2), pairs.data_ptr(), gal_idx.data_ptr(), gal_sgn.data_ptr(), ws_digits.data_ptr(),
                 ws_inter.data_ptr(), ws_acc.data_ptr(), nat.v3(c0), nat.v3(out[0]), nat.v3(out[1]),
                 rt.ntt._mode, nat.stream_ptr(dev))
        if rec is not None:
            ev1.record()
            limb = n * 8
            rec.append((ev0, ev1, limb * (2 * lq + k * 2 * lq + k * dnum * 2 * lqp)))
        return out

    def _rescale_args(self, level: int, rescale: bool) -> tuple:
        """(q_level^{-1} mod q_i table, Shoup companions, floor(q_level / 2))
        for the rows above ``level``, cached on the device; nulls without a
        rescale."""

        if not rescale:
            return None, None, 0
        cache = self.__dict__.setdefault("_rescale_tables", {})
        hit = cache.get(level)
        if hit is None:
            import numpy as np

            cfg = self.config
            start = self.rt.row_start(level)
            q0 = cfg.moduli[start]
            qs = cfg.moduli[start + 1: cfg.num_q_primes]
            inv = [pow(q0, -1, q) for q in qs]
            shoup = [(v << 64) // q for v, q in zip(inv, qs)]
            inv_t = torch.tensor(inv, dtype=torch.int64).to(self.rt.device)
            shoup_t = torch.from_numpy(np.array(shoup, dtype=np.uint64).view(np.int64)).to(self.rt.device)
            hit = (inv_t, shoup_t, q0 // 2)
            cache[level] = hit
        return hit[0].data_ptr(), hit[1].data_ptr(), hit[2]

    def tensor_product_coefficient(self, a: torch.Tensor, b: torch.Tensor, level: int) -> torch.Tensor:
        """[3, *batch, Q_l, N] coefficient-domain d0, d1, d2 of two
        coefficient-domain ciphertexts a, b ([2, *batch, Q_l, N], canonical
        standard): NTT of both, tensor product and inverse NTT in one fused
        pass (fh_ckks_hmult_tensor_coefficient, fast arithmetic)."""

        return self._tensor_product(a, b, level, False)

    def tensor_product_seeded(self, a: torch.Tensor, b: torch.Tensor,
                              level: int) -> tuple[AccumulatorSeeds, torch.Tensor]:
        """The same product with only d2 ([*batch, Q_l, N], coefficient
        domain) returned as residues; d0, d1 stay in the NTT domain as the
        ``AccumulatorSeeds`` that ``key_switch_fused(seeds=...)`` consumes."""

        out = self._tensor_product(a, b, level, True)
        return AccumulatorSeeds(out[:2], level), out[2]

    def _tensor_product(self, a: torch.Tensor, b: torch.Tensor, level: int, ntt_d01: bool) -> torch.Tensor:

        from fhelium_amd import _native as nat

        rt = self.rt
        cfg = self.config
        n = cfg.N
        start = rt.row_start(level)
        lq = cfg.num_q_primes - start
        if a.shape != b.shape or a.dim() < 3 or a.size(0) != 2 or a.shape[-2:] != (lq, n):
            raise ValueError("tensor product operands must be [2, *batch, Q_level, N] of equal shape")
        for name, t in (("a", a), ("b", b)):
            if t.dtype != torch.int64:
                raise ValueError(f"tensor product operand {name} must be int64 residues, got {t.dtype}")
            if t.device != a.device:
                raise ValueError("tensor product operands must share one device")
        batch_shape = a.shape[1:-2]
        av = a.reshape(2, -1, lq, n)
        bv = b.reshape(2, -1, lq, n)
        for t in (av, bv):
            if t.stride(-1) != 1:
                raise ValueError("tensor product operands need unit coefficient stride")
        batch = av.size(1)
        small = [q < SMALL_PRIME_LIMIT for q in cfg.moduli[start: start + lq]]
        small_rows = sum(small)
        if any(small[small_rows:]):
            raise ValueError("fused tensor product needs the binary64 rows before the integer rows")
        ws = torch.empty((4, batch, lq, n), dtype=torch.int64, device=a.device)
        out = torch.empty((3, batch, lq, n), dtype=torch.int64, device=a.device)
        nt = rt.ntt_tables
        nat.call("fh_ckks_hmult_tensor_coefficient", av.data_ptr(), bv.data_ptr(), av.stride(0), av.stride(1),
                 av.stride(2), bv.stride(0), bv.stride(1), bv.stride(2), batch, lq, cfg.logN, start, small_rows,
                 nat.t2(rt.rns_params), nt.rows.data_ptr(), nt.fwd.data_ptr(), nt.fwd_shoup.data_ptr(),
                 nt.inv.data_ptr(), nt.inv_shoup.data_ptr(), ws.data_ptr(), out.data_ptr(),
                 1 if ntt_d01 else 0, nat.stream_ptr(a.device))
        return out.reshape(3, *batch_shape, lq, n)

    def algorithmic_bytes(self, batch: int, lq: int, lqp: int, dnum: int, per_entry_keys: bool,
                          bases: int) -> int:
        """Compulsory HBM bytes of one fused key switch: the component and
        the bases read once, both outputs written once, and every key digit
        (2 x QP rows) read once per distinct key."""

        limb = self.config.N * 8
        keys = batch if per_entry_keys else 1
        return limb * (batch * lq * (1 + bases + 2) + keys * dnum * 2 * lqp)

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
