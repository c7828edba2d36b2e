"""``CkksEngine`` on gfx950: the reference's evaluation API for the hot path.

Method-for-method counterpart of fhelium/engine/ckks_engine.py for the five
north-star operations -- ``coefficient_domain_to_ntt_domain`` (:876-903),
``multiply`` (:1530-1606), ``relinearize`` (:1608-1675), ``rotate_by_step``
(:2183-2207) and ``rescale_to_next_level`` (:1131-1174) -- with the same
state checks, the same composition of primitives and the same output states,
plus the hoisted ``rotate_many_with_keys`` (:2298-2331).  Key generation,
encryption, decryption and the slot codec are provided for end-to-end use;
they use this module's own seeded sampler (the reference's ChaCha20 streams
come from an external package that is not available here).
"""

from __future__ import annotations

import math
import numbers
from dataclasses import replace

import numpy as np
import torch

from fhelium_amd import codec
from fhelium_amd import kernels as K
from fhelium_amd import rns
from fhelium_amd.config import CkksConfig, Preset
from fhelium_amd.csprng import Csprng
from fhelium_amd.core import (Ciphertext, CkksStateError, ConjugationKey, InvalidScaleError, KeySwitchKey,
                              MaximumLevelError, Plaintext, RelinearizationKey, RotationKey,
                              ScaleMismatchError, SecretKey)
from fhelium_amd.keyswitch import HybridKeySwitcher
from fhelium_amd.rotation import decompose_rotation_step
from fhelium_amd.runtime import RnsRuntime


class CkksEngine:
    def __init__(self, preset: Preset | str | CkksConfig = Preset.slots32768_scale40_levels34_int64,
                 *, device: str | torch.device = "cuda", ntt_mode: str = "fast",
                 rng_seed: int | None = None, rng_nonce: int | None = None,
                 allow_sk_gen: bool = True, fused_keyswitch: bool = True, ntt_tuning=None) -> None:
        self.config = preset if isinstance(preset, CkksConfig) else CkksConfig.from_preset(preset)
        # rng_seed / rng_nonce fix the stream for tests and benchmarks only;
        # left unset, the key and nonce come from the operating system's
        # entropy (ckks_engine.py:72-74, 86-87).
        if rng_seed is not None and type(rng_seed) is not int:
            raise TypeError("rng_seed must be an integer or None")
        if rng_nonce is not None and type(rng_nonce) is not int:
            raise TypeError("rng_nonce must be an integer or None")
        # ntt_tuning: an autotune.TuningCache of tuned NTT launch schedules
        self.runtime = RnsRuntime(self.config, device, ntt_mode, ntt_tuning)
        self.device = self.runtime.device
        self.layout = self.runtime.layout
        self.key_switcher = HybridKeySwitcher(self.runtime)
        # fused: eight-launch key switch (fh_keyswitch_fused); otherwise the
        # reference's per-digit op sequence.  Both are bit-identical.
        self.fused_keyswitch = fused_keyswitch
        self.allow_sk_gen = allow_sk_gen
        # ChaCha20 sampler on the device (fhelium_amd/csprng.py); key and
        # ciphertext randomness never leaves HBM.
        self.rng = Csprng(rng_seed, self.device, nonce=rng_nonce)
        self._moduli_cache: dict[tuple[int, ...], torch.Tensor] = {}
        self.secret_key: SecretKey | None = None
        self.relinearization_key: RelinearizationKey | None = None
        self.conjugation_key: ConjugationKey | None = None
        self.rotation_keys: dict[int, RotationKey] = {}
        self.galois_generator = 3

    # ------------------------------------------------------------------
    # Sampling (this module's own seeded sampler)
    # ------------------------------------------------------------------
    def _uniform_qp(self, shape_prefix: tuple[int, ...], rows: tuple[int, ...]) -> torch.Tensor:
        q = self._moduli_cache.get(rows)
        if q is None:
            q = torch.tensor([self.config.moduli[i] for i in rows], dtype=torch.int64, device=self.device)
            self._moduli_cache[rows] = q
        return self.rng.uniform_mod(q, tuple(shape_prefix), self.config.N)

    def _gaussian(self, shape_prefix: tuple[int, ...]) -> torch.Tensor:
        return self.rng.gaussian((*shape_prefix, self.config.N), self.config.sigma)

    def _ternary(self) -> torch.Tensor:
        return self.rng.ternary((self.config.N,))

    def _lift_signed(self, coefficients: torch.Tensor, rows: tuple[int, ...]) -> torch.Tensor:
        """Signed integer coefficients [*, N] -> canonical residues [*, rows, N] on device."""

        q = torch.tensor([self.config.moduli[i] for i in rows], dtype=torch.int64, device=self.device)
        c = coefficients.to(self.device)
        return torch.remainder(c.unsqueeze(-2), q.view(-1, 1))

    def _to_ntt_rows(self, x: torch.Tensor, row_start: int) -> torch.Tensor:
        self.runtime.ntt.forward_to_montgomery_(x, row_start)
        return x

    # ------------------------------------------------------------------
    # Keys
    # ------------------------------------------------------------------
    def generate_secret_key(self) -> SecretKey:
        cfg = self.config
        s = self._ternary()
        rows = tuple(range(cfg.total_num_primes))
        data = self._to_ntt_rows(self._lift_signed(s, rows), 0)
        self.secret_key = SecretKey(data=data, coefficients=s)
        return self.secret_key

    def set_secret_key(self, secret_key: SecretKey) -> SecretKey:
        """Install a secret key (e.g. one read from a value file).  Its signed
        coefficients, needed for rotation keys, are recovered from the first
        NTT row when absent."""

        cfg = self.config
        if tuple(secret_key.data.shape) != (cfg.total_num_primes, cfg.N):
            raise ValueError("secret key must be [QP, N]")
        data = secret_key.data.to(self.device)
        coefficients = secret_key.coefficients
        if coefficients is None:
            std = self.runtime.ntt.inverse_to_standard(data[:1].contiguous(), 0)[0]
            q0 = cfg.moduli[0]
            coefficients = torch.where(std > q0 // 2, std - q0, std)
        self.secret_key = SecretKey(data=data, coefficients=coefficients.to(self.device))
        return self.secret_key

    def set_relinearization_key(self, key: RelinearizationKey) -> RelinearizationKey:
        self.relinearization_key = RelinearizationKey(key.data.to(self.device))
        return self.relinearization_key

    def set_rotation_key(self, key: RotationKey) -> RotationKey:
        step = RotationKey.canonical_step(key.rotation_step, self.config.N)
        g = rns.rotation_galois_element(self.config.N, step, self.galois_generator)
        if key.galois_element != g:
            raise ValueError(f"rotation key for step {step} carries Galois element {key.galois_element} != {g}")
        stored = RotationKey(key.data.to(self.device), rotation_step=step, galois_element=g)
        self.rotation_keys[step] = stored
        return stored

    def _require_secret(self) -> SecretKey:
        if self.secret_key is None:
            if not self.allow_sk_gen:
                raise RuntimeError("no secret key installed")
            self.generate_secret_key()
        return self.secret_key

    def create_key_switch_key(self, source_ntt: torch.Tensor) -> torch.Tensor:
        """k_{d,0} + k_{d,1} s = P s_src + e_d on digit d's rows
        (key_generator.py:243-326).  ``source_ntt``: [QP, N] NTT/Montgomery."""

        cfg = self.config
        rt = self.runtime
        sk = self._require_secret()
        qp = tuple(range(cfg.total_num_primes))
        all_params = rt.rns_params
        src_q = source_ntt[: cfg.num_q_primes].clone()
        K.montgomery_mul_row_scalars_(src_q, rt.p_product_montgomery_q, all_params[:, : cfg.num_q_primes])
        digits = self.layout.digit_specs(0)
        key = torch.empty((len(digits), 2, cfg.total_num_primes, cfg.N), dtype=torch.int64,
                          device=self.device)
        for spec in digits:
            a = self._uniform_qp((), qp)
            e = self._to_ntt_rows(self._lift_signed(self._gaussian(()), qp), 0)
            k0 = K.sub_lazy(e, K.montgomery_mul(a, sk.data, all_params), all_params)
            lo, hi = spec.prime_ids[0], spec.prime_ids[-1] + 1
            k0[lo:hi] = K.add_lazy_with_twice_modulus(k0[lo:hi], src_q[lo:hi], rt.twice_modulus[lo:hi])
            key[spec.key_digit_index, 0] = k0
            key[spec.key_digit_index, 1] = a
        return key

    def generate_relinearization_key(self) -> RelinearizationKey:
        sk = self._require_secret()
        s2 = K.montgomery_mul(sk.data, sk.data, self.runtime.rns_params)
        self.relinearization_key = RelinearizationKey(self.create_key_switch_key(s2))
        return self.relinearization_key

    def _galois_key_data(self, galois_element: int, sk: SecretKey) -> torch.Tensor:
        """Key switch material from sigma_g(s) back to s (key_generator.py:243-326)."""

        idx, sgn = self.runtime.coefficient_galois(galois_element)
        s_rot = sk.coefficients[idx.long()] * sgn.to(torch.int64)
        src = self._to_ntt_rows(self._lift_signed(s_rot, tuple(range(self.config.total_num_primes))), 0)
        return self.create_key_switch_key(src)

    def generate_rotation_key(self, rotation_step: int) -> RotationKey:
        cfg = self.config
        sk = self._require_secret()
        step = RotationKey.canonical_step(rotation_step, cfg.N)
        g = rns.rotation_galois_element(cfg.N, step, self.galois_generator)
        key = RotationKey(self._galois_key_data(g, sk), rotation_step=step, galois_element=g)
        self.rotation_keys[step] = key
        return key

    @property
    def conjugation_galois_element(self) -> int:
        """2N - 1: the automorphism X -> X^{-1}, which conjugates every slot."""

        return 2 * self.config.N - 1

    def create_conjugation_key(self, secret_key: SecretKey | None = None) -> ConjugationKey:
        """Key switch material from sigma_{-1}(s) back to s, not installed
        (ckks_engine.py:362-376, core/keys.py:386-391); the engine's secret
        key by default."""

        sk = self._require_secret() if secret_key is None else secret_key
        if not isinstance(sk, SecretKey):
            raise TypeError(f"expected SecretKey, got {type(sk).__name__}")
        return ConjugationKey(self._galois_key_data(self.conjugation_galois_element, sk))

    def generate_conjugation_key(self) -> ConjugationKey:
        """create_conjugation_key with the engine's secret key, installed."""

        self.conjugation_key = self.create_conjugation_key()
        return self.conjugation_key

    def rotation_key(self, rotation_step: int) -> RotationKey:
        step = RotationKey.canonical_step(rotation_step, self.config.N)
        key = self.rotation_keys.get(step)
        if key is None:
            if not self.allow_sk_gen:
                raise KeyError(f"no rotation key for step {step}")
            key = self.generate_rotation_key(step)
        return key

    # ------------------------------------------------------------------
    # Encode / encrypt / decrypt / decode
    # ------------------------------------------------------------------
    def encrypt(self, values, *, level: int = 0, scale: float | None = None) -> Ciphertext:
        """Symmetric encryption of slot vectors [*batch, <=S] -> coefficient/standard."""

        cfg = self.config
        self.layout.check_level(level)
        scale = cfg.default_scale if scale is None else float(scale)
        sk = self._require_secret()
        values = np.asarray(values)
        batch = values.shape[:-1]
        coeffs = torch.from_numpy(codec.encode(values, cfg.N, scale))
        rows = self.layout.prime_ids(level)
        m = self._to_ntt_rows(self._lift_signed(coeffs, rows), level)
        e = self._to_ntt_rows(self._lift_signed(self._gaussian(batch), rows), level)
        a = self._uniform_qp(batch, rows)
        params = self.runtime.params(level)
        s = sk.data[level: cfg.num_q_primes]
        c0 = K.add_lazy(K.sub_lazy(e, K.montgomery_mul(a, s, params), params), m, params)
        data = torch.stack([c0, a])
        self.runtime.ntt.inverse_to_standard_(data, level)
        return Ciphertext(data=data, level=level, scale=scale)

    def decrypt_residues(self, ct: Ciphertext) -> torch.Tensor:
        """Canonical standard residues of <c, (1, s, s^2, ...)> over the ct's Q rows."""

        cfg = self.config
        sk = self._require_secret()
        rt = self.runtime
        params = rt.params(ct.level)
        s = sk.data[ct.level: cfg.num_q_primes]
        comps = ct.data.clone()
        if ct.polynomial_domain == "coefficient":
            rt.ntt.forward_to_montgomery_(comps, ct.level)
        elif ct.residue_representation != "montgomery":
            raise CkksStateError("NTT-domain values must be Montgomery")
        acc = comps[0]
        power = None
        for i in range(1, ct.component_count):
            power = s if power is None else K.montgomery_mul(power, s, params)
            acc = K.add_lazy(acc, K.montgomery_mul(comps[i], power, params), params)
        acc = acc.contiguous()
        rt.ntt.inverse_to_standard_(acc, ct.level)
        return acc

    def decrypt(self, ct: Ciphertext, crt_primes: int = 3) -> np.ndarray:
        """Complex slots [*batch, S] decoded at the ciphertext's scale."""

        cfg = self.config
        res = self.decrypt_residues(ct).cpu().numpy()
        rows = self.layout.prime_ids(ct.level)
        k = min(crt_primes, len(rows))
        sel = list(range(len(rows) - k, len(rows)))
        moduli = [cfg.moduli[rows[i]] for i in sel]
        values = codec.crt_centered(np.moveaxis(res[..., sel, :], -2, 0), moduli)
        return codec.embed(values) / ct.scale

    def encode(self, values, *, level: int = 0, scale: float | None = None,
               polynomial_domain: str = "coefficient") -> Plaintext:
        """Slots [*batch, <=S] -> operation-ready Montgomery plaintext in the
        requested domain (ckks_plaintext_codec.py:181-316)."""

        cfg = self.config
        self.layout.check_level(level)
        if polynomial_domain not in ("coefficient", "ntt"):
            raise ValueError("polynomial_domain must be 'coefficient' or 'ntt'")
        scale = cfg.default_scale if scale is None else float(scale)
        if not (math.isfinite(scale) and scale > 0):
            raise InvalidScaleError(f"plaintext scale must be positive and finite, got {scale}")
        values = np.asarray(values)
        coeffs = torch.from_numpy(codec.encode(values, cfg.N, scale))
        data = self._lift_signed(coeffs, self.layout.prime_ids(level)).contiguous()
        if polynomial_domain == "ntt":
            self.runtime.ntt.forward_to_montgomery_(data, level)
        else:
            K.to_montgomery_(data, self.runtime.params(level))
        return Plaintext(data=data, level=level, scale=scale, polynomial_domain=polynomial_domain)

    def _slot_positions(self) -> torch.Tensor:
        pos = self.__dict__.get("_slot_pos_dev")
        if pos is None:
            pos = torch.from_numpy(codec._slot_positions(self.config.N).astype(np.int32)).to(self.device)
            self._slot_pos_dev = pos
        return pos

    def encode_on_device(self, values: torch.Tensor, *, level: int = 0, scale: float | None = None,
                         polynomial_domain: str = "coefficient") -> Plaintext:
        """``encode`` with the embedding on the device (fh_ckks_encode):
        values [*batch, <=S] real or complex on the engine's device."""

        from fhelium_amd import _native as nat

        cfg = self.config
        self.layout.check_level(level)
        if polynomial_domain not in ("coefficient", "ntt"):
            raise ValueError("polynomial_domain must be 'coefficient' or 'ntt'")
        scale = cfg.default_scale if scale is None else float(scale)
        if not (math.isfinite(scale) and scale > 0):
            raise InvalidScaleError(f"plaintext scale must be positive and finite, got {scale}")
        v = torch.as_tensor(values, device=self.device)
        if v.dim() < 1 or v.size(-1) > cfg.slots:
            raise ValueError(f"at most {cfg.slots} slots per plaintext")
        batch_shape = tuple(v.shape[:-1])
        flat = torch.zeros((int(np.prod(batch_shape)) if batch_shape else 1, cfg.slots), dtype=torch.complex128,
                           device=self.device)
        flat[:, : v.size(-1)] = v.reshape(-1, v.size(-1)).to(torch.complex128)
        ws = torch.empty((flat.size(0), cfg.N, 2), dtype=torch.float64, device=self.device)
        rows = self.layout.prime_ids(level)
        moduli = self._moduli_cache.get(rows)
        if moduli is None:
            moduli = torch.tensor([cfg.moduli[i] for i in rows], dtype=torch.int64, device=self.device)
            self._moduli_cache[rows] = moduli
        data = torch.empty((flat.size(0), len(rows), cfg.N), dtype=torch.int64, device=self.device)
        nat.call("fh_ckks_encode", torch.view_as_real(flat).data_ptr(), flat.size(0), cfg.logN,
                 self._slot_positions().data_ptr(), scale, moduli.data_ptr(), len(rows), ws.data_ptr(),
                 data.data_ptr(), nat.stream_ptr(self.device))
        data = data.reshape(*batch_shape, len(rows), cfg.N)
        if polynomial_domain == "ntt":
            self.runtime.ntt.forward_to_montgomery_(data, level)
        else:
            K.to_montgomery_(data, self.runtime.params(level))
        return Plaintext(data=data, level=level, scale=scale, polynomial_domain=polynomial_domain)

    def decrypt_on_device(self, ct: Ciphertext) -> torch.Tensor:
        """Slots [*batch, S] complex128 on the device: decryption, then the
        centered CRT over the last two Q rows and the embedding on the
        device (fh_ckks_decode)."""

        from fhelium_amd import _native as nat

        cfg = self.config
        res = self.decrypt_residues(ct)
        lq = res.size(-2)
        if lq < 2:
            raise ValueError("device decode needs at least two Q rows")
        flat = res.reshape(-1, lq, cfg.N)
        qa, qb = cfg.moduli[ct.level + lq - 1], cfg.moduli[ct.level + lq - 2]
        ws = torch.empty((flat.size(0), cfg.N, 2), dtype=torch.float64, device=self.device)
        out = torch.empty((flat.size(0), cfg.slots), dtype=torch.complex128, device=self.device)
        nat.call("fh_ckks_decode", flat[:, lq - 1].data_ptr(), flat[:, lq - 2].data_ptr(), flat.stride(0),
                 flat.size(0), cfg.logN, qa, qb, pow(qa, -1, qb), self._slot_positions().data_ptr(), float(ct.scale),
                 ws.data_ptr(), torch.view_as_real(out).data_ptr(), nat.stream_ptr(self.device))
        return out.reshape(*res.shape[:-2], cfg.slots)

    def _check_plaintext(self, ct: Ciphertext, pt: Plaintext, domain: str) -> None:
        if pt.level != ct.level:
            raise ValueError(f"Plaintext level {pt.level} does not match ciphertext level {ct.level}.")
        if pt.polynomial_domain != domain or pt.residue_representation != "montgomery":
            raise CkksStateError(
                f"plaintext must be {domain}-domain Montgomery residues for this operation")
        if pt.modulus_basis != ct.modulus_basis or pt.data.shape[-2:] != ct.data.shape[-2:]:
            raise ValueError("Plaintext and Ciphertext RNS layouts differ")
        if pt.batch_shape and pt.batch_shape != ct.batch_shape:
            raise ValueError("a batched plaintext must match the ciphertext batch shape")

    def add_plaintext(self, ct: Ciphertext, plaintext: Plaintext, *, inplace: bool = False) -> Ciphertext:
        """c0 + p at equal scale, coefficient/standard (ckks_engine.py:1964-2066)."""

        ct.assert_state(polynomial_domain="coefficient", residue_representation="standard", components=2)
        if plaintext.scale != ct.scale:
            raise ScaleMismatchError(f"add_plaintext: ciphertext scale {ct.scale} != plaintext scale "
                                     f"{plaintext.scale}")
        self._check_plaintext(ct, plaintext, "coefficient")
        params = self.runtime.params(ct.level, ct.includes_p)
        if inplace:
            K.add_prepared_plaintext_component_(ct.c0, plaintext.data, params)
            return ct
        c0 = K.add_prepared_plaintext_component(ct.c0, plaintext.data, params)
        return Ciphertext(torch.stack([c0, ct.c1]), ct.level, ct.scale, "coefficient", "standard",
                          ct.modulus_basis)

    def multiply_plaintext(self, ct: Ciphertext, plaintext: Plaintext, *, inplace: bool = False) -> Ciphertext:
        """c_j p in the NTT domain; scales multiply (ckks_engine.py:2077-2168)."""

        ct.assert_state(polynomial_domain="ntt", residue_representation="montgomery", components=2)
        self._check_plaintext(ct, plaintext, "ntt")
        scale = ct.scale * plaintext.scale
        if not (math.isfinite(scale) and scale > 0):
            raise InvalidScaleError(f"multiply_plaintext result scale {scale} is not positive and finite")
        params = self.runtime.params(ct.level, ct.includes_p)
        pt = plaintext.data
        if pt.dim() < ct.data.dim() - 1:
            pt = pt.expand(ct.data.shape[1:])
        out = torch.empty_like(ct.data)
        out[0] = K.montgomery_mul(ct.c0, pt, params)
        out[1] = K.montgomery_mul(ct.c1, pt, params)
        result = Ciphertext(out, ct.level, scale, "ntt", "montgomery", ct.modulus_basis)
        if inplace:
            ct.data, ct.scale = result.data, result.scale
            return ct
        return result

    def multiply_constants_and_add(self, cts: list[Ciphertext], values: list[float], *,
                                   product_scale: float) -> Ciphertext:
        """sum_t values[t] * cts[t], each constant taken at scale
        product_scale / cts[t].scale, so that every product carries
        product_scale; one pass over the terms (fh_ckks_scalar_combination).

        A constant plaintext (every slot ``v``) at scale s encodes as the
        constant polynomial m = round(v s) (the encoder of
        ckks_engine.py:2077-2168 gives it), and its NTT is m in every slot,
        so each product here has the residues multiply_plaintext gives for
        that plaintext -- in either domain and representation, which the
        terms share.  Terms at lower levels are read at the output's level
        (the highest among them) by dropping their leading rows, as
        mod_switch_to_level does.  The result is not rescaled."""

        if not cts or len(cts) != len(values):
            raise ValueError("one value per ciphertext term is required")
        if len(cts) > K.MAX_COMBINATION_TERMS:
            raise ValueError(f"at most {K.MAX_COMBINATION_TERMS} terms per combination")
        first = cts[0]
        level = max(ct.level for ct in cts)
        for ct in cts:
            if ct.component_count != 2 or ct.batch_shape != first.batch_shape or ct.includes_p:
                raise CkksStateError("terms must be two-component Q-basis ciphertexts of one batch shape")
            if (ct.polynomial_domain, ct.residue_representation) != (first.polynomial_domain,
                                                                      first.residue_representation):
                raise CkksStateError("terms must share polynomial domain and residue representation")
        scale = float(product_scale)
        if not (math.isfinite(scale) and scale > 0):
            raise InvalidScaleError(f"product scale {product_scale} is not positive and finite")
        ints = tuple(int(round(float(v) * (scale / ct.scale))) for v, ct in zip(values, cts))
        scalars = self._combination_scalars(level, ints)
        terms = []
        for ct in cts:
            data = ct.data[..., level - ct.level:, :]
            terms.append(data.reshape(2, -1, data.size(-2), data.size(-1)))
        params = self.runtime.params(level, False)
        out = K.scalar_combination(terms, scalars, params)
        return Ciphertext(out.view(first.data.shape[:-2] + out.shape[-2:]), level, scale,
                          first.polynomial_domain, first.residue_representation, first.modulus_basis)

    def _combination_scalars(self, level: int, ints: tuple[int, ...]) -> torch.Tensor:
        """[T, rows] Montgomery factors m_t R mod q of the rows at ``level``,
        cached per (level, integers): uploading them is a host-to-device copy."""

        cache = self.__dict__.setdefault("_combination_cache", {})
        key = (level, ints)
        hit = cache.get(key)
        if hit is None:
            moduli = [self.config.moduli[i] for i in self.layout.prime_ids(level)]
            rows = [[(m % q) * rns.R % q for q in moduli] for m in ints]
            if len(cache) > 1024:
                cache.clear()
            hit = torch.tensor(rows, dtype=torch.int64).pin_memory().to(self.device, non_blocking=True)
            cache[key] = hit
        return hit

    def add(self, lhs: Ciphertext, rhs: Ciphertext, *, inplace: bool = False) -> Ciphertext:
        """Component-wise canonical sum at equal scale and state (ckks_engine.py:1708-1745, 3352-3377)."""

        self._same_state(lhs, rhs, "add")
        params = self.runtime.params(lhs.level, lhs.includes_p)
        result = lhs if inplace else lhs.clone()
        flat = result.data.view(-1, result.data.size(-2), result.data.size(-1))
        K.add_canonical_(flat, self._operand_for(result, rhs), params)
        return result

    @staticmethod
    def _operand_for(result: Ciphertext, rhs: Ciphertext) -> torch.Tensor:
        """``rhs``'s payload in ``result``'s flat shape; a copy when it shares
        storage with ``result`` (``add_(a, a)``), since the in-place kernels
        refuse an output that overlaps a read-only operand."""

        other = rhs.data
        if other.untyped_storage().data_ptr() == result.data.untyped_storage().data_ptr():
            other = other.clone()
        return other.reshape(-1, result.data.size(-2), result.data.size(-1))

    def _same_state(self, lhs: Ciphertext, rhs: Ciphertext, operation: str) -> None:
        """Equal layout, scale and arithmetic state (ckks_engine.py
        _assert_same_cipher_layout)."""

        self._same_layout(lhs, rhs)
        if lhs.scale != rhs.scale:
            raise ScaleMismatchError(f"{operation}: scales differ ({lhs.scale} != {rhs.scale})")
        if lhs.component_count != rhs.component_count:
            raise CkksStateError(f"{operation}: component counts differ")
        if (lhs.polynomial_domain, lhs.residue_representation) != (rhs.polynomial_domain,
                                                                  rhs.residue_representation):
            raise CkksStateError(f"{operation}: operand states differ ({lhs.polynomial_domain}/"
                                 f"{lhs.residue_representation} vs {rhs.polynomial_domain}/"
                                 f"{rhs.residue_representation})")

    def subtract(self, lhs: Ciphertext, rhs: Ciphertext, *, inplace: bool = False) -> Ciphertext:
        """Component-wise canonical difference at equal scale and state
        (ckks_engine.py:1871-1917)."""

        self._same_state(lhs, rhs, "subtract")
        params = self.runtime.params(lhs.level, lhs.includes_p)
        result = lhs if inplace else lhs.clone()
        flat = result.data.view(-1, result.data.size(-2), result.data.size(-1))
        K.sub_canonical_(flat, self._operand_for(result, rhs), params)
        return result

    def sum_ciphertexts(self, ciphertexts) -> Ciphertext:
        """Canonical sum of equal-state ciphertexts into new storage
        (ckks_engine.py:1761-1797)."""

        if not ciphertexts:
            raise ValueError("sum_ciphertexts needs at least one ciphertext (an empty sum has no level or scale)")
        if not all(isinstance(ct, Ciphertext) for ct in ciphertexts):
            raise TypeError("sum_ciphertexts sums Ciphertext values")
        result = ciphertexts[0].clone()
        for ct in ciphertexts[1:]:
            self.add(result, ct, inplace=True)
        return result

    def sum_ciphertext_batch(self, batch: Ciphertext, *, dim: int = 0) -> Ciphertext:
        """Sum over one batch axis (ckks_engine.py:1801-1869).  The axis is
        moved to the front of one working copy, then halves are added in
        place, so each round is one canonical-add launch over all remaining
        entries; an odd count's last entry is added at the end.  Modular
        addition is exact, so the residues equal any other summation order."""

        if not isinstance(batch, Ciphertext):
            raise TypeError(f"sum_ciphertext_batch expects a Ciphertext, got {type(batch).__name__}")
        axes = len(batch.batch_shape)
        if axes == 0:
            raise ValueError("sum_ciphertext_batch requires a batched value")
        d = dim + axes if dim < 0 else dim
        if not 0 <= d < axes:
            raise IndexError(f"batch dimension {dim} is outside shape {batch.batch_shape}")
        params = self.runtime.params(batch.level, batch.includes_p)
        limbs, n = batch.data.size(-2), batch.data.size(-1)
        work = batch.data.movedim(d + 1, 0).contiguous()
        tails = []
        while work.size(0) > 1:
            count = work.size(0)
            half = count // 2
            if count % 2:
                tails.append(work[count - 1])  # beyond every later round's rows
            K.add_canonical_(work[:half].view(-1, limbs, n), work[half: 2 * half].reshape(-1, limbs, n), params)
            work = work[:half]
        total = work[0].clone()
        flat = total.view(-1, limbs, n)
        for tail in tails:
            K.add_canonical_(flat, tail.reshape(-1, limbs, n), params)
        return replace(batch, data=total)

    def subtract_(self, lhs: Ciphertext, rhs: Ciphertext) -> Ciphertext:
        """subtract in place into ``lhs`` (ckks_engine.py:1910)."""

        return self.subtract(lhs, rhs, inplace=True)

    def negate(self, ct: Ciphertext, *, inplace: bool = False) -> Ciphertext:
        """-c_j mod q, canonical, state unchanged (ckks_engine.py:1919-1957)."""

        params = self.runtime.params(ct.level, ct.includes_p)
        zero = torch.zeros_like(ct.data)
        flat = zero.view(-1, zero.size(-2), zero.size(-1))
        K.sub_canonical_(flat, ct.data.reshape(flat.shape), params)
        if inplace:
            ct.data.copy_(zero)
            return ct
        return replace(ct, data=zero)

    def negate_(self, ct: Ciphertext) -> Ciphertext:
        """negate in place (ckks_engine.py:1959)."""

        return self.negate(ct, inplace=True)

    # ------------------------------------------------------------------
    # Domain transitions (ckks_engine.py:867-985, 2790-2835): strict source
    # states re-checked at call time; functional forms return new storage,
    # the in-place forms (trailing underscore) transform the payload in its
    # storage and update the state fields.  Ciphertexts move between
    # (coefficient, standard) and (ntt, montgomery); plaintexts keep
    # Montgomery residues in both domains.
    # ------------------------------------------------------------------
    def _domain_source(self, value, source_domain: str) -> None:
        operation = ("coefficient_domain_to_ntt_domain" if source_domain == "coefficient"
                     else "ntt_domain_to_coefficient_domain")
        article = "a coefficient-domain" if source_domain == "coefficient" else "an NTT-domain"
        if isinstance(value, Plaintext):
            kind = "plaintext"
            if value.modulus_basis not in ("Q", "QP"):
                raise ValueError(f"Unsupported plaintext modulus_basis: {value.modulus_basis!r}")
            if value.polynomial_domain != source_domain:
                raise CkksStateError(f"{operation} requires {article} {kind}")
            if value.residue_representation != "montgomery":
                raise CkksStateError(f"{operation} requires Montgomery plaintext residues")
        elif isinstance(value, Ciphertext):
            if value.modulus_basis not in ("Q", "QP"):
                raise ValueError(f"Unsupported ciphertext modulus_basis: {value.modulus_basis!r}")
            if value.polynomial_domain != source_domain:
                raise CkksStateError(f"{operation} requires {article} ciphertext")
            if source_domain == "coefficient" and value.residue_representation != "standard":
                raise CkksStateError(f"{operation} requires standard ciphertext residues")
            if source_domain == "ntt" and value.residue_representation != "montgomery":
                raise CkksStateError(f"{operation} requires Montgomery ciphertext residues")
        else:
            raise TypeError(f"Domain transition expects Ciphertext or Plaintext, got {type(value).__name__}")
        self._check_rows(value, operation)

    def _check_rows(self, value, operation: str) -> None:
        """The payload holds exactly the rows of its level and basis."""

        self.layout.check_level(value.level)
        rows = value.data.size(-2)
        start = self.runtime.row_start(value.level)
        want = self.config.num_q_primes - start + (self.config.num_p_primes if value.modulus_basis == "QP" else 0)
        if rows != want or value.data.size(-1) != self.config.N:
            raise ValueError(f"{operation}: {rows} limbs of {value.data.size(-1)} coefficients do not match "
                             f"level {value.level} basis {value.modulus_basis} ({want} x {self.config.N})")

    def _forward_ntt_(self, value):
        rt = self.runtime
        start = rt.row_start(value.level)
        if isinstance(value, Ciphertext):
            rt.ntt.forward_to_montgomery_(value.data, start)
            value.residue_representation = "montgomery"
        else:
            rt.ntt.forward_montgomery_(value.data, start)
        value.polynomial_domain = "ntt"
        return value

    def _inverse_ntt_(self, value):
        rt = self.runtime
        start = rt.row_start(value.level)
        if isinstance(value, Ciphertext):
            rt.ntt.inverse_to_standard_(value.data, start)
            value.residue_representation = "standard"
        else:
            rt.ntt.inverse_montgomery_(value.data, start)
        value.polynomial_domain = "coefficient"
        return value

    def coefficient_domain_to_ntt_domain_(self, value):
        """coefficient_domain_to_ntt_domain in place (ckks_engine.py:914-928)."""

        self._domain_source(value, "coefficient")
        return self._forward_ntt_(value)

    def ntt_domain_to_coefficient_domain(self, value):
        """NTT -> coefficient domain (ckks_engine.py:940-961): a ciphertext's
        (ntt, montgomery) becomes canonical (coefficient, standard), a
        plaintext's (ntt, montgomery) becomes (coefficient, montgomery)."""

        self._domain_source(value, "ntt")
        if isinstance(value, Plaintext):
            return self._inverse_ntt_(replace(value, data=value.data.clone()))
        rt = self.runtime
        data = rt.ntt.inverse_to_standard(value.data, rt.row_start(value.level))
        return Ciphertext(data, value.level, value.scale, "coefficient", "standard", value.modulus_basis,
                          value.context_id)

    def add_(self, lhs: Ciphertext, rhs: Ciphertext) -> Ciphertext:
        """add in place into ``lhs`` (ckks_engine.py:1752)."""

        return self.add(lhs, rhs, inplace=True)

    def add_plaintext_(self, ct: Ciphertext, plaintext: Plaintext) -> Ciphertext:
        """add_plaintext in place into ``ct`` (ckks_engine.py:2068)."""

        return self.add_plaintext(ct, plaintext, inplace=True)

    def multiply_plaintext_(self, ct: Ciphertext, plaintext: Plaintext) -> Ciphertext:
        """multiply_plaintext in place into ``ct`` (ckks_engine.py:2170)."""

        return self.multiply_plaintext(ct, plaintext, inplace=True)

    def ntt_domain_to_coefficient_domain_(self, value):
        """ntt_domain_to_coefficient_domain in place (ckks_engine.py:973-985)."""

        self._domain_source(value, "ntt")
        return self._inverse_ntt_(value)

    def _on_device(self, t: torch.Tensor) -> bool:
        """``t`` lives on the engine's device (an index-less device means the
        current one of its type)."""

        if t.device.type != self.device.type:
            return False
        index = self.device.index
        if index is None and self.device.type == "cuda":
            index = torch.cuda.current_device()
        return index is None or t.device.index == index

    def validate_ciphertext(self, ciphertext: Ciphertext) -> None:
        """Re-check a ciphertext's construction invariants and that it is
        complete for this engine (device, dtype, level, basis rows, N);
        no kernel runs and nothing is mutated (ckks_engine.py:2560-2570)."""

        if not isinstance(ciphertext, Ciphertext):
            raise TypeError(f"Expected Ciphertext, got {type(ciphertext).__name__}")
        ct = ciphertext
        Ciphertext(ct.data, ct.level, ct.scale, ct.polynomial_domain, ct.residue_representation,
                   ct.modulus_basis, ct.context_id)
        if ct.data.dtype != torch.int64 or not self._on_device(ct.data):
            raise ValueError(f"ciphertext payload must be int64 on {self.device}, got {ct.data.dtype} "
                             f"on {ct.data.device}")
        self._check_rows(ct, "validate_ciphertext")

    def validate_key_switch_key(self, key: KeySwitchKey) -> None:
        """Key switch material complete for this engine: [digits at level 0,
        2, QP rows, N] int64 on the engine's device; a rotation key's step
        canonical and its Galois element the step's (ckks_engine.py:2593-2630)."""

        if not isinstance(key, KeySwitchKey):
            raise TypeError(f"Expected KeySwitchKey, got {type(key).__name__}")
        cfg = self.config
        want = (len(self.layout.digit_specs(0)), 2, cfg.total_num_primes, cfg.N)
        data = key.data
        if not isinstance(data, torch.Tensor) or data.dtype != torch.int64 or not self._on_device(data):
            raise ValueError(f"key material must be an int64 tensor on {self.device}")
        if tuple(data.shape) != want:
            raise ValueError(f"key material must be {list(want)}, got {list(data.shape)}")
        if isinstance(key, RotationKey):
            if type(key.rotation_step) is not int:
                raise TypeError(f"RotationKey rotation_step must be an integer, got "
                                f"{type(key.rotation_step).__name__}")
            step = RotationKey.canonical_step(key.rotation_step, cfg.N)
            if key.rotation_step != step:
                raise ValueError(f"RotationKey rotation_step is not canonical: {key.rotation_step} != {step}")
            g = rns.rotation_galois_element(cfg.N, step, self.galois_generator)
            if key.galois_element != g:
                raise ValueError(f"RotationKey for step {step} carries Galois element {key.galois_element} != {g}")

    def _residue_source(self, plaintext, source: str) -> None:
        """A coefficient-domain plaintext in ``source`` residues
        (ckks_engine.py:2835-2857)."""

        operation = ("standard_residues_to_montgomery_residues" if source == "standard"
                     else "montgomery_residues_to_standard_residues")
        if not isinstance(plaintext, Plaintext):
            raise TypeError(f"{operation} expects a Plaintext, got {type(plaintext).__name__}")
        if plaintext.modulus_basis not in ("Q", "QP"):
            raise ValueError(f"Unsupported plaintext modulus_basis: {plaintext.modulus_basis!r}")
        if plaintext.polynomial_domain != "coefficient":
            raise ValueError(f"{operation} requires a coefficient-domain plaintext")
        if plaintext.residue_representation != source:
            raise ValueError(f"{operation} requires {'standard' if source == 'standard' else 'Montgomery'} residues")
        self._check_rows(plaintext, operation)

    def _convert_residues_(self, plaintext: Plaintext, target: str) -> Plaintext:
        params = self.runtime.params(plaintext.level, plaintext.modulus_basis == "QP")
        rows = plaintext.data.view(-1, plaintext.data.size(-2), plaintext.data.size(-1))
        (K.to_montgomery_ if target == "montgomery" else K.from_montgomery_)(rows, params)
        plaintext.residue_representation = target
        return plaintext

    def standard_residues_to_montgomery_residues(self, plaintext: Plaintext) -> Plaintext:
        """r -> r R mod q on every row of a coefficient-domain plaintext, into
        new storage (ckks_engine.py:987-1003, 3134-3150)."""

        self._residue_source(plaintext, "standard")
        return self._convert_residues_(replace(plaintext, data=plaintext.data.clone()), "montgomery")

    def standard_residues_to_montgomery_residues_(self, plaintext: Plaintext) -> Plaintext:
        """standard_residues_to_montgomery_residues in place (ckks_engine.py:1005-1019)."""

        self._residue_source(plaintext, "standard")
        return self._convert_residues_(plaintext, "montgomery")

    def montgomery_residues_to_standard_residues(self, plaintext: Plaintext) -> Plaintext:
        """Montgomery reduction of every residue to its canonical value, into
        new storage (ckks_engine.py:1021-1039, 3170-3186)."""

        self._residue_source(plaintext, "montgomery")
        return self._convert_residues_(replace(plaintext, data=plaintext.data.clone()), "standard")

    def montgomery_residues_to_standard_residues_(self, plaintext: Plaintext) -> Plaintext:
        """montgomery_residues_to_standard_residues in place (ckks_engine.py:1041-1054)."""

        self._residue_source(plaintext, "montgomery")
        return self._convert_residues_(plaintext, "standard")

    # ------------------------------------------------------------------
    # The hot path
    # ------------------------------------------------------------------
    def coefficient_domain_to_ntt_domain(self, ct):
        """Coefficient -> NTT domain (ckks_engine.py:876-903, 3094-3109): a
        ciphertext's (coefficient, standard) becomes (ntt, montgomery), a
        plaintext's (coefficient, montgomery) becomes (ntt, montgomery)."""

        self._domain_source(ct, "coefficient")
        if isinstance(ct, Plaintext):
            return self._forward_ntt_(replace(ct, data=ct.data.clone()))
        data = self.runtime.ntt.forward_to_montgomery(ct.data, self.runtime.row_start(ct.level))
        return Ciphertext(data, ct.level, ct.scale, "ntt", "montgomery", ct.modulus_basis, ct.context_id)

    def _same_layout(self, a: Ciphertext, b: Ciphertext) -> None:
        if a.level != b.level or a.data.shape[1:] != b.data.shape[1:] or a.modulus_basis != b.modulus_basis:
            raise ValueError("ciphertext layouts differ (level, batch shape or basis)")

    def multiply(self, lhs: Ciphertext, rhs: Ciphertext) -> Ciphertext:
        """d0 = a0 b0, d1 = a0 b1 + a1 b0, d2 = a1 b1 (ckks_engine.py:1530-1606)."""

        lhs.assert_state(polynomial_domain="ntt", residue_representation="montgomery", components=2)
        rhs.assert_state(polynomial_domain="ntt", residue_representation="montgomery", components=2)
        self._same_layout(lhs, rhs)
        scale = lhs.scale * rhs.scale
        if not math.isfinite(scale) or scale <= 0:
            raise InvalidScaleError("product scale is not positive and finite")
        out = torch.empty((3, *lhs.data.shape[1:]), dtype=torch.int64, device=self.device)
        K.tensor_product(out, lhs.data, rhs.data, self.runtime.params(lhs.level, lhs.includes_p))
        return Ciphertext(out, lhs.level, scale, "ntt", "montgomery", lhs.modulus_basis)

    def relinearize(self, ct: Ciphertext, relinearization_key: RelinearizationKey | None = None) -> Ciphertext:
        """Key-switch d2 from s^2 to s (ckks_engine.py:1608-1675)."""

        ct.assert_state(polynomial_domain="ntt", modulus_basis="Q", residue_representation="montgomery",
                        components=3)
        key = relinearization_key or self.relinearization_key
        if key is None:
            key = self.generate_relinearization_key()
        rt = self.runtime
        d = rt.ntt.inverse_to_standard(ct.data, rt.row_start(ct.level))
        if self.fused_keyswitch:
            out = self.key_switcher.key_switch_fused(d[2], key, ct.level, base0=d[0], base1=d[1])
        else:
            corr0, corr1 = self.key_switcher.apply_key_switch(d[2], key, ct.level)
            out = torch.empty((2, *d.shape[1:]), dtype=torch.int64, device=self.device)
            K.add_two_canonicalize(out, d[0], d[1], corr0, corr1, rt.params(ct.level))
        return Ciphertext(out, ct.level, ct.scale, "coefficient", "standard", ct.modulus_basis)

    def _rescale_tables(self, ct: Ciphertext, rounding: str):
        if rounding not in ("nearest", "floor"):
            raise ValueError(f"rounding must be 'nearest' or 'floor', got {rounding!r}")
        ct.assert_state(polynomial_domain="coefficient", residue_representation="standard")
        cfg = self.config
        next_level = ct.level + 1
        if next_level >= cfg.public_level_count:
            raise MaximumLevelError(ct.level, cfg.public_level_count)
        rt = self.runtime
        start = rt.row_start(ct.level)
        stop = cfg.total_num_primes if ct.includes_p else cfg.num_q_primes
        if ct.data.size(-2) != stop - start:
            raise ValueError(f"rescale: {ct.data.size(-2)} limbs do not match level {ct.level} basis "
                             f"{ct.modulus_basis} ({stop - start})")
        scale = ct.scale / float(cfg.moduli[start])
        if not math.isfinite(scale) or scale <= 0:
            raise InvalidScaleError("rescaled scale is not positive and finite")
        return (next_level, scale, rt.params_rows(start + 1, stop), rt.rescale_inverses[ct.level][: stop - start - 1],
                cfg.moduli[start])

    def rescale_to_next_level(self, ct: Ciphertext, *, rounding: str = "nearest") -> Ciphertext:
        """Divide-round-drop of the leading Q prime (ckks_engine.py:1131-1174,
        ckks_rescale.py:84-131, 224-377): ``rounding="nearest"`` rounds the
        quotient to the nearest integer, ``"floor"`` subtracts the least
        non-negative dropped residue before the exact division."""

        next_level, scale, remaining, inv, dropped_prime = self._rescale_tables(ct, rounding)
        if rounding == "nearest":
            out = K.rescale_components(ct.data, inv, remaining, dropped_prime // 2)
        else:
            out = K.rescale_drop_leading_prime_truncate(ct.data[..., 1:, :], inv, ct.data[..., 0, :], remaining)
        return Ciphertext(out, next_level, scale, ct.polynomial_domain, ct.residue_representation, ct.modulus_basis,
                          ct.context_id)

    def rescale_to_next_level_(self, ct: Ciphertext, *, rounding: str = "nearest") -> Ciphertext:
        """rescale_to_next_level in place (ckks_engine.py:1176-1246): the
        remaining rows are updated in the original allocation, ``ct.data``
        becomes the view of those rows, and level and scale advance."""

        next_level, scale, remaining, inv, dropped_prime = self._rescale_tables(ct, rounding)
        rows = ct.data[..., 1:, :]
        # the dropped row is read while the rows after it are written: a copy
        # of it keeps the operands apart
        dropped = ct.data[..., 0, :].clone()
        if rounding == "nearest":
            K.rescale_drop_leading_prime_nearest_(rows, inv, dropped, remaining, dropped_prime // 2)
        else:
            K.rescale_drop_leading_prime_truncate_(rows, inv, dropped, remaining)
        ct.data = rows
        ct.level = next_level
        ct.scale = scale
        return ct

    def mod_switch_to_level(self, ct: Ciphertext, target_level: int) -> Ciphertext:
        """Restrict ``ct`` to the basis at ``target_level`` by dropping its
        leading Q rows; scale and every other state axis are kept, and the
        result owns its storage (ckks_engine.py:1278-1320, 3209-3262)."""

        drop = self._mod_switch_drop(ct, target_level)
        data = ct.data[..., drop:, :].contiguous()
        return Ciphertext(data, target_level, ct.scale, ct.polynomial_domain, ct.residue_representation,
                          ct.modulus_basis, ct.context_id)

    def mod_switch_to_level_(self, ct: Ciphertext, target_level: int) -> Ciphertext:
        """mod_switch_to_level in place (ckks_engine.py:1317-1344): ``ct.data``
        becomes the view of the kept rows of its own allocation; the current
        level is a no-op."""

        drop = self._mod_switch_drop(ct, target_level)
        ct.data = ct.data[..., drop:, :]
        ct.level = target_level
        return ct

    def _mod_switch_drop(self, ct: Ciphertext, target_level: int) -> int:
        """Leading Q rows a mod switch to ``target_level`` drops."""

        if not isinstance(target_level, int) or isinstance(target_level, bool):
            raise TypeError("target_level must be an integer")
        self.layout.check_level(target_level)
        if target_level < ct.level:
            raise ValueError(f"cannot mod-switch from level {ct.level} back to {target_level}")
        return target_level - ct.level

    @property
    def public_level_count(self) -> int:
        return self.config.public_level_count

    @property
    def final_public_level(self) -> int:
        return self.config.public_level_count - 1

    @property
    def num_slots(self) -> int:
        return self.config.slots

    def rescale_to_next_drop_prime(self, *, level: int) -> int:
        """The Q prime a rescale at ``level`` divides by (ckks_engine.py:1056-1087)."""

        if type(level) is not int:
            raise TypeError("level must be an integer")
        if level < 0:
            raise ValueError(f"level must be non-negative, got {level}")
        if level >= self.final_public_level:
            raise MaximumLevelError(level, self.config.public_level_count)
        return int(self.config.moduli[self.runtime.row_start(level)])

    def rescale_to_next_output_scale(self, input_scale: float, *, level: int) -> float:
        """input_scale / q_drop in binary64, as rescale_to_next_level sets it
        (ckks_engine.py:1089-1129)."""

        real = isinstance(input_scale, numbers.Real) and not isinstance(input_scale, bool)
        scale = float(input_scale) if real else math.nan
        if not (math.isfinite(scale) and scale > 0):
            raise InvalidScaleError(f"rescale input scale must be positive and finite, got {input_scale!r}")
        out = scale / float(self.rescale_to_next_drop_prime(level=level))
        if not (math.isfinite(out) and out > 0):
            raise InvalidScaleError(f"rescale result scale {out!r} is not positive and finite")
        return out

    def _next_level(self, ct: Ciphertext) -> int:
        if ct.level + 1 >= self.config.public_level_count:
            raise MaximumLevelError(ct.level, self.config.public_level_count)
        return ct.level + 1

    def mod_switch_to_next_level(self, ct: Ciphertext) -> Ciphertext:
        """Drop the leading Q row, keeping the scale; MaximumLevelError at the
        final level (ckks_engine.py:1208-1246)."""

        return self.mod_switch_to_level(ct, self._next_level(ct))

    def mod_switch_to_next_level_(self, ct: Ciphertext) -> Ciphertext:
        """mod_switch_to_next_level in place (ckks_engine.py:1248-1276)."""

        return self.mod_switch_to_level_(ct, self._next_level(ct))

    def reinterpret_at_scale(self, ct: Ciphertext, target_scale: float, *,
                             max_relative_change: float | None = None) -> Ciphertext:
        """Same residues read at ``target_scale``; with a bound, the symmetric
        relative change max(old/new, new/old) - 1 must not exceed it
        (ckks_engine.py:1346-1393, 3264-3300)."""

        out = ct.clone()
        out.scale = self._reinterpreted_scale(ct, target_scale, max_relative_change)
        return out

    def reinterpret_at_scale_(self, ct: Ciphertext, target_scale: float, *,
                              max_relative_change: float | None = None) -> Ciphertext:
        """reinterpret_at_scale on ``ct`` itself: only its scale changes
        (ckks_engine.py:1396-1430)."""

        ct.scale = self._reinterpreted_scale(ct, target_scale, max_relative_change)
        return ct

    @staticmethod
    def _reinterpreted_scale(ct: Ciphertext, target_scale: float, max_relative_change: float | None) -> float:
        target = float(target_scale)
        if not (math.isfinite(target) and target > 0):
            raise InvalidScaleError(f"target scale must be positive and finite, got {target_scale}")
        if max_relative_change is not None:
            limit = float(max_relative_change)
            if not (math.isfinite(limit) and limit >= 0.0):
                raise ValueError("max_relative_change must be a non-negative finite float")
            if max(ct.scale / target, target / ct.scale) - 1.0 > limit:
                raise ScaleMismatchError(f"reinterpret_at_scale: {ct.scale} -> {target} exceeds {limit}")
        return target

    def _assert_rotation_ciphertext(self, ct: Ciphertext) -> None:
        ct.assert_state(polynomial_domain="coefficient", modulus_basis="Q", residue_representation="standard",
                        components=2)

    def rotate_by_step(self, ct: Ciphertext, rotation_step: int) -> Ciphertext:
        """m'_j = m_{(j - r) mod S} (ckks_engine.py:2183-2207, 3822-3897)."""

        if not isinstance(rotation_step, int):
            raise TypeError("rotation_step must be an integer")
        self._assert_rotation_ciphertext(ct)
        step = RotationKey.canonical_step(rotation_step, self.config.N)
        result = ct.clone() if step == 0 else ct
        for hop in self.rotation_path(step):
            result = self.rotate_with_key(result, self.rotation_key(hop))
        return result

    def rotate_by_step_(self, ct: Ciphertext, rotation_step: int) -> Ciphertext:
        """``ct`` replaced by rotate_by_step's result (ckks_engine.py:2333-2346):
        the object keeps its identity, ``ct.data`` becomes the new storage."""

        result = self.rotate_by_step(ct, rotation_step)
        ct.data = result.data
        ct.level, ct.scale = result.level, result.scale
        ct.polynomial_domain, ct.residue_representation = result.polynomial_domain, result.residue_representation
        ct.modulus_basis = result.modulus_basis
        return ct

    def switch_key(self, ct: Ciphertext, key: KeySwitchKey) -> Ciphertext:
        """c0 + c1 s_src -> c0' + c1' s_dst with a key generated for that
        direction (ckks_engine.py:1677-1706): a two-component coefficient
        standard Q ciphertext in, a new one out, same level and scale."""

        ct.assert_state(polynomial_domain="coefficient", modulus_basis="Q", residue_representation="standard",
                        components=2)
        if not isinstance(key, KeySwitchKey):
            raise TypeError(f"switch_key expects a KeySwitchKey, got {type(key).__name__}")
        self.layout.check_level(ct.level)
        return self._switch_rotated(ct, ct.data, key)

    def rotation_path(self, rotation_step: int) -> list[int]:
        """Canonical steps whose rotations, applied in order, realise
        ``rotation_step`` (ckks_engine.py:3726-3751): the step itself when
        its key is installed or may be generated, otherwise the shortest
        composition over the installed keys (rotation.decompose_rotation_step).
        Raises KeyError when neither exists."""

        step = RotationKey.canonical_step(rotation_step, self.config.N)
        if step == 0:
            return []
        if step in self.rotation_keys or self.allow_sk_gen:
            return [step]
        return [RotationKey.canonical_step(s, self.config.N)
                for s in decompose_rotation_step(step, self.config.slots, self.rotation_keys)]

    def rotate_with_key(self, ct: Ciphertext, key: RotationKey) -> Ciphertext:
        self._assert_rotation_ciphertext(ct)
        rt = self.runtime
        g = key.galois_element
        if self.fused_keyswitch:
            entries = int(np.prod(ct.batch_shape)) if ct.batch_shape else 1
            return self._switch_rotated_fused(ct, key, [g] * entries)
        idx, sgn = rt.coefficient_galois(g)
        twice = rt.twice_modulus[rt.row_start(ct.level): self.config.num_q_primes]
        rotated = K.apply_coefficient_galois_automorphism(ct.data, idx, sgn, twice)
        return self._switch_rotated(ct, rotated, key)

    def conjugate(self, ct: Ciphertext, key: ConjugationKey | None = None) -> Ciphertext:
        """m'_j = conj(m_j): X -> X^{-1} on both components, then the key
        switch from sigma_{-1}(s) back to s (ckks_engine.py:2348-2400); the
        installed conjugation key by default."""

        self._assert_rotation_ciphertext(ct)
        key = self.conjugation_key if key is None else key
        if not isinstance(key, ConjugationKey):
            raise TypeError("conjugate needs a ConjugationKey (generate_conjugation_key installs one)")
        g = self.conjugation_galois_element
        if self.fused_keyswitch:
            entries = int(np.prod(ct.batch_shape)) if ct.batch_shape else 1
            return self._switch_rotated_fused(ct, key, [g] * entries)
        rt = self.runtime
        idx, sgn = rt.coefficient_galois(g)
        twice = rt.twice_modulus[rt.row_start(ct.level): self.config.num_q_primes]
        return self._switch_rotated(ct, K.apply_coefficient_galois_automorphism(ct.data, idx, sgn, twice), key)

    def _switch_rotated_fused(self, ct: Ciphertext, key, galois_elements: list[int]) -> Ciphertext:
        """Automorphism folded into the key switch: the kernels read c1 and
        c0 through each entry's Galois gather."""

        out = self.key_switcher.key_switch_fused(ct.data[1], key, ct.level, base0=ct.data[0],
                                                 galois_elements=galois_elements)
        return Ciphertext(out, ct.level, ct.scale, "coefficient", "standard", ct.modulus_basis)

    def _switch_rotated(self, ct: Ciphertext, rotated: torch.Tensor, key) -> Ciphertext:
        rt = self.runtime
        if self.fused_keyswitch:
            out = self.key_switcher.key_switch_fused(rotated[1], key, ct.level, base0=rotated[0])
        else:
            corr0, corr1 = self.key_switcher.apply_key_switch(rotated[1], key, ct.level)
            out = torch.empty_like(rotated)
            out[0] = K.add_canonical(rotated[0], corr0, rt.params(ct.level))
            out[1] = corr1
        return Ciphertext(out, ct.level, ct.scale, "coefficient", "standard", ct.modulus_basis)

    def rotate_by_steps(self, ct: Ciphertext, rotation_steps: list[int]) -> Ciphertext:
        """Rotate batch entry b of a batched ciphertext [2, B, limb, N] by
        rotation_steps[b] in one batched key switch (each entry uses its own
        rotation key; no reference counterpart -- the reference batches only
        homogeneous operations)."""

        self._assert_rotation_ciphertext(ct)
        if len(ct.batch_shape) != 1 or ct.batch_shape[0] != len(rotation_steps):
            raise ValueError("rotate_by_steps needs a [2, B, limb, N] ciphertext and B steps")
        paths = [self.rotation_path(s) for s in rotation_steps]
        if any(len(p) != 1 for p in paths):
            return self._rotate_by_paths(ct, paths)
        rt = self.runtime
        keys = [self.rotation_key(p[0]) for p in paths]
        if self.fused_keyswitch:
            return self._switch_rotated_fused(ct, keys, [k.galois_element for k in keys])
        twice = rt.twice_modulus[rt.row_start(ct.level): self.config.num_q_primes]
        rotated = torch.empty_like(ct.data)
        for b, key in enumerate(keys):
            idx, sgn = rt.coefficient_galois(key.galois_element)
            rotated[:, b] = K.apply_coefficient_galois_automorphism(ct.data[:, b], idx, sgn, twice)
        return self._switch_rotated(ct, rotated, keys)

    def _rotate_by_paths(self, ct: Ciphertext, paths: list[list[int]]) -> Ciphertext:
        """Per-entry composed rotations: hop k of every entry whose path has
        one runs in one batched key switch; entry b's result equals its hops
        applied one at a time."""

        data = ct.data.clone()
        for hop in range(max(len(p) for p in paths)):
            active = [b for b, p in enumerate(paths) if len(p) > hop]
            index = torch.tensor(active, dtype=torch.long, device=self.device)
            sub = Ciphertext(data.index_select(1, index), ct.level, ct.scale, ct.polynomial_domain,
                             ct.residue_representation, ct.modulus_basis, ct.context_id)
            out = self.rotate_by_steps(sub, [paths[b][hop] for b in active])
            data.index_copy_(1, index, out.data)
        return Ciphertext(data, ct.level, ct.scale, ct.polynomial_domain, ct.residue_representation,
                          ct.modulus_basis, ct.context_id)

    def rotate_many_with_keys(self, ct: Ciphertext, keys: list[RotationKey], *,
                              use_hoisting: bool = True) -> list[Ciphertext]:
        """Hoisted rotations: one ModUp + NTT of c1 shared by every key
        (ckks_engine.py:2298-2331, 3767-3820).  As in the reference, fewer
        than two non-zero steps fall back to direct rotations.  Every result
        owns its storage."""

        self._assert_rotation_ciphertext(ct)
        for key in keys:
            if RotationKey.canonical_step(key.rotation_step, self.config.N) != key.rotation_step:
                raise ValueError("rotation key must carry a canonical signed step")
        return self._apply_rotation_key_sequence(ct, [(k.rotation_step, k) for k in keys],
                                                 use_hoisting=use_hoisting)

    def rotate_many_by_steps(self, ct: Ciphertext, rotation_steps: list[int], *,
                             use_hoisting: bool = True) -> list[Ciphertext]:
        """Rotations by ordered signed steps with engine-owned keys
        (ckks_engine.py:2247-2297): steps with a direct key (installed or
        generated) share one hoisted decomposition; a step that needs a
        composed key path is evaluated on its own."""

        self._assert_rotation_ciphertext(ct)
        for s in rotation_steps:
            if not isinstance(s, int):
                raise TypeError("All rotation steps must be integers")
        steps = [RotationKey.canonical_step(s, self.config.N) for s in rotation_steps]
        paths = [self.rotation_path(s) for s in steps]
        if not use_hoisting:
            return [self._apply_path(ct, p) for p in paths]
        direct = [(i, s, self.rotation_key(p[0]) if p else None)
                  for i, (s, p) in enumerate(zip(steps, paths)) if len(p) <= 1]
        outputs = {i: self._apply_path(ct, p) for i, p in enumerate(paths) if len(p) > 1}
        results = self._apply_rotation_key_sequence(ct, [(s, k) for _, s, k in direct], use_hoisting=True)
        outputs.update(zip([i for i, _, _ in direct], results))
        return [outputs[i] for i in range(len(paths))]

    def _apply_path(self, ct: Ciphertext, path: list[int]) -> Ciphertext:
        result = ct.clone()
        for hop in path:
            result = self.rotate_with_key(result, self.rotation_key(hop))
        return result

    def _apply_rotation_key_sequence(self, ct: Ciphertext, entries: list[tuple[int, RotationKey | None]], *,
                                     use_hoisting: bool) -> list[Ciphertext]:
        """Ordered direct keys (ckks_engine.py:3767-3820): hoisted when at
        least two steps are non-zero, otherwise one rotation per key; step 0
        returns a copy."""

        nonzero = [k for s, k in entries if s != 0]
        if not use_hoisting or len(nonzero) < 2:
            return [ct.clone() if s == 0 else self.rotate_with_key(ct, k) for s, k in entries]
        if self.fused_keyswitch and ct.batch_shape in ((), (1,)):
            batch = self._rotate_hoisted_fused(ct, nonzero)
            it = iter(range(len(nonzero)))
            outs = []
            for s, _ in entries:
                if s == 0:
                    outs.append(ct.clone())
                    continue
                e = next(it)
                data = batch.data[:, e].reshape(ct.data.shape).clone()
                outs.append(Ciphertext(data, ct.level, ct.scale, "coefficient", "standard", ct.modulus_basis))
            return outs
        hoisted = iter(self._rotate_hoisted_unfused(ct, nonzero))
        return [ct.clone() if s == 0 else Ciphertext(next(hoisted), ct.level, ct.scale, "coefficient", "standard",
                                                     ct.modulus_basis) for s, _ in entries]

    def _rotate_hoisted_unfused(self, ct: Ciphertext, keys: list[RotationKey]) -> list[torch.Tensor]:
        """The reference's hoisted op sequence (prepare_rotation_digits, then
        per key the NTT-domain gather, MAC and ModDown) on the per-op kernels."""

        rt = self.runtime
        prepared = self.key_switcher.prepare_rotation_digits(ct.c1, ct.level)
        twice = rt.twice_modulus[rt.row_start(ct.level): self.config.num_q_primes]
        params = rt.params(ct.level)
        outs = []
        for key in keys:
            g = key.galois_element
            idx, sgn = rt.coefficient_galois(g)
            corr0, corr1 = self.key_switcher.apply_prepared_rotation(prepared, key, ct.level, g)
            c0 = K.apply_coefficient_galois_automorphism(ct.c0, idx, sgn, twice)
            outs.append(torch.stack([K.add_canonical(c0, corr0, params), corr1]))
        return outs

    def _rotate_hoisted_fused(self, ct: Ciphertext, keys: list[RotationKey]) -> Ciphertext:
        data = self.key_switcher.key_switch_hoisted(ct.data[1], ct.data[0], keys, ct.level)
        return Ciphertext(data, ct.level, ct.scale, "coefficient", "standard", ct.modulus_basis)

    def rotate_hoisted(self, ct: Ciphertext, rotation_steps: list[int]) -> Ciphertext:
        """Entry e of the result ([2, K, Q_level, N]) is the unbatched ``ct``
        rotated by rotation_steps[e] with its direct key, all through one
        shared decomposition and basis extension of c1 (the hoisted residues
        of rotate_many_with_keys, returned as one batched ciphertext; no
        reference counterpart).  Step 0 entries hold ``ct`` itself."""

        self._assert_rotation_ciphertext(ct)
        if ct.batch_shape not in ((), (1,)):
            raise ValueError("rotate_hoisted takes one unbatched ciphertext")
        steps = [RotationKey.canonical_step(s, self.config.N) for s in rotation_steps]
        nonzero = [s for s in steps if s != 0]
        lq, n = ct.data.shape[-2:]
        if self.fused_keyswitch and nonzero and len(nonzero) == len(steps):
            return self._rotate_hoisted_fused(ct, [self.rotation_key(s) for s in steps])
        out = torch.empty((2, len(steps), lq, n), dtype=torch.int64, device=self.device)
        if nonzero:
            keys = [self.rotation_key(s) for s in nonzero]
            if self.fused_keyswitch:
                rotated = self._rotate_hoisted_fused(ct, keys).data
            else:
                flat = ct.clone()
                flat.data = ct.data.reshape(2, lq, n)
                rotated = torch.stack(self._rotate_hoisted_unfused(flat, keys), dim=1)
            pos = [e for e, s in enumerate(steps) if s != 0]
            out[:, pos] = rotated
        zero = [e for e, s in enumerate(steps) if s == 0]
        if zero:
            out[:, zero] = ct.data.reshape(2, 1, lq, n)
        return Ciphertext(out, ct.level, ct.scale, "coefficient", "standard", ct.modulus_basis)

    # convenience: HMult as the reference's examples compose it (README.md:66-69)
    def hmult(self, a: Ciphertext, b: Ciphertext) -> Ciphertext:
        """rescale(relinearize(multiply(NTT(a), NTT(b)))).  With the fused
        key switch and fast NTT arithmetic the first three steps run as one
        fused tensor product (fh_ckks_hmult_tensor_coefficient) feeding the
        key switch; the result is bit-identical to the composition."""

        from fhelium_amd.ntt import MODE_FAST

        if not (self.fused_keyswitch and self.runtime.ntt._mode == MODE_FAST and self.config.logN >= 12):
            return self.rescale_to_next_level(
                self.relinearize(self.multiply(self.coefficient_domain_to_ntt_domain(a),
                                               self.coefficient_domain_to_ntt_domain(b))))
        for ct in (a, b):
            ct.assert_state(polynomial_domain="coefficient", residue_representation="standard",
                            modulus_basis="Q", components=2)
        self._same_layout(a, b)
        scale = a.scale * b.scale
        if not math.isfinite(scale) or scale <= 0:
            raise InvalidScaleError("product scale is not positive and finite")
        key = self.relinearization_key
        if key is None:
            key = self.generate_relinearization_key()
        next_level = a.level + 1
        if next_level >= self.config.public_level_count:
            raise MaximumLevelError(a.level, self.config.public_level_count)
        # d0, d1 stay in the NTT domain and seed the key switch's accumulators
        # (the result gains exactly d0, d1); relinearization and rescale run
        # in one key switch (ModDown writes the rescaled rows)
        seeds, d2 = self.key_switcher.tensor_product_seeded(a.data, b.data, a.level)
        out = self.key_switcher.key_switch_fused(d2, key, a.level, rescale=True, seeds=seeds)
        dropped = self.config.moduli[self.runtime.row_start(a.level)]
        return Ciphertext(out, next_level, scale / float(dropped), "coefficient", "standard", "Q")
