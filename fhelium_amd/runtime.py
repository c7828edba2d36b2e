"""Device-resident RNS runtime for one CKKS config on one ROCm device.

Holds every table of the hot path in HBM (parameter rows, NTT tables, ModUp
tables per hybrid digit, rescale / ModDown inverses) and exposes the
level-aware operations the engine composes, in the spirit of the reference's
``RnsRuntime`` (fhelium/engine/rns/runtime.py:24-975): an operand's limbs
are always a contiguous interval of canonical QP rows, and the interval is
passed explicitly rather than inferred from the limb count.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from fhelium_amd import kernels as K
from fhelium_amd import ntt as gntt
from fhelium_amd import rns
from fhelium_amd.config import CkksConfig


@dataclass
class DeviceDigitTables:
    prime_ids: tuple[int, ...]
    normalizers: torch.Tensor
    propagation: torch.Tensor
    modulus_lo: torch.Tensor
    modulus_hi: torch.Tensor
    neg_inv_lo: torch.Tensor
    neg_inv_hi: torch.Tensor
    basis_extension: torch.Tensor  # [D-1, QP0]
    basis_extension_std: torch.Tensor  # [D-1, QP0]
    basis_extension_shoup: torch.Tensor  # [D-1, QP0]
    garner_f64: torch.Tensor  # binary64 Garner record or empty

    @classmethod
    def from_host(cls, key, t, dev) -> "DeviceDigitTables":
        return cls(key, t.normalizers.to(dev), t.propagation.to(dev).contiguous(), t.modulus_lo.to(dev),
                   t.modulus_hi.to(dev), t.neg_inv_lo.to(dev), t.neg_inv_hi.to(dev),
                   t.basis_extension.to(dev).contiguous(), t.basis_extension_std.to(dev).contiguous(),
                   t.basis_extension_shoup.to(dev).contiguous(), t.garner_f64.to(dev))


class RnsRuntime:
    def __init__(self, config: CkksConfig, device: torch.device | str = "cuda",
                 ntt_mode: str = "fast", ntt_tuning=None) -> None:
        self.config = config
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("the gfx950 runtime needs a ROCm device")
        self.host = rns.HostTables(config)
        self.layout = self.host.layout
        dev = self.device
        self.rns_params = self.host.rns_params.to(dev)
        self.twice_modulus = self.rns_params[0].contiguous()
        fwd, inv = rns.compact_twiddles_montgomery(config.moduli, config.logN)
        self.forward_twiddles = fwd.to(dev)
        self.inverse_twiddles = inv.to(dev)
        self.ntt_tables = gntt.NttTables(self.rns_params, self.forward_twiddles,
                                         self.inverse_twiddles, config.logN)
        # ntt_tuning: an autotune.TuningCache whose per-shape schedules the
        # NTT backend follows (None: one launch per transform)
        self.ntt = gntt.Gfx950NttBackend(self.ntt_tables, ntt_mode, moduli=tuple(config.moduli),
                                         tuning=ntt_tuning)
        self.digit_tables = {
            key: DeviceDigitTables.from_host(key, t, dev) for key, t in self.host.digit_tables.items()
        }
        self.rescale_inverses = [v.to(dev) for v in self.host.rescale_inverses]
        self.moddown_inverses = [v.to(dev) for v in self.host.moddown_inverses]
        self.p_product_montgomery_q = self.host.p_product_montgomery_q.to(dev)
        self.p_digit = DeviceDigitTables.from_host(self.layout.p_prime_ids, self.host.p_digit, dev)
        self.p_inverse_q = self.host.p_inverse_q.to(dev)
        self.seed_scale = self.host.seed_scale.to(dev)
        self.moddown_binary64 = self.host.moddown_binary64.to(dev).contiguous()
        self._galois_cache: dict[int, tuple[torch.Tensor, torch.Tensor]] = {}
        self._ntt_galois_cache: dict[int, torch.Tensor] = {}

    # -- row intervals ----------------------------------------------------
    def row_start(self, level: int) -> int:
        return self.layout.start_row(level)

    def params(self, level: int, include_p: bool = False) -> torch.Tensor:
        start = self.row_start(level)
        stop = self.config.total_num_primes if include_p else self.config.num_q_primes
        return self.rns_params[:, start:stop]

    def params_rows(self, start: int, stop: int) -> torch.Tensor:
        return self.rns_params[:, start:stop]

    # -- NTT ----------------------------------------------------------------
    def forward_to_montgomery_(self, x: torch.Tensor, level: int) -> None:
        self.ntt.forward_to_montgomery_(x, self.row_start(level))

    def forward_montgomery_(self, x: torch.Tensor, level: int) -> None:
        self.ntt.forward_montgomery_(x, self.row_start(level))

    def inverse_to_standard_(self, x: torch.Tensor, level: int) -> None:
        self.ntt.inverse_to_standard_(x, self.row_start(level))

    # -- Galois tables ------------------------------------------------------
    def coefficient_galois(self, galois_element: int) -> tuple[torch.Tensor, torch.Tensor]:
        hit = self._galois_cache.get(galois_element)
        if hit is None:
            idx, sgn = rns.coefficient_galois_tables(self.config.N, galois_element)
            hit = (idx.to(self.device), sgn.to(self.device))
            self._galois_cache[galois_element] = hit
        return hit

    def ntt_galois(self, galois_element: int) -> torch.Tensor:
        hit = self._ntt_galois_cache.get(galois_element)
        if hit is None:
            hit = rns.ntt_galois_indices(self.config.logN, galois_element).to(self.device)
            self._ntt_galois_cache[galois_element] = hit
        return hit

    # -- arithmetic shortcuts ---------------------------------------------
    def montgomery_mul(self, a, b, level, include_p=False):
        return K.montgomery_mul(a, b, self.params(level, include_p))

    def add_lazy(self, a, b, level, include_p=False):
        return K.add_lazy(a, b, self.params(level, include_p))

    def add_canonical(self, a, b, level, include_p=False):
        return K.add_canonical(a, b, self.params(level, include_p))

    def canonicalize_(self, x, level, include_p=False):
        K.canonicalize_residues_(x, self.params(level, include_p))
