"""CKKS parameter presets and the ordered QP modulus chain.

Mirrors the reference's ``Preset``/``CkksConfig`` (fhelium/config/ckks.py:26-158
for the preset table, :379-535 for modulus selection): the chain is
``[scale primes q_0..q_{L-1}, structural base q_L, P primes p_0..p_{k-1}]``
with the scale primes taken from the ``scale_bits`` catalog sequence and the
base and P primes from the 60-bit message sequence, both for ring dimension N.
Only the int64 (62-bit Montgomery radix) presets with logN <= 16 are offered
here; the catalog data is the reference's own (``primes_v1.json``).
"""

from __future__ import annotations

import enum
import functools
import hashlib
import json
import math
import pathlib
from dataclasses import dataclass, field

import torch

_PRIMES = pathlib.Path(__file__).with_name("primes_v1.json")


@functools.lru_cache(maxsize=1)
def _catalog() -> dict:
    return json.loads(_PRIMES.read_text())


def scale_primes(scale_bits: int, ring_dimension: int) -> list[int]:
    try:
        return list(_catalog()["scale"][f"{scale_bits}:{ring_dimension}"])
    except KeyError as exc:
        raise KeyError(f"no scale primes for {scale_bits} bits at N={ring_dimension}") from exc


def message_primes(ring_dimension: int) -> list[int]:
    try:
        return list(_catalog()["message"][f"60:{ring_dimension}"])
    except KeyError as exc:
        raise KeyError(f"no 60-bit message primes at N={ring_dimension}") from exc


class Preset(enum.Enum):
    """Supported reference presets (fhelium/config/ckks.py:40-47, int64 only)."""

    slots8192_scale30_levels9_int64 = "slots8192-scale30-levels9-int64"
    slots8192_scale40_levels7_int64 = "slots8192-scale40-levels7-int64"
    slots8192_scale50_levels5_int64 = "slots8192-scale50-levels5-int64"
    slots16384_scale30_levels21_int64 = "slots16384-scale30-levels21-int64"
    slots16384_scale40_levels16_int64 = "slots16384-scale40-levels16-int64"
    slots16384_scale50_levels12_int64 = "slots16384-scale50-levels12-int64"
    slots32768_scale30_levels45_int64 = "slots32768-scale30-levels45-int64"
    slots32768_scale40_levels34_int64 = "slots32768-scale40-levels34-int64"
    slots32768_scale50_levels27_int64 = "slots32768-scale50-levels27-int64"


_PRESETS: dict[Preset, tuple[int, int, int, int]] = {
    # logN, scale_bits, num_scale_primes, num_p_primes (ckks.py:60-106)
    Preset.slots8192_scale30_levels9_int64: (14, 30, 9, 1),
    Preset.slots8192_scale40_levels7_int64: (14, 40, 7, 1),
    Preset.slots8192_scale50_levels5_int64: (14, 50, 5, 1),
    Preset.slots16384_scale30_levels21_int64: (15, 30, 21, 2),
    Preset.slots16384_scale40_levels16_int64: (15, 40, 16, 2),
    Preset.slots16384_scale50_levels12_int64: (15, 50, 12, 2),
    Preset.slots32768_scale30_levels45_int64: (16, 30, 45, 4),
    Preset.slots32768_scale40_levels34_int64: (16, 40, 34, 4),
    Preset.slots32768_scale50_levels27_int64: (16, 50, 27, 4),
}


@dataclass(frozen=True)
class CkksConfig:
    """One CKKS parameter set.

    ``ring_log`` may be lowered below the preset's logN (``with_ring_log``)
    for small parity tests: the catalog primes satisfy q = 1 mod 2N for the
    preset's N and therefore for every smaller power of two.
    """

    logN: int
    scale_bits: int
    num_scale_primes: int
    num_p_primes: int
    catalog_ring_dimension: int
    buffer_bit_length: int = 62
    sigma: float = 3.19
    moduli: tuple[int, ...] = field(init=False)

    def __post_init__(self) -> None:
        scale = scale_primes(self.scale_bits, self.catalog_ring_dimension)[: self.num_scale_primes]
        message = message_primes(self.catalog_ring_dimension)[: 1 + self.num_p_primes]
        if len(scale) != self.num_scale_primes or len(message) != 1 + self.num_p_primes:
            raise ValueError("prime catalog cannot supply the requested chain")
        moduli = tuple(scale) + tuple(message)
        if len(set(moduli)) != len(moduli):
            raise ValueError("duplicate primes in the modulus chain")
        radix = 1 << self.buffer_bit_length
        for q in moduli:
            if 4 * q >= radix or q % 2 == 0 or (q - 1) % (2 << self.logN):
                raise ValueError(f"modulus {q} unusable at logN={self.logN}")
        object.__setattr__(self, "moduli", moduli)

    @classmethod
    def from_preset(cls, preset: Preset | str) -> "CkksConfig":
        preset = Preset(preset) if not isinstance(preset, Preset) else preset
        log_n, bits, scale_count, p_count = _PRESETS[preset]
        return cls(log_n, bits, scale_count, p_count, 1 << log_n)

    def with_ring_log(self, log_n: int) -> "CkksConfig":
        if log_n > math.log2(self.catalog_ring_dimension):
            raise ValueError("ring dimension cannot exceed the catalog's")
        return CkksConfig(log_n, self.scale_bits, self.num_scale_primes, self.num_p_primes,
                          self.catalog_ring_dimension, self.buffer_bit_length, self.sigma)

    # -- derived sizes (ckks.py:549-570) ------------------------------------
    @property
    def N(self) -> int:
        return 1 << self.logN

    @property
    def slots(self) -> int:
        return self.N // 2

    @property
    def num_q_primes(self) -> int:
        return self.num_scale_primes + 1

    @property
    def total_num_primes(self) -> int:
        return self.num_q_primes + self.num_p_primes

    @property
    def q_moduli(self) -> tuple[int, ...]:
        return self.moduli[: self.num_q_primes]

    @property
    def p_moduli(self) -> tuple[int, ...]:
        return self.moduli[self.num_q_primes :]

    @property
    def public_level_count(self) -> int:
        return self.num_scale_primes

    @property
    def default_scale(self) -> float:
        return float(1 << self.scale_bits)

    @property
    def context_id(self) -> str:
        """SHA-256 fingerprint of the mathematical context, computed from the
        same payload text as the reference (fhelium/core/context.py:72-84),
        so value files name the same context on both backends."""

        payload = (
            "representation=direct_per_value_scale_v1;"
            f"logN={self.logN};default_scale={self.default_scale!r};"
            f"Q={','.join(str(q) for q in self.q_moduli)};"
            f"P={','.join(str(p) for p in self.p_moduli)};galois_generator=3"
        )
        return hashlib.sha256(payload.encode("utf-8")).hexdigest()

    @property
    def torch_dtype(self) -> torch.dtype:
        return torch.int64
