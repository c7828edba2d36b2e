"""CKKS value types and the errors raised by the hot-path methods.

Mirrors the state model of the reference's ``Ciphertext`` and key classes
(fhelium/core/ciphertext.py:26-384, fhelium/core/keys.py:222-395): a value is
a dense int64 tensor ``[component, *batch, limb, N]`` plus level, scale,
polynomial domain ("coefficient" | "ntt"), residue representation
("standard" | "montgomery") and modulus basis ("Q" | "QP").  Limb ``j`` of a
level-``l`` Q value is prime row ``l + j``.
"""

from __future__ import annotations

import math
import operator
from dataclasses import dataclass, field, replace

import torch


class CkksStateError(ValueError):
    """A value's domain / representation / component count is not accepted."""


class MaximumLevelError(ValueError):
    """Rescale or mod switch requested past the last public level (errors.py,
    ckks_rescale.py:166-173)."""

    def __init__(self, level: int, maximum_level: int) -> None:
        super().__init__(f"cannot move past level {level}: the last public level is {maximum_level - 1}")
        self.level = level
        self.maximum_level = maximum_level


class InvalidScaleError(ValueError):
    pass


def _check_scale(scale, value_name: str) -> float:
    """A positive finite binary64 scale (core/scale.py:11-29)."""

    if isinstance(scale, (bool, str, bytes)):
        raise InvalidScaleError(f"{value_name} scale must be a positive finite number, got {scale!r}")
    try:
        value = float(scale)
    except (TypeError, ValueError, OverflowError) as error:
        raise InvalidScaleError(f"{value_name} scale must be a positive finite number, got {scale!r}") from error
    if not math.isfinite(value) or value <= 0.0:
        raise InvalidScaleError(f"{value_name} scale must be a positive finite number, got {scale!r}")
    return value


def _check_payload(data, value_name: str, min_dims: int) -> None:
    """Dense integral tensor (core/validation.py:62-77) with non-empty limb
    and coefficient axes and nonzero batch extents."""

    if not isinstance(data, torch.Tensor):
        raise TypeError(f"{value_name} data must be a torch.Tensor")
    if data.layout != torch.strided:
        raise TypeError(f"{value_name} data must use dense strided storage")
    if data.dtype == torch.bool or data.is_floating_point() or data.is_complex():
        raise TypeError(f"{value_name} data must use an integral scalar dtype")
    if data.dim() < min_dims:
        raise ValueError(f"{value_name} data has too few axes: shape {tuple(data.shape)}")
    if data.size(-2) == 0 or data.size(-1) == 0:
        raise ValueError(f"{value_name} limb and coefficient axes cannot be empty")


def _check_level(level, value_name: str) -> int:
    if isinstance(level, bool):
        raise ValueError(f"{value_name} level must be a non-negative integer, got {level!r}")
    try:
        value = operator.index(level)
    except TypeError as error:
        raise ValueError(f"{value_name} level must be a non-negative integer, got {level!r}") from error
    if value < 0:
        raise ValueError(f"{value_name} level must be a non-negative integer, got {level!r}")
    return value


@dataclass
class Ciphertext:
    """``data`` is [component, *batch, limb, coeff]; construction applies the
    structural checks of the reference's Ciphertext
    (fhelium/core/ciphertext.py:67-127): 2 or 3 components, non-empty limb
    and coefficient axes, nonzero batch dimensions, a known domain, basis
    and representation (coefficient values standard, NTT values Montgomery),
    a non-negative level and a positive finite scale."""

    data: torch.Tensor
    level: int
    scale: float
    polynomial_domain: str = "coefficient"
    residue_representation: str = "standard"
    modulus_basis: str = "Q"
    context_id: int = 0

    def __post_init__(self) -> None:
        self.scale = _check_scale(self.scale, "Ciphertext")
        self.level = _check_level(self.level, "Ciphertext")
        _check_payload(self.data, "Ciphertext", 3)
        if self.data.size(0) not in (2, 3):
            raise ValueError(f"Ciphertext component axis must have size 2 or 3, got {self.data.size(0)}")
        if any(extent == 0 for extent in self.data.shape[1:-2]):
            raise ValueError("Ciphertext batch dimensions must be nonzero")
        if self.polynomial_domain not in ("coefficient", "ntt"):
            raise ValueError(f"Unsupported ciphertext polynomial_domain: {self.polynomial_domain!r}")
        if self.modulus_basis not in ("Q", "QP"):
            raise ValueError(f"Unsupported ciphertext modulus_basis: {self.modulus_basis!r}")
        if self.residue_representation not in ("standard", "montgomery"):
            raise ValueError(f"Unsupported ciphertext residue representation: {self.residue_representation!r}")
        if self.polynomial_domain == "coefficient" and self.residue_representation != "standard":
            raise ValueError("a coefficient-domain Ciphertext holds standard residues")
        if self.polynomial_domain == "ntt" and self.residue_representation != "montgomery":
            raise ValueError("an NTT-domain Ciphertext holds Montgomery residues")

    @property
    def component_count(self) -> int:
        return self.data.size(0)

    def component(self, i: int) -> torch.Tensor:
        return self.data[i]

    @property
    def c0(self) -> torch.Tensor:
        return self.data[0]

    @property
    def c1(self) -> torch.Tensor:
        return self.data[1]

    @property
    def c2(self) -> torch.Tensor:
        return self.data[2]

    @property
    def batch_shape(self) -> tuple[int, ...]:
        return tuple(self.data.shape[1:-2])

    @property
    def limb_count(self) -> int:
        return self.data.size(-2)

    @property
    def includes_p(self) -> bool:
        return self.modulus_basis == "QP"

    def prime_ids(self, num_q: int, num_p: int) -> tuple[int, ...]:
        ids = tuple(range(self.level, num_q))
        return ids + tuple(range(num_q, num_q + num_p)) if self.includes_p else ids

    def clone(self) -> "Ciphertext":
        return replace(self, data=self.data.clone())

    def assert_state(self, *, polynomial_domain: str | None = None,
                     residue_representation: str | None = None, components: int | None = None,
                     modulus_basis: str | None = None) -> None:
        if polynomial_domain is not None and self.polynomial_domain != polynomial_domain:
            raise CkksStateError(f"expected a {polynomial_domain}-domain value, got {self.polynomial_domain}")
        if residue_representation is not None and self.residue_representation != residue_representation:
            raise CkksStateError(
                f"expected {residue_representation} residues, got {self.residue_representation}")
        if components is not None and self.component_count != components:
            raise CkksStateError(f"expected {components} components, got {self.component_count}")
        if modulus_basis is not None and self.modulus_basis != modulus_basis:
            raise CkksStateError(f"expected basis {modulus_basis}, got {self.modulus_basis}")


class ScaleMismatchError(ValueError):
    """Operands whose binary64 scales are not exactly equal (errors.py)."""


@dataclass
class Plaintext:
    """Operation-ready RNS plaintext ``[*batch, limb, N]`` in Montgomery form,
    coefficient or NTT domain (fhelium/core/plaintext.py)."""

    data: torch.Tensor
    level: int
    scale: float
    polynomial_domain: str = "coefficient"
    residue_representation: str = "montgomery"
    modulus_basis: str = "Q"

    def __post_init__(self) -> None:
        # the structural checks the reference applies to its values
        # (fhelium/core/ciphertext.py:67-100, core/scale.py:11-29)
        self.scale = _check_scale(self.scale, "Plaintext")
        self.level = _check_level(self.level, "Plaintext")
        _check_payload(self.data, "Plaintext", 2)
        if any(extent == 0 for extent in self.data.shape[:-2]):
            raise ValueError("Plaintext batch dimensions must be nonzero")
        if self.polynomial_domain not in ("coefficient", "ntt"):
            raise ValueError(f"Unsupported plaintext polynomial_domain: {self.polynomial_domain!r}")
        if self.modulus_basis not in ("Q", "QP"):
            raise ValueError(f"Unsupported plaintext modulus_basis: {self.modulus_basis!r}")

    @property
    def batch_shape(self) -> tuple[int, ...]:
        return tuple(self.data.shape[:-2])


@dataclass
class SecretKey:
    """s on every QP row, NTT domain, Montgomery form: [QP, N]."""

    data: torch.Tensor
    coefficients: torch.Tensor  # signed ternary coefficients [N] (host copy for tests)


@dataclass
class KeySwitchKey:
    """Hybrid key ``[key_digit, 2, QP, N]`` NTT/Montgomery (keys.py:222-292)."""

    data: torch.Tensor

    def digit(self, key_digit_index: int) -> torch.Tensor:
        return self.data[key_digit_index]


@dataclass
class RelinearizationKey(KeySwitchKey):
    pass


@dataclass
class ConjugationKey(KeySwitchKey):
    """Key switch material from sigma_{-1}(s) back to s (keys.py:386-391):
    after X -> X^{-1} it restores the key relation, conjugating every slot."""


@dataclass
class RotationKey(KeySwitchKey):
    rotation_step: int = 0
    galois_element: int = 1

    @staticmethod
    def canonical_step(rotation_step: int, ring_dimension: int) -> int:
        """Step modulo S = N/2 into [-S/2, S/2) (keys.py:339-351)."""

        slots = ring_dimension // 2
        step = int(rotation_step) % slots
        return step - slots if step >= slots // 2 else step
