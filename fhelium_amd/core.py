"""CKKS value types and the errors raised by the hot-path methods.

Mirrors the state model of the reference's ``Ciphertext`` and key classes
(fhelium/core/ciphertext.py:26-384, fhelium/core/keys.py:222-395): a value is
a dense int64 tensor ``[component, *batch, limb, N]`` plus level, scale,
polynomial domain ("coefficient" | "ntt"), residue representation
("standard" | "montgomery") and modulus basis ("Q" | "QP").  Limb ``j`` of a
level-``l`` Q value is prime row ``l + j``.
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace

import torch


class CkksStateError(ValueError):
    """A value's domain / representation / component count is not accepted."""


class MaximumLevelError(ValueError):
    """Rescale requested past the last public level (errors.py, ckks_rescale.py:166-173)."""

    def __init__(self, level: int, maximum_level: int) -> None:
        super().__init__(f"cannot rescale from level {level}: the last public level is {maximum_level - 1}")
        self.level = level
        self.maximum_level = maximum_level


class InvalidScaleError(ValueError):
    pass


@dataclass
class Ciphertext:
    data: torch.Tensor
    level: int
    scale: float
    polynomial_domain: str = "coefficient"
    residue_representation: str = "standard"
    modulus_basis: str = "Q"
    context_id: int = 0

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
class RotationKey(KeySwitchKey):
    rotation_step: int = 0
    galois_element: int = 1

    @staticmethod
    def canonical_step(rotation_step: int, ring_dimension: int) -> int:
        """Step modulo S = N/2 into [-S/2, S/2) (keys.py:339-351)."""

        slots = ring_dimension // 2
        step = int(rotation_step) % slots
        return step - slots if step >= slots // 2 else step
