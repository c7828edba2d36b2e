"""Binary64 CKKS slot codec and small-basis CRT decode (host side).

Slot order: slot ``j`` is the evaluation at ``zeta^{e_j}``, ``zeta = e^{i pi / N}``
and ``e_j = 3^{-j} mod 2N``, so the Galois map ``X -> X^{3^r}`` rotates slots as
``torch.roll(m, r)`` -- the reference's generator-3 convention
(fhelium/engine/galois.py:10-26, slot_embedding.py).  Encoding rounds to the
nearest integer (the reference rounds stochastically; both are unbiased to
within the decode tolerance).  Encode/decode are outside the evaluated hot
path and run on the host in binary64.
"""

from __future__ import annotations

import functools

import numpy as np


@functools.lru_cache(maxsize=8)
def _slot_positions(n: int) -> np.ndarray:
    """k-index of slot j in the odd-exponent evaluation vector."""

    two_n = 2 * n
    g_inv = pow(3, -1, two_n)
    e = 1
    pos = np.empty(n // 2, dtype=np.int64)
    for j in range(n // 2):
        pos[j] = (e - 1) // 2
        e = e * g_inv % two_n
    return pos


@functools.lru_cache(maxsize=8)
def _twist(n: int) -> np.ndarray:
    return np.exp(1j * np.pi * np.arange(n) / n)


def embed_inverse(slots: np.ndarray, n: int) -> np.ndarray:
    """[*, <=N/2] complex/real slots -> [*, N] real coefficients (unscaled)."""

    slots = np.asarray(slots)
    s = n // 2
    batch = slots.shape[:-1]
    m = np.zeros(batch + (s,), dtype=np.complex128)
    m[..., : slots.shape[-1]] = slots
    pos = _slot_positions(n)
    z = np.zeros(batch + (n,), dtype=np.complex128)
    z[..., pos] = m
    z[..., n - 1 - pos] = np.conj(m)
    a = np.fft.fft(z, axis=-1) / n * np.conj(_twist(n))
    return a.real


def embed(coefficients: np.ndarray) -> np.ndarray:
    """[*, N] real coefficients -> [*, N/2] complex slots (unscaled)."""

    a = np.asarray(coefficients, dtype=np.float64)
    n = a.shape[-1]
    z = np.fft.ifft(a * _twist(n), axis=-1) * n
    return z[..., _slot_positions(n)]


def encode(slots: np.ndarray, n: int, scale: float) -> np.ndarray:
    """Slots -> rounded int64 coefficients at ``scale``."""

    coeffs = np.rint(embed_inverse(slots, n) * scale)
    if np.abs(coeffs).max(initial=0) >= 2.0**62:
        raise ValueError("encoded coefficients exceed the int64 residue range")
    return coeffs.astype(np.int64)


def crt_centered(residues: np.ndarray, moduli: list[int]) -> np.ndarray:
    """[k, *, N] canonical residues over ``moduli`` -> centered values as float64.

    Exact CRT over the k primes with Python integers; valid when the true
    value's magnitude is below half the product of the primes.
    """

    residues = np.asarray(residues)
    big_m = 1
    for q in moduli:
        big_m *= q
    acc = np.zeros(residues.shape[1:], dtype=object)
    for r, q in zip(residues, moduli):
        mi = big_m // q
        coef = mi * pow(mi, -1, q)
        acc = acc + r.astype(object) * coef
    acc = acc % big_m
    half = big_m // 2
    acc = np.where(acc > half, acc - big_m, acc)
    return acc.astype(np.float64)
