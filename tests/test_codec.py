"""Host codec: encode/decode round trip, generator-3 rotation convention and
small-basis CRT decode (CPU only)."""

from __future__ import annotations

import numpy as np
import pytest

from fhelium_amd import codec, rns


@pytest.mark.parametrize("n", [16, 1024, 1 << 14])
def test_round_trip(n):
    rng = np.random.default_rng(n)
    m = rng.uniform(-1, 1, n // 2) + 1j * rng.uniform(-1, 1, n // 2)
    a = codec.embed_inverse(m, n)
    np.testing.assert_allclose(codec.embed(a), m, atol=1e-9)


@pytest.mark.parametrize("step", [1, 3, -2, 100])
def test_galois_rotates_slots_like_roll(step):
    n = 256
    rng = np.random.default_rng(step + 1000)
    m = rng.uniform(-1, 1, n // 2)
    a = codec.encode(m, n, 2.0**30)
    g = rns.rotation_galois_element(n, step)
    idx, sgn = rns.coefficient_galois_tables(n, g)
    rotated = a[idx.numpy()] * sgn.numpy().astype(np.int64)
    # brute-force X -> X^g on the integer polynomial
    brute = np.zeros(n, dtype=np.int64)
    for j, c in enumerate(a):
        e = (j * g) % (2 * n)
        if e < n:
            brute[e] += c
        else:
            brute[e - n] -= c
    np.testing.assert_array_equal(rotated, brute)
    np.testing.assert_allclose(codec.embed(rotated).real / 2.0**30, np.roll(m, step), atol=1e-6)


def test_crt_centered():
    moduli = [1099510054913, 1152921504606584833]
    rng = np.random.default_rng(3)
    vals = rng.integers(-(1 << 62), 1 << 62, size=64)
    res = np.stack([np.mod(vals, q) for q in moduli])
    np.testing.assert_array_equal(codec.crt_centered(res, moduli), vals.astype(np.float64))
