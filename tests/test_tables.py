"""Pins of the engine's scalar and index tables by their defining properties.

Each table is checked by what it must do, computed independently of how the
table builders compute it:

* rescale inverses (ckks_engine.py:2952-2989): x * t, reduced by the
  oracle's REDC (pinned by the golden REDC vectors), equals x / q_drop mod q;
* ModDown P-drop inverses (ckks_engine.py:2991-3048): row d drops
  p_{k-1-d}, and REDC(x * t) equals x / p mod q on every surviving row;
* P R mod q (ckks_engine.py:3050-3072): REDC(x * t) equals x P mod q;
* the fused key switch's seed scales: adding scale * seed to an accumulator
  adds P * seed in the accumulator's own scaling (plain products on binary64
  rows, one REDC on integer rows);
* the coefficient Galois tables (galois.py:50-83): gathering with them equals
  the substitution X -> X^g computed by polynomial algebra modulo X^N + 1;
* the NTT-domain Galois indices (hybrid_keyswitch.py:506-528): gathering
  evaluations with them equals evaluating the substituted polynomial at the
  NTT points psi^(2 br(k) + 1), computed by direct evaluation.
"""

from __future__ import annotations

import math
import random

import pytest

from fhelium_amd import rns
from oracle import oracle as O
from tests import helpers as H

R = 1 << 62


def _redc(a: int, b: int, q: int) -> int:
    k = (-pow(q, -1, R)) % R
    return O.mont_mul_scalar(a, b, q, k) % q


@pytest.fixture(scope="module")
def tables():
    return H.host_tables(12)


def test_rescale_inverses_divide_by_the_dropped_prime(tables):
    rnd = random.Random(1)
    moduli = tables.moduli
    for level in range(0, tables.config.public_level_count, 5):
        ids = tables.layout.prime_ids(level, include_p=True)
        drop = moduli[ids[0]]
        inv = tables.rescale_inverses[level].tolist()
        assert len(inv) == len(ids) - 1
        for t, i in zip(inv, ids[1:]):
            q = moduli[i]
            x = rnd.randrange(q)
            assert (_redc(x, t, q) * drop - x) % q == 0, (level, i)


def test_moddown_inverses_drop_p_in_reverse_order(tables):
    cfg = tables.config
    moduli = tables.moduli
    k = cfg.num_p_primes
    qp0 = tables.layout.prime_ids(0, include_p=True)
    rnd = random.Random(2)
    for level in (0, 7, cfg.public_level_count - 1):
        packed = tables.moddown_inverses[level]
        start = tables.layout.start_row(level)
        assert packed.shape[0] == k
        for d in range(k):
            p = cfg.p_moduli[k - 1 - d]  # row d drops the last remaining P prime
            survivors = qp0[start: len(qp0) - d - 1]
            row = packed[d].tolist()
            for c, i in enumerate(survivors):
                q = moduli[i]
                x = rnd.randrange(q)
                assert (_redc(x, row[c], q) * p - x) % q == 0, (level, d, c)
            assert all(v == 0 for v in row[len(survivors):])


def test_p_product_montgomery_multiplies_by_p(tables):
    p_prod = math.prod(tables.config.p_moduli)
    rnd = random.Random(3)
    for t, i in zip(tables.p_product_montgomery_q.tolist(), tables.layout.prime_ids(0)):
        q = tables.moduli[i]
        x = rnd.randrange(q)
        assert _redc(x, t, q) == x * p_prod % q


def test_seed_scale_adds_p_times_the_seed(tables):
    p_prod = math.prod(tables.config.p_moduli)
    rnd = random.Random(4)
    for t, i in zip(tables.seed_scale.tolist(), tables.layout.prime_ids(0)):
        q = tables.moduli[i]
        seed = rnd.randrange(q)
        if q < rns.BINARY64_PRIME_LIMIT:
            added = t * seed % q  # binary64 rows: plain residue products
        else:
            added = _redc(seed, t, q)  # integer rows: one REDC per sum
        assert added == p_prod * seed % q, i
        assert 0 <= t < q


def _poly_mul(a, b, n):
    """Product modulo X^n + 1 over the integers (schoolbook)."""
    out = [0] * n
    for i, x in enumerate(a):
        if x == 0:
            continue
        for j, y in enumerate(b):
            k = i + j
            if k < n:
                out[k] += x * y
            else:
                out[k - n] -= x * y
    return out


def _substitute(a, g, n):
    """a(X^g) modulo X^n + 1 by Horner's rule on Y = X^g (computed as a
    polynomial power, not by an index formula)."""
    y = [1] + [0] * (n - 1)
    x = [0, 1] + [0] * (n - 2)
    for _ in range(g):
        y = _poly_mul(y, x, n)
    out = [0] * n
    for c in reversed(a):
        out = _poly_mul(out, y, n)
        out[0] += c
    return out


@pytest.mark.parametrize("n", [8, 16])
def test_coefficient_galois_tables_match_substitution(n):
    rnd = random.Random(n)
    for g in (3, 5, 2 * n - 1, pow(3, 5, 2 * n)):
        idx, sgn = rns.coefficient_galois_tables(n, g)
        a = [rnd.randrange(-50, 50) for _ in range(n)]
        want = _substitute(a, g, n)
        got = [int(s) * a[int(i)] for i, s in zip(idx.tolist(), sgn.tolist())]
        assert got == want, g


def test_ntt_galois_indices_match_evaluation():
    n, q = 16, 97  # 97 = 1 mod 32: primitive 32nd roots exist
    log_n = 4
    psi = next(r for r in (pow(c, (q - 1) // (2 * n), q) for c in range(2, q)) if pow(r, n, q) != 1)
    points = [pow(psi, 2 * rns.bit_reverse_table(log_n)[k] + 1, q) for k in range(n)]
    rnd = random.Random(7)
    a = [rnd.randrange(q) for _ in range(n)]

    def evaluate(poly):
        return [sum(c * pow(x, e, q) for e, c in enumerate(poly)) % q for x in points]

    spectrum = evaluate(a)
    for g in (3, 9, 2 * n - 1, 27):
        src = rns.ntt_galois_indices(log_n, g).tolist()
        want = evaluate([c % q for c in _substitute(a, g, n)])
        assert [spectrum[s] for s in src] == want, g
