"""Host tables of the fused key switch (CPU only).

The decompose kernel computes the mixed-radix (Garner) digits of binary64
digits from a per-digit standard-form record (``DigitTables.garner_f64``).
These tests pin that record: its constants are the Montgomery tables without
their R factor, and the kernel's binary64 recurrence -- emulated here with
correctly rounded binary64 operations -- reproduces the exact mixed-radix
digits of random and boundary integers for the preset's digits.
"""

from __future__ import annotations

import math
import random
from fractions import Fraction

import pytest

from fhelium_amd import rns
from fhelium_amd.config import CkksConfig, Preset

MAGIC = 1.5 * 2.0**52


def fma(a: float, b: float, c: float) -> float:
    """a * b + c with one rounding (Fraction arithmetic is exact)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def mulmod_signed(w: float, x: float, q: float, qinv: float) -> float:
    """The device's f64_mulmod_signed: w x - round(w x / q) q, exactly."""
    h = w * x
    lo = fma(w, x, -h)
    t = fma(h, qinv, MAGIC) - MAGIC
    return fma(-t, q, h) + lo


def garner_f64(record: list[float], residues: list[int]) -> list[int]:
    """The decompose kernel's binary64 Garner recurrence on one coefficient."""
    r = len(residues)
    norm = record[: r - 1]
    prop = record[r - 1: r - 1 + (r - 1) * r]
    qs = record[r - 1 + (r - 1) * r: r - 1 + (r - 1) * r + r]
    qinv = record[-r:]
    first = float(residues[0])
    first = first if first < qs[0] else first - qs[0]
    acc = [first] * r
    digits = [first]
    for row in range(1, r):
        d = mulmod_signed(norm[row - 1], float(residues[row]) - acc[row], qs[row], qinv[row])
        d = d + qs[row] if d < 0.0 else d
        digits.append(d)
        for t in range(row + 1, r):
            acc[t] += mulmod_signed(prop[(row - 1) * r + t], d, qs[t], qinv[t])
    assert all(float(int(d)) == d for d in digits)
    return [int(d) for d in digits]


def mixed_radix(x: int, moduli: list[int]) -> list[int]:
    out = []
    for q in moduli:
        out.append(x % q)
        x //= q
    return out


@pytest.fixture(scope="module")
def preset():
    cfg = CkksConfig.from_preset(Preset.slots32768_scale40_levels34_int64)
    return cfg, rns.RnsLayout(cfg)


def test_garner_record_is_the_standard_form_of_the_montgomery_tables(preset):
    cfg, layout = preset
    qp0 = layout.prime_ids(0, include_p=True)
    seen = 0
    for level in range(cfg.public_level_count):
        for spec in layout.digit_specs(level):
            d = len(spec.prime_ids)
            t = rns._digit_tables(cfg.moduli, spec.prime_ids, qp0)
            src = [cfg.moduli[i] for i in spec.prime_ids]
            if d == 1 or not all(q < rns.F64_DIGIT_LIMIT for q in src):
                assert t.garner_f64.numel() == 0
                continue
            rec = t.garner_f64.tolist()
            assert len(rec) == (d - 1) + (d - 1) * d + 2 * d
            for c in range(d - 1):
                q = src[c + 1]
                assert int(rec[c]) == t.normalizers[c].item() * pow(rns.R, -1, q) % q
                for tgt in range(c + 2, d):
                    q = src[tgt]
                    want = t.propagation[c, tgt].item() * pow(rns.R, -1, q) % q
                    assert int(rec[(d - 1) + c * d + tgt]) == want
            assert rec[(d - 1) + (d - 1) * d:][:d] == [float(q) for q in src]
            seen += 1
    assert seen > 0


def test_garner_record_reproduces_mixed_radix_digits(preset):
    cfg, layout = preset
    qp0 = layout.prime_ids(0, include_p=True)
    rng = random.Random(20260807)
    specs = {spec.prime_ids for lvl in (0, 20, 31) for spec in layout.digit_specs(lvl) if len(spec.prime_ids) > 1}
    assert specs
    for ids in sorted(specs)[:4]:
        t = rns._digit_tables(cfg.moduli, ids, qp0)
        src = [cfg.moduli[i] for i in ids]
        record = t.garner_f64.tolist()
        top = math.prod(src)
        values = [0, 1, top - 1, top // 2, src[0] - 1, src[0]] + [rng.randrange(top) for _ in range(120)]
        for x in values:
            assert garner_f64(record, [x % q for q in src]) == mixed_radix(x, src), (ids, x)
