"""Generate tests/golden/ckks_golden_v1.json from Python-integer closed forms.

The reference ships no golden-vector files; its own tests compare against
closed forms computed inline (tests/test_ntt_backend.py:110-243 brute-force
and round-trip checks, tests/test_ckks_operation_correctness.py:404-419
``_mixed_radix_integer_reference``, tests/test_cpu_native_strides.py:321-370
``_moddown_reference``).  This script computes the same kind of expected
values with plain Python integers -- no oracle, no product code -- for small
seeded inputs, and records inputs, the reference-format tables and expected
outputs.  Where the reference's output is a lazy representative in [0, 2q)
the fixture stores the canonical value and the test compares modulo q.

Run: python tests/golden/make_golden.py  (writes the JSON next to this file)
"""

from __future__ import annotations

import json
import pathlib
import random

OUT = pathlib.Path(__file__).with_name("ckks_golden_v1.json")
R = 1 << 62
HALF = (1 << 31) - 1

# First primes of the reference catalog for N = 65536 (scale 40-bit and
# message 60-bit sequences, fhelium/config/resources/*_primes_v1.safetensors);
# each is 1 mod 2^17, hence NTT-friendly for every N <= 65536.
Q40 = [1099510054913, 1099515691009, 1099507695617]
Q60 = [1152921504606584833, 1152921504598720513, 1152921504597016577, 1152921504595968001,
       1152921504592429057]


def params_rows(moduli, n):
    """The reference's [8, limb] parameter rows (rns_parameters.h:6-16)."""
    rows = [[] for _ in range(8)]
    for q in moduli:
        k = (-pow(q, -1, R)) % R
        rows[0].append(2 * q)
        rows[1].append(q & HALF)
        rows[2].append(q >> 31)
        rows[3].append(k & HALF)
        rows[4].append(k >> 31)
        rows[5].append(pow(R, 2, q))
        rows[6].append((pow(R, 2, q) << 40) % q)
        rows[7].append(pow(n, -1, q) * R % q)
    return rows


def psi(q, n):
    """First c = 2, 3, ... with c^((q-1)/2N) of order 2N (twiddles.py:10-20)."""
    e = (q - 1) // (2 * n)
    for c in range(2, n):
        r = pow(c, e, q)
        if pow(r, n, q) != 1:
            return r
    raise RuntimeError("no root")


def br(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0


def crt(residues, moduli):
    m = 1
    for q in moduli:
        m *= q
    x = 0
    for r, q in zip(residues, moduli):
        mi = m // q
        x += r * mi * pow(mi, -1, q)
    return x % m, m


def main():
    rnd = random.Random(0x5EED)
    out = {"version": 1, "radix_bits": 62}

    # -- Montgomery multiply: canonical a b R^{-1} mod q for lazy inputs --------
    cases = []
    for q in Q40[:2] + Q60[:2]:
        k = (-pow(q, -1, R)) % R
        a = [rnd.randrange(2 * q) for _ in range(16)] + [0, 2 * q - 1, q, q - 1]
        b = [rnd.randrange(2 * q) for _ in range(16)] + [2 * q - 1, 2 * q - 1, 1, q]
        cases.append({"q": q, "k": k, "a": a, "b": b,
                      "expected": [x * y * pow(R, -1, q) % q for x, y in zip(a, b)]})
    out["mont_mul"] = cases

    # -- negacyclic NTT: A[k] = sum a[n] psi^{(2 br(k)+1) n} ------------------
    ntt = []
    for log_n in (4, 5, 6):
        n = 1 << log_n
        moduli = [Q40[0], Q40[1], Q60[0]]
        fwd_tw, inv_tw, x_rows, fwd_rows, std_rows = [], [], [], [], []
        for q in moduli:
            p = psi(q, n)
            pinv = pow(p, -1, q)
            fwd_tw.append([pow(p, br(i, log_n), q) * R % q for i in range(n)])
            inv_tw.append([pow(pinv, br(i, log_n), q) * R % q for i in range(n)])
            x = [rnd.randrange(q) for _ in range(n)]
            x_rows.append(x)
            fwd_rows.append([sum(x[j] * pow(p, (2 * br(kk, log_n) + 1) * j, q) for j in range(n)) % q
                             for kk in range(n)])
            # standard-form input to forward_to_montgomery: output is (A * R) mod q
            std_rows.append([(v * R) % q for v in fwd_rows[-1]])
        ntt.append({"log_n": log_n, "moduli": moduli, "params": params_rows(moduli, n),
                    "forward_twiddles_montgomery": fwd_tw, "inverse_twiddles_montgomery": inv_tw,
                    "x": x_rows, "forward_canonical": fwd_rows, "forward_to_montgomery_canonical": std_rows})
    out["ntt"] = ntt

    # -- mixed-radix decomposition and basis extension ---------------------
    mr = []
    for digit in ([Q40[0], Q40[1], Q40[2], Q60[0]], [Q40[2], Q60[1]], [Q60[2]]):
        d = len(digit)
        products = [digit[0]]
        for i in range(1, d - 1):
            products.append(products[-1] * digit[i])
        norm = [pow(products[c], -1, digit[c + 1]) * R % digit[c + 1] for c in range(d - 1)]
        prop = [[(products[c] * R) % digit[t] if t >= c + 2 else 0 for t in range(d)] for c in range(d - 1)]
        dest = Q40 + Q60
        basis = [[(products[c] * pow(R, 2, p)) % p for p in dest] for c in range(d - 1)]
        n = 8
        src, digits_expected, ext_expected = [], [], []
        for _ in range(n):
            vals = [rnd.randrange(q) for q in digit]
            x, _ = crt(vals, digit)
            ds, rest = [], x
            for q in digit:
                ds.append(rest % q)
                rest //= q
            src.append(vals)
            digits_expected.append(ds)
            ext_expected.append([(x % p) * R % p for p in dest])
        mr.append({"moduli": digit, "normalizers": norm, "propagation": prop,
                   "lo": [q & HALF for q in digit], "hi": [q >> 31 for q in digit],
                   "klo": [((-pow(q, -1, R)) % R) & HALF for q in digit],
                   "khi": [((-pow(q, -1, R)) % R) >> 31 for q in digit],
                   "destination_moduli": dest, "destination_params": params_rows(dest, 8),
                   "basis_extension": basis,
                   "source": [list(col) for col in zip(*src)],
                   "digits": [list(col) for col in zip(*digits_expected)],
                   "extended_canonical": [list(col) for col in zip(*ext_expected)]})
    out["mixed_radix"] = mr

    # -- ModDown: floor(X / P) mod q_i over canonical QP residues -------------
    qm, pm = Q40, Q60[:3]
    all_m = qm + pm
    inv_rows = []
    for step, p in enumerate(reversed(pm)):
        survivors = all_m[: len(all_m) - step - 1]
        row = [pow(p, -1, m) * R % m for m in survivors]
        inv_rows.append(row + [0] * (len(all_m) - 1 - len(row)))
    n = 8
    qres, pres, exp = [], [], []
    P = pm[0] * pm[1] * pm[2]
    for _ in range(n):
        vals = [rnd.randrange(m) for m in all_m]
        x, _ = crt(vals, all_m)
        qres.append(vals[: len(qm)])
        pres.append(vals[len(qm):])
        exp.append([(x // P) % q for q in qm])
    out["moddown"] = {"q_moduli": qm, "p_moduli": pm, "params": params_rows(all_m, n),
                      "inverses": inv_rows, "q": [list(c) for c in zip(*qres)],
                      "p": [list(c) for c in zip(*pres)], "expected": [list(c) for c in zip(*exp)]}

    # -- rescale: round / floor of X / q_drop over the CRT value ---------------
    chain = [Q40[0], Q40[1], Q40[2], Q60[0]]
    drop = chain[0]
    rem = chain[1:]
    inv = [pow(drop, -1, m) * R % m for m in rem]
    res, near, trunc = [], [], []
    for t in range(n):
        vals = [rnd.randrange(m) for m in chain]
        if t == 0:
            vals[0] = drop // 2          # ties: exactly half rounds down
        if t == 1:
            vals[0] = drop // 2 + 1
        x, _ = crt(vals, chain)
        res.append(vals)
        near.append([(x // drop + (1 if x % drop > drop // 2 else 0)) % m for m in rem])
        trunc.append([(x // drop) % m for m in rem])
    out["rescale"] = {"moduli": chain, "params_remaining": params_rows(rem, n), "inverse": inv,
                      "half": drop // 2, "residues": [list(c) for c in zip(*res)],
                      "nearest": [list(c) for c in zip(*near)], "truncate": [list(c) for c in zip(*trunc)]}

    # -- Galois X -> X^g on integer polynomials -------------------------------
    gal = []
    n = 16
    for step in (1, 3, 7):
        g = pow(3, step % n, 2 * n)
        coeffs = [rnd.randrange(-1000, 1000) for _ in range(n)]
        img = [0] * n
        for j, c in enumerate(coeffs):
            e = j * g % (2 * n)
            if e < n:
                img[e] += c
            else:
                img[e - n] -= c
        idx = [0] * n
        sgn = [0] * n
        for j in range(n):
            e = j * g % (2 * n)
            idx[e % n] = j
            sgn[e % n] = -1 if e >= n else 1
        gal.append({"n": n, "g": g, "coefficients": coeffs, "image": img, "source_indices": idx,
                    "source_sign": sgn})
    out["galois"] = gal

    OUT.write_text(json.dumps(out, separators=(",", ":")) + "\n")
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
