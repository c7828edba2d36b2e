"""numpy front end of the C oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use
this module, and only as the checker.  Functions take and return dense int64
numpy arrays in the reference layouts ([batch, limb, n] residues, [8, limb]
rns_params, compact Montgomery twiddles [limb, n]) and restate the reference
CPU kernels cited in ``fh_oracle.c``.

Pinning (tests/test_oracle.py): every function is checked against the
golden vectors in tests/golden/ (generated from Python-integer closed forms
by tests/golden/make_golden.py) and against closed-form properties:
brute-force negacyclic DFT, exact round trips, Garner reconstruction,
floor(x/P), round(x/q), X -> X^g on integer polynomials.
"""

from __future__ import annotations

import ctypes
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "_build" / "libfh_oracle.so"

_lib = None


def build() -> pathlib.Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
        _lib.or_mont_mul.restype = ctypes.c_int64
        _lib.or_mont_mul.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint64]
    return _lib


_KEEP: list = []


def _p(a: np.ndarray) -> ctypes.c_void_p:
    """Pointer to a C-contiguous array.  The array is kept alive until the
    next oracle call finishes: callers often pass a temporary copy whose only
    reference would otherwise die before the C function runs."""

    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    _KEEP.append(a)
    return ctypes.c_void_p(a.ctypes.data)


class _Call:
    """Context that releases the arrays pinned by _p after a C call."""

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        _KEEP.clear()
        return False


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


_I = ctypes.c_int64
_C = ctypes.c_int


def default_threads() -> int:
    return int(os.environ.get("FH_ORACLE_THREADS", "0"))


def mont_mul_scalar(a: int, b: int, q: int, k: int) -> int:
    return int(lib().or_mont_mul(a, b, q, k))


def ntt_forward(x, twiddles, params, to_montgomery: bool, threads: int | None = None) -> np.ndarray:
    v = _i64(x).copy()
    b, limbs, n = v.shape
    log_n = n.bit_length() - 1
    with _Call():
        lib().or_ntt_forward(_p(v), _I(b), _I(limbs), _I(log_n), _p(_i64(twiddles)), _p(_i64(params)),
                             _C(int(to_montgomery)), _C(default_threads() if threads is None else threads))
    return v


def ntt_inverse(x, twiddles, params, out_form: int, threads: int | None = None) -> np.ndarray:
    v = _i64(x).copy()
    b, limbs, n = v.shape
    log_n = n.bit_length() - 1
    with _Call():
        lib().or_ntt_inverse(_p(v), _I(b), _I(limbs), _I(log_n), _p(_i64(twiddles)), _p(_i64(params)),
                             _C(out_form), _C(default_threads() if threads is None else threads))
    return v


BINARY_OPS = {"mul": 0, "add_lazy": 1, "sub_lazy": 2, "add_canonical": 3, "sub_canonical": 4,
              "mul_cyclic": 5, "mul_contiguous": 6}


def binary(op: str, a, b, params, threads: int | None = None) -> np.ndarray:
    a = _i64(a)
    b = _i64(b)
    out = np.empty_like(a)
    batch, limbs, n = a.shape
    support = b.shape[2]
    with _Call():
        lib().or_binary(_C(BINARY_OPS[op]), _p(out), _p(a), _p(b), _I(batch), _I(b.shape[0]), _I(limbs),
                        _I(n), _p(_i64(params)), _I(support),
                        _C(default_threads() if threads is None else threads))
    return out


UNARY_OPS = {"to_montgomery": 0, "from_montgomery": 1, "canonicalize": 2, "center": 3, "shift_positive": 4}


def unary(op: str, x, params) -> np.ndarray:
    v = _i64(x).copy()
    batch, limbs, n = v.shape
    with _Call():
        lib().or_unary(_C(UNARY_OPS[op]), _p(v), _I(batch), _I(limbs), _I(n), _p(_i64(params)))
    return v


def row_scalars(x, scalars, params, canonical: bool) -> np.ndarray:
    x = _i64(x)
    out = np.empty_like(x)
    batch, limbs, n = x.shape
    with _Call():
        lib().or_row_scalars(_p(out), _p(x), _p(_i64(scalars)), _p(_i64(params)), _I(batch), _I(limbs),
                             _I(n), _C(int(canonical)))
    return out


def add_lazy_twice(a, b, twice) -> np.ndarray:
    a = _i64(a)
    b = _i64(b)
    out = np.empty_like(a)
    batch, limbs, n = a.shape
    with _Call():
        lib().or_add_lazy_twice(_p(out), _p(a), _p(b), _p(_i64(twice)), _I(batch), _I(b.shape[0]),
                                _I(limbs), _I(n))
    return out


def lift(coefficients, twice) -> np.ndarray:
    c = _i64(coefficients)
    twice = _i64(twice)
    batch, n = c.shape
    out = np.empty((batch, twice.shape[0], n), dtype=np.int64)
    with _Call():
        lib().or_lift(_p(out), _p(c), _p(twice), _I(batch), _I(twice.shape[0]), _I(n))
    return out


def mixed_radix_decompose(src, norm, prop, lo, hi, klo, khi) -> np.ndarray:
    s = _i64(src)
    out = np.empty_like(s)
    batch, rows, n = s.shape
    norm = _i64(norm) if len(norm) else np.zeros(1, np.int64)
    prop = _i64(prop) if np.size(prop) else np.zeros((1, rows), np.int64)
    with _Call():
        lib().or_mixed_radix_decompose(_p(out), _p(s), _I(batch), _I(rows), _I(n), _p(norm), _p(prop),
                                       _p(_i64(lo)), _p(_i64(hi)), _p(_i64(klo)), _p(_i64(khi)))
    return out


def basis_extend(digits, coef, params, dest_rows: int) -> np.ndarray:
    d = _i64(digits)
    batch, dcount, n = d.shape
    coef = _i64(coef) if np.size(coef) else np.zeros((1, dest_rows), np.int64)
    out = np.empty((batch, dest_rows, n), dtype=np.int64)
    with _Call():
        lib().or_basis_extend(_p(out), _p(d), _I(batch), _I(dcount), _I(dest_rows), _I(n), _p(coef),
                              _I(coef.shape[1]), _p(_i64(params)))
    return out


def rescale(remaining, dropped, inverse, params, half_drop: int, nearest: bool) -> np.ndarray:
    r = _i64(remaining)
    out = np.empty_like(r)
    batch, limbs, n = r.shape
    with _Call():
        lib().or_rescale(_p(out), _p(r), _p(_i64(dropped)), _p(_i64(inverse)), _p(_i64(params)),
                         _I(batch), _I(limbs), _I(n), _I(half_drop), _C(int(nearest)))
    return out


def galois_coefficient(x, src, sign, twice) -> np.ndarray:
    x = _i64(x)
    out = np.empty_like(x)
    batch, limbs, n = x.shape
    src = np.ascontiguousarray(src, dtype=np.int32)
    sign = np.ascontiguousarray(sign, dtype=np.int8)
    with _Call():
        lib().or_galois_coef(_p(out), _p(x), _p(src), _p(sign), _p(_i64(twice)), _I(batch), _I(limbs), _I(n))
    return out


def galois_ntt(x, src) -> np.ndarray:
    x = _i64(x)
    out = np.empty_like(x)
    batch, limbs, n = x.shape
    src = np.ascontiguousarray(src, dtype=np.int32)
    with _Call():
        lib().or_galois_ntt(_p(out), _p(x), _p(src), _I(batch), _I(limbs), _I(n))
    return out


def keyswitch_mac(acc0, acc1, digit, key, params, key_row_start: int,
                  threads: int | None = None) -> tuple[np.ndarray, np.ndarray]:
    a0 = _i64(acc0).copy()
    a1 = _i64(acc1).copy()
    d = _i64(digit)
    key = _i64(key)
    batch, limbs, n = d.shape
    with _Call():
        lib().or_keyswitch_mac(_p(a0), _p(a1), _p(d), _p(key), _I(key.shape[1]), _p(_i64(params)),
                               _I(key_row_start), _I(batch), _I(limbs), _I(n),
                               _C(default_threads() if threads is None else threads))
    return a0, a1


def moddown(q, p, inverse, params, threads: int | None = None) -> np.ndarray:
    q = _i64(q)
    p = _i64(p)
    inverse = _i64(inverse)
    batch, qc, n = q.shape
    pc = p.shape[1]
    out = np.empty_like(q)
    with _Call():
        lib().or_moddown(_p(out), _p(q), _p(p), _p(inverse), _I(inverse.shape[1]), _p(_i64(params)),
                         _I(batch), _I(qc), _I(pc), _I(n),
                         _C(default_threads() if threads is None else threads))
    return out


PLAINTEXT_LAYOUTS = {"dense": 0, "cyclic": 1, "contiguous": 2, "strided": 3}


def add_plaintext(layout: str, ct, pt, params, implicit=None) -> np.ndarray:
    ct = _i64(ct)
    pt = _i64(pt)
    out = np.empty_like(ct)
    batch, limbs, n = ct.shape
    imp = _i64(implicit) if implicit is not None else np.zeros(1, np.int64)
    with _Call():
        lib().or_add_plaintext(_C(PLAINTEXT_LAYOUTS[layout]), _p(out), _p(ct), _p(pt), _p(imp),
                               _I(pt.shape[0]), _p(_i64(params)), _I(batch), _I(limbs), _I(n),
                               _I(pt.shape[2]))
    return out
