"""Host-side RNS tables: Montgomery parameters, the Q/P layout, hybrid
key-switch digits and every engine-owned scalar table of the hot path.

All values are computed with Python integers from closed forms and stored as
int64 CPU tensors in the reference's layouts, so a table built here is
interchangeable with the reference's:

* ``rns_params`` [8, QP rows]: 2q, q_lo, q_hi, (-q^-1)_lo, (-q^-1)_hi, R^2,
  Delta_0 R^2, N^-1 R  (csrc/ops/common/rns_parameters.h:6-16,
  fhelium/engine/rns/runtime.py:136-216, montgomery.py:59-114).
* hybrid digits: scale primes split into chunks of ``num_p_primes``, then the
  base prime alone; a digit keeps its level-0 key index as Q rows drop
  (fhelium/engine/rns/decomposition.py:30-53).
* mixed-radix normalizers / propagation / basis-extension coefficients
  (fhelium/engine/rns/parameters.py:172-259).
* rescale inverses, ModDown P-drop inverses, P R mod q
  (fhelium/engine/ckks_engine.py:2952-3072).
* Galois gathers (fhelium/engine/galois.py:10-83,
  hybrid_keyswitch.py:489-528).
"""

from __future__ import annotations

import functools
import math
from dataclasses import dataclass

import torch

from fhelium_amd.config import CkksConfig

RADIX_BITS = 62
HALF_BITS = 31
R = 1 << RADIX_BITS
_HALF_MASK = (1 << HALF_BITS) - 1


def montgomery_rows(moduli: tuple[int, ...], log_n: int, scale_bits: int) -> list[list[int]]:
    """The eight reference parameter rows for ``moduli`` in order."""

    n = 1 << log_n
    rows = [[] for _ in range(8)]
    for q in moduli:
        k = (-pow(q, -1, R)) % R
        r2 = pow(R, 2, q)
        rows[0].append(2 * q)
        rows[1].append(q & _HALF_MASK)
        rows[2].append(q >> HALF_BITS)
        rows[3].append(k & _HALF_MASK)
        rows[4].append(k >> HALF_BITS)
        rows[5].append(r2)
        rows[6].append((r2 << scale_bits) % q)
        rows[7].append((pow(n, -1, q) * R) % q)
    return rows


@dataclass(frozen=True)
class DigitSpec:
    """One active hybrid digit (fhelium/engine/rns/layout.py:13-27)."""

    level: int
    digit_index: int
    key_digit_index: int
    prime_ids: tuple[int, ...]
    component_row_ids: tuple[int, ...]


class RnsLayout:
    """Q/P row layout of one chain.

    Level ``l`` keeps Q rows ``l .. num_q-1``; P rows follow.  Tensors are
    ``[..., limb, N]`` with limb ``j`` = ``prime_ids(l, include_p)[j]``.
    """

    def __init__(self, config: CkksConfig) -> None:
        self.config = config
        self.num_q = config.num_q_primes
        self.num_p = config.num_p_primes
        self.q_prime_ids = tuple(range(self.num_q))
        self.p_prime_ids = tuple(range(self.num_q, self.num_q + self.num_p))
        scale_ids = self.q_prime_ids[:-1]
        width = self.num_p
        self.level_zero_digits = tuple(
            tuple(scale_ids[s : s + width]) for s in range(0, len(scale_ids), width)
        ) + ((self.q_prime_ids[-1],),)
        self._specs = tuple(self._build_specs(level) for level in range(config.public_level_count))

    @property
    def digit_count(self) -> int:
        return len(self.level_zero_digits)

    def check_level(self, level: int) -> None:
        if not 0 <= level < self.config.public_level_count:
            raise ValueError(f"level {level} outside [0, {self.config.public_level_count})")

    def prime_ids(self, level: int, include_p: bool = False) -> tuple[int, ...]:
        self.check_level(level)
        ids = tuple(range(level, self.num_q))
        return ids + self.p_prime_ids if include_p else ids

    def start_row(self, level: int) -> int:
        return level

    def _build_specs(self, level: int) -> tuple[DigitSpec, ...]:
        specs = []
        for key_index, digit in enumerate(self.level_zero_digits):
            ids = tuple(p for p in digit if p >= level)
            if ids:
                specs.append(
                    DigitSpec(level, len(specs), key_index, ids, tuple(p - level for p in ids))
                )
        return tuple(specs)

    def digit_specs(self, level: int) -> tuple[DigitSpec, ...]:
        self.check_level(level)
        return self._specs[level]


@dataclass
class DigitTables:
    """Per-digit ModUp tables (parameters.py:172-259)."""

    prime_ids: tuple[int, ...]
    normalizers: torch.Tensor          # [D-1]       (prod_{t<=r} b_t)^{-1} R mod b_{r+1}
    propagation: torch.Tensor          # [D-1, D]    prod_{t<=r} b_t R mod b_target
    modulus_lo: torch.Tensor           # [D]
    modulus_hi: torch.Tensor
    neg_inv_lo: torch.Tensor
    neg_inv_hi: torch.Tensor
    basis_extension: torch.Tensor      # [D-1, QP0]  M_r R^2 mod p_j over level-0 QP rows
    basis_extension_std: torch.Tensor  # [D-1, QP0]  M_r R mod p_j (REDC of the above)
    basis_extension_shoup: torch.Tensor  # [D-1, QP0] floor(std * 2^64 / p_j) as two's complement
    # binary64 Garner record when every prime is below 2^45 (else empty):
    # [D-1 normalizers, (D-1) x D propagation, D moduli, D reciprocals], the
    # constants in standard form (the tables above without their R factor)
    garner_f64: torch.Tensor


# Digits whose primes are all below this are carried as binary64 (exact).
F64_DIGIT_LIMIT = 1 << 45


def _digit_tables(moduli: tuple[int, ...], prime_ids: tuple[int, ...], qp_ids: tuple[int, ...]) -> DigitTables:
    d = len(prime_ids)
    src = [moduli[i] for i in prime_ids]
    products = [src[0]]
    for idx in range(1, d - 1):
        products.append(products[-1] * src[idx])
    normalizers = []
    propagation = [[0] * d for _ in range(max(d - 1, 0))]
    for c in range(d - 1):
        normalizers.append((pow(products[c], -1, src[c + 1]) * R) % src[c + 1])
        for target in range(c + 2, d):
            propagation[c][target] = (products[c] * R) % src[target]
    dest = [moduli[i] for i in qp_ids]
    r2 = [pow(R, 2, p) for p in dest]
    basis = [[(products[c] * r2[j]) % dest[j] for j in range(len(dest))] for c in range(d - 1)]
    basis_std = [[(products[c] * R) % dest[j] for j in range(len(dest))] for c in range(d - 1)]
    basis_shoup = [[_as_signed64((w << 64) // dest[j]) for j, w in enumerate(row)] for row in basis_std]
    ks = [(-pow(q, -1, R)) % R for q in src]
    garner = []
    if d > 1 and all(q < F64_DIGIT_LIMIT for q in src):
        garner += [float(pow(products[c], -1, src[c + 1])) for c in range(d - 1)]
        garner += [float(products[c] % src[t]) if t >= c + 2 else 0.0 for c in range(d - 1) for t in range(d)]
        garner += [float(q) for q in src] + [1.0 / q for q in src]
    as_t = lambda v, shape=None: torch.tensor(v, dtype=torch.int64).reshape(shape) if shape else torch.tensor(v, dtype=torch.int64)
    return DigitTables(
        prime_ids=prime_ids,
        normalizers=as_t(normalizers) if normalizers else torch.empty(0, dtype=torch.int64),
        propagation=as_t(propagation, (d - 1, d)) if d > 1 else torch.empty(0, d, dtype=torch.int64),
        modulus_lo=as_t([q & _HALF_MASK for q in src]),
        modulus_hi=as_t([q >> HALF_BITS for q in src]),
        neg_inv_lo=as_t([k & _HALF_MASK for k in ks]),
        neg_inv_hi=as_t([k >> HALF_BITS for k in ks]),
        basis_extension=as_t(basis, (d - 1, len(dest))) if d > 1 else torch.empty(0, len(dest), dtype=torch.int64),
        basis_extension_std=(as_t(basis_std, (d - 1, len(dest))) if d > 1
                             else torch.empty(0, len(dest), dtype=torch.int64)),
        basis_extension_shoup=(as_t(basis_shoup, (d - 1, len(dest))) if d > 1
                               else torch.empty(0, len(dest), dtype=torch.int64)),
        garner_f64=torch.tensor(garner, dtype=torch.float64),
    )


def _as_signed64(v: int) -> int:
    return v - (1 << 64) if v >= (1 << 63) else v


# Rows whose prime is below this limit run in binary64 arithmetic; larger
# primes take the integer (Montgomery / Shoup) path.  2^42 is what the fast
# transform's growth bound covers: values reach 11q over 16 unreduced
# butterfly stages and the key-switch accumulators up to 0.53 (digits + 1) q,
# and every binary64 product needs |w x| / q below 2^46.4 to stay exact.
BINARY64_PRIME_LIMIT = 1 << 42

# Binary64 ModDown split: a P digit c < 2^60 is taken as c_hi 2^30 + c_lo.
MODDOWN_SPLIT_BITS = 30


MODDOWN_RECORD_WORDS = 16


def moddown_binary64_table(moduli: tuple[int, ...], p_ids: tuple[int, ...],
                           q_ids: tuple[int, ...]) -> torch.Tensor:
    """[Q0, 16] float64 ModDown records, one per Q row whose prime is below
    BINARY64_PRIME_LIMIT: words 0..P-1 hold M_r mod q, words P..2P-1 hold 2^30 M_r mod q
    (M_r = prod_{t<r} p_t), word 2P holds P^{-1} mod q, word 14 holds q and
    word 15 holds 1/q.  Rows of larger primes are all zero (they take the
    integer path).  One 128-byte record per row lets a wave fetch a row's
    constants with two scalar loads."""

    p = [moduli[i] for i in p_ids]
    if 2 * len(p) + 1 > 14:
        raise ValueError("the ModDown record holds at most 6 P primes")
    prefix = [1]
    for v in p[:-1]:
        prefix.append(prefix[-1] * v)
    p_product = math.prod(p)
    out = torch.zeros((len(q_ids), MODDOWN_RECORD_WORDS), dtype=torch.float64)
    for j, qi in enumerate(q_ids):
        q = moduli[qi]
        if q >= BINARY64_PRIME_LIMIT:
            continue
        for r, m in enumerate(prefix):
            out[j, r] = float(m % q)
            out[j, len(p) + r] = float((m << MODDOWN_SPLIT_BITS) % q)
        out[j, 2 * len(p)] = float(pow(p_product, -1, q))
        out[j, 14] = float(q)
        out[j, 15] = 1.0 / q
    return out


class HostTables:
    """Every scalar table of the hot path for one config, as CPU tensors."""

    def __init__(self, config: CkksConfig) -> None:
        self.config = config
        self.layout = RnsLayout(config)
        moduli = config.moduli
        self.moduli = moduli
        self.rns_params = torch.tensor(
            montgomery_rows(moduli, config.logN, config.scale_bits), dtype=torch.int64
        ).contiguous()
        qp0 = self.layout.prime_ids(0, include_p=True)
        self.digit_tables: dict[tuple[int, ...], DigitTables] = {}
        for level in range(config.public_level_count):
            for spec in self.layout.digit_specs(level):
                if spec.prime_ids not in self.digit_tables:
                    self.digit_tables[spec.prime_ids] = _digit_tables(moduli, spec.prime_ids, qp0)
        self.rescale_inverses = [self._rescale_inverse(level) for level in range(config.public_level_count)]
        self.moddown_inverses = self._moddown_inverses()
        p_product = math.prod(config.p_moduli)
        self.p_product_montgomery_q = torch.tensor(
            [(p_product * R) % moduli[i] for i in self.layout.prime_ids(0)], dtype=torch.int64
        )
        # Fused ModDown tables: mixed-radix digits of the P residues (the P
        # primes treated as one digit) extended onto every QP row, and
        # P^{-1} mod q for the Q rows.  floor(X / P) = (X - (X mod P)) / P.
        self.p_digit = _digit_tables(moduli, self.layout.p_prime_ids, qp0)
        self.p_inverse_q = torch.tensor(
            [pow(p_product, -1, moduli[i]) for i in self.layout.prime_ids(0)], dtype=torch.int64
        )
        # Key-switch accumulator seed scales: P mod q for the binary64 rows
        # (their accumulators hold x R^2 and the seeds a b R^2), P R mod q for
        # the integer rows (exact sums reduced by one REDC: x R, seeds a b R).
        self.seed_scale = torch.tensor(
            [p_product % moduli[i] if moduli[i] < BINARY64_PRIME_LIMIT else (p_product * R) % moduli[i]
             for i in self.layout.prime_ids(0)], dtype=torch.int64
        )
        self.moddown_binary64 = moddown_binary64_table(moduli, self.layout.p_prime_ids,
                                                       self.layout.prime_ids(0))

    def _rescale_inverse(self, level: int) -> torch.Tensor:
        ids = self.layout.prime_ids(level, include_p=True)
        drop = self.moduli[ids[0]]
        return torch.tensor(
            [(pow(drop, -1, self.moduli[i]) * R) % self.moduli[i] for i in ids[1:]], dtype=torch.int64
        )

    def _moddown_inverses(self) -> list[torch.Tensor]:
        # ckks_engine.py:2991-3048: row d drops p_{k-1-d}; entries cover the
        # surviving QP rows of level 0, then each level slices from its start.
        cfg = self.config
        k = cfg.num_p_primes
        dropped = list(cfg.p_moduli)[::-1]
        base_rows = []
        qp0 = self.layout.prime_ids(0, include_p=True)
        for step, p in enumerate(dropped):
            survivors = qp0[: len(qp0) - step - 1]
            base_rows.append([(pow(p, -1, self.moduli[i]) * R) % self.moduli[i] for i in survivors])
        out = []
        for level in range(cfg.public_level_count):
            start = self.layout.start_row(level)
            rows = [r[start:] for r in base_rows]
            width = max(len(r) for r in rows)
            packed = torch.zeros((k, width), dtype=torch.int64)
            for idx, r in enumerate(rows):
                packed[idx, : len(r)] = torch.tensor(r, dtype=torch.int64)
            out.append(packed)
        return out


# ---------------------------------------------------------------------------
# NTT twiddles (fhelium/engine/ntt/plans/twiddles.py:10-84)
# ---------------------------------------------------------------------------


def primitive_2nth_root(q: int, n: int) -> int:
    """First c = 2, 3, ... whose c^((q-1)/2N) is a primitive 2N-th root."""

    exponent = (q - 1) // (2 * n)
    for candidate in range(2, max(n, 3)):
        root = pow(candidate, exponent, q)
        if pow(root, n, q) != 1:
            return root
    raise RuntimeError(f"no primitive {2 * n}-th root modulo {q}")


@functools.lru_cache(maxsize=None)
def bit_reverse_table(log_n: int) -> tuple[int, ...]:
    n = 1 << log_n
    out = [0] * n
    for i in range(n):
        out[i] = int(f"{i:0{log_n}b}"[::-1], 2) if log_n else 0
    return tuple(out)


def _power_series(root: int, n: int, q: int) -> list[int]:
    out = [1] * n
    acc = 1
    for i in range(1, n):
        acc = acc * root % q
        out[i] = acc
    return out


@functools.lru_cache(maxsize=8)
def compact_twiddles_montgomery(moduli: tuple[int, ...], log_n: int) -> tuple[torch.Tensor, torch.Tensor]:
    """Forward psi^{br(k)} R and inverse psi^{-br(k)} R mod q, [rows, N]."""

    n = 1 << log_n
    br = bit_reverse_table(log_n)
    fwd = torch.empty((len(moduli), n), dtype=torch.int64)
    inv = torch.empty((len(moduli), n), dtype=torch.int64)
    for row, q in enumerate(moduli):
        psi = primitive_2nth_root(q, n)
        f = _power_series(psi, n, q)
        i = _power_series(pow(psi, -1, q), n, q)
        fwd[row] = torch.tensor([f[br[k]] * R % q for k in range(n)], dtype=torch.int64)
        inv[row] = torch.tensor([i[br[k]] * R % q for k in range(n)], dtype=torch.int64)
    return fwd, inv


def power_of_two_radix_twiddles(moduli: tuple[int, ...], log_n: int,
                                radix: int) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """The reference's fixed-radix NTT tables in Montgomery form
    (fhelium/engine/ntt/plans/twiddles.py:118-228, converted as its runtime
    does): forward and inverse outer twists [rows, N-1] and forward and
    inverse radix-root powers [rows, radix].  Digit d of the transform
    (stage_start = d * b, b = log2(radix); mirrored for the inverse) stores,
    for each of its 2^stage_start groups, beta^1 .. beta^(radix-1) with beta
    = root^{br((1 << (stage_start + b - 1)) + group * radix / 2)}; the root
    powers are root^{i 2N / radix}."""

    if radix not in (4, 8, 16):
        raise ValueError("radix must be 4, 8 or 16")
    bits = {4: 2, 8: 3, 16: 4}[radix]
    if log_n % bits:
        raise ValueError("the radix width must divide log_n")
    n = 1 << log_n
    order = 2 * n

    def exponents(inverse: bool) -> list[int]:
        out = []
        for stage_start in range(0, log_n, bits):
            first = log_n - stage_start - bits if inverse else stage_start
            leaf = first + bits - 1
            for group in range(1 << first):
                beta = int(f"{(1 << leaf) + group * (radix // 2):0{log_n}b}"[::-1], 2)
                out.extend(beta * lane % order for lane in range(1, radix))
        assert len(out) == n - 1
        return out

    fwd_e, inv_e = exponents(False), exponents(True)
    rows = len(moduli)
    fo = torch.empty((rows, n - 1), dtype=torch.int64)
    io = torch.empty((rows, n - 1), dtype=torch.int64)
    fr = torch.empty((rows, radix), dtype=torch.int64)
    ir = torch.empty((rows, radix), dtype=torch.int64)
    for row, q in enumerate(moduli):
        psi = primitive_2nth_root(q, n)
        f = _power_series(psi, order, q)
        i = _power_series(pow(psi, -1, q), order, q)
        fo[row] = torch.tensor([f[e] * R % q for e in fwd_e], dtype=torch.int64)
        io[row] = torch.tensor([i[e] * R % q for e in inv_e], dtype=torch.int64)
        fr[row] = torch.tensor([f[k * order // radix] * R % q for k in range(radix)], dtype=torch.int64)
        ir[row] = torch.tensor([i[k * order // radix] * R % q for k in range(radix)], dtype=torch.int64)
    return fo, io, fr, ir


# ---------------------------------------------------------------------------
# Galois tables
# ---------------------------------------------------------------------------


def rotation_galois_element(n: int, step: int, generator: int = 3) -> int:
    """g = generator^(step mod N) mod 2N (galois.py:10-26)."""

    exponent = -int(step) if generator == 5 else int(step)
    return pow(generator, exponent % n, 2 * n)


@functools.lru_cache(maxsize=256)
def coefficient_galois_tables(n: int, galois_element: int) -> tuple[torch.Tensor, torch.Tensor]:
    """Destination j reads source_indices[j] times source_sign[j] (galois.py:50-83)."""

    two_n = 2 * n
    g = galois_element % two_n
    if g % 2 == 0:
        raise ValueError("galois_element must be odd")
    source = torch.arange(n, dtype=torch.int64)
    mapped = (g * source) % two_n
    dest = mapped % n
    sign = torch.where((mapped // n) % 2 == 1, -1, 1).to(torch.int8)
    idx = torch.empty(n, dtype=torch.int32)
    idx[dest] = source.to(torch.int32)
    sgn = torch.empty(n, dtype=torch.int8)
    sgn[dest] = sign
    return idx, sgn


@functools.lru_cache(maxsize=256)
def ntt_galois_indices(log_n: int, galois_element: int) -> torch.Tensor:
    """src[k] = br(((2 br(k) + 1) g mod 2N - 1) / 2) (hybrid_keyswitch.py:506-528)."""

    n = 1 << log_n
    br = torch.tensor(bit_reverse_table(log_n), dtype=torch.int64)
    dest_exp = 2 * br + 1
    src_br = ((dest_exp * galois_element) % (2 * n) - 1) // 2
    return br.index_select(0, src_br).to(torch.int32)
