"""ChaCha20 keystream and the samplers built on it -- TEST INFRASTRUCTURE ONLY.

Only tests/ use this module, as the checker of fhelium_amd/csrc/csprng.hip.
It restates the block function of RFC 8439 section 2.3 (vectorised over
blocks with numpy uint32 arithmetic) and the three samplers' use of the
keystream words (uniform mod q from 128 bits, ternary from 64 bits, rounded
Box-Muller pairs from two 53-bit uniforms).

Pinning (tests/test_csprng.py): the block function against keystreams that
OpenSSL computed (tests/golden/chacha20_vectors.json, written by
tests/golden/make_chacha20_vectors.py), which include the RFC 8439
section 2.3.2 block and the appendix A.1 all-zero vector.  The samplers are
closed forms of the keystream words and are checked by the properties the
tests state (ranges, frequencies, moments).
"""

from __future__ import annotations

import numpy as np

_CONST = np.array([0x61707865, 0x3320646E, 0x79622D32, 0x6B206574], dtype=np.uint32)


def _rotl(x: np.ndarray, r: int) -> np.ndarray:
    return (x << np.uint32(r)) | (x >> np.uint32(32 - r))


def _quarter(x: np.ndarray, a: int, b: int, c: int, d: int) -> None:
    x[a] += x[b]; x[d] ^= x[a]; x[d] = _rotl(x[d], 16)
    x[c] += x[d]; x[b] ^= x[c]; x[b] = _rotl(x[b], 12)
    x[a] += x[b]; x[d] ^= x[a]; x[d] = _rotl(x[d], 8)
    x[c] += x[d]; x[b] ^= x[c]; x[b] = _rotl(x[b], 7)


def keystream(key: list[int], nonce: list[int], counter: int, blocks: int) -> np.ndarray:
    """uint32 [blocks * 16]: blocks counter .. counter + blocks - 1."""

    s = np.empty((16, blocks), dtype=np.uint32)
    s[0:4] = _CONST[:, None]
    s[4:12] = np.asarray(key, dtype=np.uint32)[:, None]
    s[12] = (np.arange(blocks, dtype=np.uint64) + np.uint64(counter)).astype(np.uint32)
    s[13:16] = np.asarray(nonce, dtype=np.uint32)[:, None]
    x = s.copy()
    with np.errstate(over="ignore"):
        for _ in range(10):
            _quarter(x, 0, 4, 8, 12); _quarter(x, 1, 5, 9, 13)
            _quarter(x, 2, 6, 10, 14); _quarter(x, 3, 7, 11, 15)
            _quarter(x, 0, 5, 10, 15); _quarter(x, 1, 6, 11, 12)
            _quarter(x, 2, 7, 8, 13); _quarter(x, 3, 4, 9, 14)
        x += s
    return np.ascontiguousarray(x.T).reshape(-1)


def words_to_bytes(words: np.ndarray) -> bytes:
    return words.astype("<u4").tobytes()


def uniform_mod(words: np.ndarray, moduli: list[int], outer: int, count: int) -> np.ndarray:
    rows = len(moduli)
    total = outer * rows * count
    w = words[: 4 * total].astype(np.uint64).reshape(total, 4)
    out = np.empty(total, dtype=np.int64)
    for g in range(total):
        x = int(w[g, 0]) | int(w[g, 1]) << 32 | int(w[g, 2]) << 64 | int(w[g, 3]) << 96
        out[g] = x % moduli[(g // count) % rows]
    return out.reshape(outer, rows, count)


def ternary(words: np.ndarray, total: int) -> np.ndarray:
    w = words[: 2 * total].astype(np.uint64).reshape(total, 2)
    x = w[:, 0] | (w[:, 1] << np.uint64(32))
    return (x % np.uint64(3)).astype(np.int64) - 1


def gaussian(words: np.ndarray, total: int, sigma: float) -> np.ndarray:
    pairs = -(-total // 2)
    w = words[: 4 * pairs].astype(np.uint64).reshape(pairs, 4)

    def unit(lo, hi):
        return (((hi << np.uint64(32)) | lo) >> np.uint64(11)).astype(np.float64) * 2.0**-53 + 2.0**-54

    u1, u2 = unit(w[:, 0], w[:, 1]), unit(w[:, 2], w[:, 3])
    rad = sigma * np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * pairs, dtype=np.int64)
    out[0::2] = np.rint(rad * np.cos(2.0 * np.pi * u2))
    out[1::2] = np.rint(rad * np.sin(2.0 * np.pi * u2))
    return out[:total]
