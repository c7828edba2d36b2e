"""Seeded ChaCha20 sampling on the device (this backend's generator).

The reference samples keys and ciphertexts from ChaCha20 streams
(fhelium/rng/csprng.py:74-250, via the external triton-csprng package).
``Csprng`` keeps a 256-bit key, a 96-bit nonce and a block counter; every
request takes the next run of keystream blocks, so the stream never repeats
within one nonce.  The seed is turned into key words as the reference does
(fhelium/rng/csprng.py:32-51: SHA-256 of the 32-byte little-endian signed
seed, read as eight little-endian words); the nonce is (low 32 bits, high 32
bits, stream word), and the stream word advances when a request would wrap
the 32-bit block counter.

Without a seed the 256-bit key and the 64-bit nonce come from the operating
system's entropy (``secrets``), as the reference leaves production entropy to
its generator when ``rng_seed`` is None (fhelium/engine/ckks_engine.py:72-74);
fixed seeds are for tests and benchmarks.

The reference's own streams cannot be reproduced here (its package is not
available), so keys and ciphertexts differ from the reference's for equal
seeds; the block function itself is checked against RFC 8439.
"""

from __future__ import annotations

import ctypes
import hashlib
import secrets

import torch

from fhelium_amd import _native as nat

_MASK = 0xFFFFFFFF


def key_words_from_seed(seed: int) -> list[int]:
    digest = hashlib.sha256(int(seed).to_bytes(32, "little", signed=True)).digest()
    return [int.from_bytes(digest[i: i + 4], "little") for i in range(0, 32, 4)]


def entropy_key_words() -> list[int]:
    """Eight 32-bit key words from the operating system's CSPRNG."""

    raw = secrets.token_bytes(32)
    return [int.from_bytes(raw[i: i + 4], "little") for i in range(0, 32, 4)]


class Csprng:
    def __init__(self, seed: int | None, device: torch.device | str, nonce: int | None = None) -> None:
        self.device = torch.device(device)
        words = entropy_key_words() if seed is None else key_words_from_seed(seed)
        self.key = (ctypes.c_uint32 * 8)(*words)
        if nonce is None:
            nonce = 0 if seed is not None else secrets.randbits(64)
        self._nonce = [nonce & _MASK, (nonce >> 32) & _MASK, 0]
        self.counter = 0

    @property
    def nonce(self) -> tuple[int, int, int]:
        return tuple(self._nonce)

    def _take(self, blocks: int) -> tuple[ctypes.Array, int]:
        if blocks > 1 << 32:
            raise ValueError("one request may use at most 2^32 ChaCha20 blocks")
        if self.counter + blocks > 1 << 32:
            self._nonce[2] = (self._nonce[2] + 1) & _MASK
            self.counter = 0
        start = self.counter
        self.counter += blocks
        return (ctypes.c_uint32 * 3)(*self._nonce), start

    def keystream(self, blocks: int) -> torch.Tensor:
        out = torch.empty(blocks * 16, dtype=torch.uint32, device=self.device)
        nonce, start = self._take(blocks)
        nat.call("fh_chacha20_keystream", self.key, nonce, start, blocks, out.data_ptr(),
                 nat.stream_ptr(self.device))
        return out

    def uniform_mod(self, moduli: torch.Tensor, outer: tuple[int, ...], count: int) -> torch.Tensor:
        """[*outer, rows, count] uniform residues; moduli: device int64 [rows]."""

        rows = moduli.numel()
        n_outer = 1
        for d in outer:
            n_outer *= int(d)
        out = torch.empty((*outer, rows, count), dtype=torch.int64, device=self.device)
        total = n_outer * rows * count
        nonce, start = self._take(-(-total // 4))
        nat.call("fh_sample_uniform_mod", self.key, nonce, start, moduli.contiguous().data_ptr(), n_outer, rows,
                 count, out.data_ptr(), nat.stream_ptr(self.device))
        return out

    def ternary(self, shape: tuple[int, ...]) -> torch.Tensor:
        out = torch.empty(shape, dtype=torch.int64, device=self.device)
        nonce, start = self._take(-(-out.numel() // 8))
        nat.call("fh_sample_ternary", self.key, nonce, start, out.numel(), out.data_ptr(),
                 nat.stream_ptr(self.device))
        return out

    def gaussian(self, shape: tuple[int, ...], sigma: float) -> torch.Tensor:
        out = torch.empty(shape, dtype=torch.int64, device=self.device)
        nonce, start = self._take(-(-out.numel() // 8))
        nat.call("fh_sample_gaussian", self.key, nonce, start, out.numel(), float(sigma), out.data_ptr(),
                 nat.stream_ptr(self.device))
        return out
