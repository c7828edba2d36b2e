"""Plaintext matrix times encrypted vector by the diagonal method.

The reference's packed matrix-vector benchmark
(fhelium/benchmarks/standalone/packed_matrix_vector.py:120-170) computes
y = sum_i d_i * rot(x, i) with hoisted rotations (rotate_many_with_keys),
plaintext products and additions.  Here the K - 1 rotations of the one
ciphertext run as one hoisted fused key switch (``CkksEngine.rotate_hoisted``:
c1 is decomposed and extended once, and every rotation key's MAC reads that
intermediate through its NTT-domain Galois map), so the rotated ciphertexts
carry exactly the reference's hoisted residues.  The products run as one
batched plaintext multiply and the sum as a pairwise tree of batched
canonical additions; one inverse NTT and one rescale finish.
"""

from __future__ import annotations

import numpy as np
import torch

from fhelium_amd import kernels as K
from fhelium_amd.core import Ciphertext, Plaintext


def prepare_diagonals(engine, diagonals, level: int) -> Plaintext:
    """[K, <=S] diagonals -> one batched NTT-domain plaintext [K, Q_l, N]."""

    d = torch.as_tensor(np.asarray(diagonals), device=engine.device)
    return engine.encode_on_device(d, level=level, polynomial_domain="ntt")


def _sum_entries(engine, data: torch.Tensor, level: int) -> torch.Tensor:
    """[2, B, Q, N] canonical residues -> [2, Q, N] sum over the batch."""

    params = engine.runtime.params(level)
    while data.size(1) > 1:
        half = data.size(1) // 2
        summed = torch.empty_like(data[:, :half])
        flat = lambda t: t.reshape(-1, t.size(-2), t.size(-1))  # noqa: E731
        summed.copy_(data[:, :half])
        K.add_canonical_(flat(summed), flat(data[:, half: 2 * half].contiguous()), params)
        if data.size(1) % 2:
            summed = torch.cat([summed, data[:, -1:]], dim=1)
        data = summed
    return data[:, 0]


def diagonal_matvec(engine, ct: Ciphertext, diagonals: Plaintext, *, chunk: int = 128) -> Ciphertext:
    """y = sum_{i<K} d_i * rot(x, i), rescaled once: ``ct`` is one
    coefficient-domain ciphertext [2, Q_l, N]; ``diagonals`` the batched
    NTT-domain plaintext from prepare_diagonals (K entries).  Rotations run
    ``chunk`` at a time, each chunk one hoisted key switch."""

    engine._assert_rotation_ciphertext(ct)
    if ct.batch_shape:
        raise ValueError("diagonal_matvec takes one unbatched ciphertext")
    if len(diagonals.batch_shape) != 1 or diagonals.level != ct.level:
        raise ValueError("diagonals must be a [K, Q_l, N] plaintext at the ciphertext's level")
    k_total = diagonals.batch_shape[0]
    acc = None
    for start in range(0, k_total, chunk):
        steps = list(range(start, min(start + chunk, k_total)))
        b = len(steps)
        # entry i holds rot(x, steps[i]); step 0 is the ciphertext itself
        rotated = engine.rotate_hoisted(ct, steps)
        pt = Plaintext(diagonals.data[start: start + b], diagonals.level, diagonals.scale, "ntt")
        prod = engine.multiply_plaintext(engine.coefficient_domain_to_ntt_domain(rotated), pt)
        part = _sum_entries(engine, prod.data, ct.level)
        if acc is None:
            acc = part
        else:
            K.add_canonical_(acc, part, engine.runtime.params(ct.level))
    summed = Ciphertext(acc, ct.level, prod.scale, "ntt", "montgomery")
    return engine.rescale_to_next_level(engine.ntt_domain_to_coefficient_domain(summed))
