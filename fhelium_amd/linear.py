"""Plaintext matrix times encrypted vector by the diagonal method.

The reference's packed matrix-vector benchmark
(fhelium/benchmarks/standalone/packed_matrix_vector.py:120-170) computes
y = sum_i d_i * rot(x, i) with hoisted rotations (rotate_many_with_keys),
plaintext products and additions.  Here the K - 1 rotations of the one
ciphertext run as one hoisted fused key switch (``CkksEngine.rotate_hoisted``:
c1 is decomposed and extended once, and every rotation key's MAC reads that
intermediate through its NTT-domain Galois map), so the rotated ciphertexts
carry exactly the reference's hoisted residues.  The plaintext products and
their sum run as one inner-product pass over the rotated batch (canonical
modular sums do not depend on the order of addition, so this equals
multiply_plaintext followed by additions); one inverse NTT and one rescale
finish.
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


def diagonal_matvec(engine, ct: Ciphertext, diagonals: Plaintext, *, chunk: int = 128) -> Ciphertext:
    """y = sum_{i<K} d_i * rot(x, i), rescaled once: ``ct`` is one
    coefficient-domain ciphertext [2, Q_l, N]; ``diagonals`` the batched
    NTT-domain plaintext from prepare_diagonals (K entries).  Rotations run
    ``chunk`` at a time, each chunk one hoisted key switch; each chunk's
    products and their sum are one inner-product pass
    (kernels.plaintext_inner_product)."""

    engine._assert_rotation_ciphertext(ct)
    if ct.batch_shape:
        raise ValueError("diagonal_matvec takes one unbatched ciphertext")
    if len(diagonals.batch_shape) != 1 or diagonals.level != ct.level:
        raise ValueError("diagonals must be a [K, Q_l, N] plaintext at the ciphertext's level")
    k_total = diagonals.batch_shape[0]
    params = engine.runtime.params(ct.level)
    acc = None

    def accumulate(part: torch.Tensor) -> None:
        nonlocal acc
        if acc is None:
            acc = part
        else:
            K.add_canonical_(acc, part, params)

    # step 0: the ciphertext itself times d_0
    x0 = engine.coefficient_domain_to_ntt_domain(ct)
    accumulate(K.plaintext_inner_product(x0.data.unsqueeze(1), diagonals.data[0:1], params))
    for start in range(1, k_total, chunk):
        stop = min(start + chunk, k_total)
        rotated = engine.rotate_hoisted(ct, list(range(start, stop)))  # entry i: rot(x, start + i)
        ntt = engine.coefficient_domain_to_ntt_domain(rotated)
        accumulate(K.plaintext_inner_product(ntt.data, diagonals.data[start:stop], params))
    summed = Ciphertext(acc, ct.level, ct.scale * diagonals.scale, "ntt", "montgomery")
    return engine.rescale_to_next_level(engine.ntt_domain_to_coefficient_domain(summed))
