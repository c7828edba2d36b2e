"""Device slot codec (fh_ckks_encode / fh_ckks_decode) against the binary64
host codec (fhelium_amd/codec.py, numpy's FFT): encoded coefficients agree
except for rounding ties, and encode -> decrypt round trips on the device
recover the slots within the reference's tolerance
(tests/test_ckks_operation_correctness.py:48-51)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd import codec
from fhelium_amd.engine import CkksEngine
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[12, 16])
def eng(request):
    e = CkksEngine(H.config(request.param), rng_seed=71)
    e.generate_secret_key()
    return e


def test_encode_matches_host_codec(eng):
    rng = np.random.default_rng(1)
    z = rng.uniform(-1, 1, (2, eng.config.slots)) + 1j * rng.uniform(-1, 1, (2, eng.config.slots))
    scale = eng.config.default_scale
    want = codec.encode(z, eng.config.N, scale)  # [2, N] int64 coefficients
    pt = eng.encode_on_device(torch.from_numpy(z).to("cuda"), level=3, scale=scale)
    # coefficient-domain plaintexts hold Montgomery residues: back to
    # standard residues of the first row, centred
    from fhelium_amd import kernels as K
    data = pt.data.clone()
    K.from_montgomery_(data, eng.runtime.params(3))
    q0 = eng.config.moduli[3]
    row = data[:, 0].cpu().numpy()
    got = np.where(row > q0 // 2, row - q0, row)
    diff = np.abs(got - want)
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-3


def test_device_round_trip(eng):
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (3, eng.config.slots))
    pt = eng.encode_on_device(torch.from_numpy(x).to("cuda"), level=5)
    host = eng.encode(x, level=5)
    ct = eng.encrypt(x, level=5)
    dec = eng.decrypt_on_device(ct).cpu().numpy()
    assert np.abs(dec.real - x).max() < 2e-5 and np.abs(dec.imag).max() < 2e-5
    # the device decode agrees with the host decode of the same ciphertext
    assert np.abs(dec - eng.decrypt(ct)).max() < 1e-9
    # device-encoded plaintext added to a ciphertext decrypts as the host one
    a = eng.decrypt_on_device(eng.add_plaintext(ct, pt)).cpu().numpy()
    b = eng.decrypt(eng.add_plaintext(ct, host))
    assert np.abs(a - b).max() < 1e-9 and np.abs(a.real - 2 * x).max() < 2e-5


def test_decode_after_hmult(eng):
    eng.generate_relinearization_key()
    rng = np.random.default_rng(3)
    x, y = rng.uniform(-1, 1, eng.config.slots), rng.uniform(-1, 1, eng.config.slots)
    out = eng.hmult(eng.encrypt(x, level=30), eng.encrypt(y, level=30))
    dec = eng.decrypt_on_device(out).cpu().numpy()
    assert np.abs(dec.real - x * y).max() < 2e-5
