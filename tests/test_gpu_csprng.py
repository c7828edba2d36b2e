"""gfx950 ChaCha20 keystream and samplers against the checker: the
keystream bit for bit against OpenSSL's (RFC 8439), the samplers bit for bit
against their closed forms of the same keystream words."""

from __future__ import annotations

import ctypes
import json
import pathlib

import numpy as np
import pytest
import torch

from fhelium_amd import _native as nat
from fhelium_amd.csprng import Csprng
from oracle import chacha20 as C

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "chacha20_vectors.json").read_text())


def words(hexstr: str) -> list[int]:
    b = bytes.fromhex(hexstr)
    return [int.from_bytes(b[i: i + 4], "little") for i in range(0, len(b), 4)]


def arr(ws, n):
    return (ctypes.c_uint32 * n)(*ws)


def test_keystream_matches_openssl():
    for v in GOLDEN["vectors"]:
        out = torch.empty(v["blocks"] * 16, dtype=torch.uint32, device="cuda")
        nat.call("fh_chacha20_keystream", arr(words(v["key"]), 8), arr(words(v["nonce"]), 3), v["counter"],
                 v["blocks"], out.data_ptr(), nat.stream_ptr(out.device))
        got = out.cpu().numpy()
        assert C.words_to_bytes(got).hex() == v["keystream"]


def test_counter_wrap_is_rejected():
    out = torch.empty(32, dtype=torch.uint32, device="cuda")
    with pytest.raises(RuntimeError, match="wrap"):
        nat.call("fh_chacha20_keystream", arr([0] * 8, 8), arr([0] * 3, 3), 0xFFFFFFFF, 2, out.data_ptr(),
                 nat.stream_ptr(out.device))


def test_samplers_match_checker():
    rng = Csprng(seed=77, device="cuda", nonce=0x1234_5678_9ABC)
    key = list(rng.key)
    q = [(1 << 40) - 87, 1099511627689, (1 << 60) - 93]
    moduli = torch.tensor(q, dtype=torch.int64, device="cuda")
    # each request starts where the previous one ended
    u = rng.uniform_mod(moduli, (3,), 1000)        # 9000 values -> 2250 blocks
    t = rng.ternary((5000,))                          # 625 blocks
    g = rng.gaussian((7, 1000), 3.19)                 # 875 blocks
    torch.cuda.synchronize()
    nonce = list(rng.nonce)
    ks = C.keystream(key, nonce, 0, 2250 + 625 + 875)
    np.testing.assert_array_equal(u.cpu().numpy(), C.uniform_mod(ks, q, 3, 1000))
    np.testing.assert_array_equal(t.cpu().numpy(), C.ternary(ks[2250 * 16:], 5000))
    want_g = C.gaussian(ks[(2250 + 625) * 16:], 7000, 3.19).reshape(7, 1000)
    mism = int((g.cpu().numpy() != want_g).sum())
    assert mism <= 1, mism  # libm last-ulp differences can flip a rounding tie
    assert rng.counter == 2250 + 625 + 875


def test_stream_word_advances_instead_of_wrapping():
    rng = Csprng(seed=5, device="cuda")
    rng.counter = (1 << 32) - 10
    before = rng.nonce
    rng.ternary((8 * 20,))
    assert rng.nonce[2] == before[2] + 1 and rng.counter == 20
