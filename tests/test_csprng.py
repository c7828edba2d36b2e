"""The ChaCha20 checker against OpenSSL's keystreams (RFC 8439 block
function) and the seed-to-key rule (fhelium/rng/csprng.py:32-51)."""

from __future__ import annotations

import hashlib
import json
import pathlib

import numpy as np

from oracle import chacha20 as C

GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "chacha20_vectors.json").read_text())


def words(hexstr: str) -> list[int]:
    b = bytes.fromhex(hexstr)
    return [int.from_bytes(b[i: i + 4], "little") for i in range(0, len(b), 4)]


def test_block_function_matches_openssl():
    for v in GOLDEN["vectors"]:
        got = C.keystream(words(v["key"]), words(v["nonce"]), v["counter"], v["blocks"])
        assert C.words_to_bytes(got).hex() == v["keystream"]


def test_rfc8439_section_2_3_2_block():
    v = GOLDEN["vectors"][0]
    got = C.words_to_bytes(C.keystream(words(v["key"]), words(v["nonce"]), 1, 1)).hex()
    assert got.startswith("10f1e7e4d13b5915500fdd1fa32071c4")
    assert got.endswith("b5129cd1de164eb9cbd083e8a2503c4e")


def test_seed_to_key_words():
    from fhelium_amd.csprng import key_words_from_seed

    d = hashlib.sha256((12345).to_bytes(32, "little", signed=True)).digest()
    assert key_words_from_seed(12345) == [int.from_bytes(d[i: i + 4], "little") for i in range(0, 32, 4)]
    assert key_words_from_seed(-1) != key_words_from_seed(1)


def test_sampler_checkers_properties():
    ks = C.keystream(list(range(8)), [1, 2, 3], 0, 4096)
    t = C.ternary(ks, 20000)
    counts = np.bincount(t + 1, minlength=3) / t.size
    assert set(np.unique(t)) <= {-1, 0, 1} and np.all(np.abs(counts - 1 / 3) < 0.02)
    g = C.gaussian(ks, 20000, 3.19)
    assert abs(g.mean()) < 0.1 and abs(g.std() - np.sqrt(3.19**2 + 1 / 12)) < 0.1
    q = [(1 << 40) - 87, (1 << 60) - 93]
    u = C.uniform_mod(ks, q, 2, 100)
    assert u.shape == (2, 2, 100) and np.all(u[:, 0] < q[0]) and np.all(u >= 0)
    assert u[:, 1].max() > (1 << 59)


def test_default_generators_draw_fresh_entropy():
    """Without rng_seed the key and nonce come from the OS: two generators
    differ, while an explicit seed reproduces the seeded key words."""

    from fhelium_amd.csprng import Csprng, key_words_from_seed

    a, b = Csprng(None, "cpu"), Csprng(None, "cpu")
    assert list(a.key) != list(b.key)
    assert a.nonce != b.nonce
    s = Csprng(20260807, "cpu")
    assert list(s.key) == key_words_from_seed(20260807)
    assert s.nonce == (0, 0, 0)


def test_engine_rejects_non_integer_seed():
    import pytest

    from fhelium_amd.engine import CkksEngine

    with pytest.raises(TypeError):
        CkksEngine(rng_seed="7", device="cpu")
