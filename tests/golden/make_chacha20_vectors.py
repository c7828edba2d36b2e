"""Write chacha20_vectors.json: ChaCha20 keystream blocks computed by OpenSSL
(`openssl enc -chacha20`, RFC 8439 block function; OpenSSL's 16-byte IV is
the 32-bit little-endian block counter followed by the 96-bit nonce).

Run from the repository root: python tests/golden/make_chacha20_vectors.py
"""

from __future__ import annotations

import json
import pathlib
import subprocess

CASES = [
    # RFC 8439 section 2.3.2 key and nonce, counters 0..3
    (bytes(range(32)).hex(), "000000090000004a00000000", 0, 4),
    # RFC 8439 section 2.4.2 nonce
    (bytes(range(32)).hex(), "000000000000004a00000000", 1, 2),
    # an all-zero key and nonce (RFC 8439 appendix A.1 inputs)
    ("00" * 32, "00" * 12, 0, 2),
    # arbitrary key, nonce with a high stream word, counter near the top
    ("c0ffee" * 10 + "0102", "0123456789abcdef00000007", 0xFFFFFFFC, 4),
]


def keystream(key_hex: str, nonce_hex: str, counter: int, blocks: int) -> str:
    iv = counter.to_bytes(4, "little").hex() + nonce_hex
    out = subprocess.run(["openssl", "enc", "-chacha20", "-K", key_hex, "-iv", iv, "-nosalt", "-nopad"],
                         input=bytes(64 * blocks), capture_output=True, check=True).stdout
    return out.hex()


def main() -> None:
    vectors = [{"key": k, "nonce": n, "counter": c, "blocks": b, "keystream": keystream(k, n, c, b)}
               for k, n, c, b in CASES]
    path = pathlib.Path(__file__).with_name("chacha20_vectors.json")
    path.write_text(json.dumps({"source": "openssl enc -chacha20", "vectors": vectors}, indent=1) + "\n")


if __name__ == "__main__":
    main()
