"""Regenerate ``primes_v1.json`` from the reference's packaged prime catalog.

The reference ships its NTT-friendly prime sequences as two safetensors data
files (fhelium/config/resources/{scale,message}_primes_v1.safetensors, format
validated by fhelium/config/_prime_catalog.py:93-167).  Presets select the
first ``num_scale_primes`` scale primes and the first ``1 + num_p_primes``
60-bit message primes (fhelium/config/ckks.py:379-535).  This script copies
the sequences this backend supports into the package's JSON data file
(fhelium_amd/config/primes_v1.json) so the package never needs the reference
at run time.  It is run by hand, once, in a tree where the reference is
present; the JSON is committed.
"""

from __future__ import annotations

import json
import pathlib
import sys

from safetensors import safe_open

REFERENCE = pathlib.Path("/root/reference/fhelium/config/resources")
OUT = pathlib.Path(__file__).resolve().parents[2] / "fhelium_amd" / "config" / "primes_v1.json"
RING_DIMENSIONS = (1 << 14, 1 << 15, 1 << 16)


def _read(name: str) -> dict[str, list[int]]:
    with safe_open(str(REFERENCE / name), framework="pt", device="cpu") as handle:
        return {key: handle.get_tensor(key).tolist() for key in handle.keys()}


def main() -> int:
    scale = _read("scale_primes_v1.safetensors")
    message = _read("message_primes_v1.safetensors")
    out = {"scale": {}, "message": {}}
    for n in RING_DIMENSIONS:
        for bits in (30, 40, 50):
            out["scale"][f"{bits}:{n}"] = [int(p) for p in scale[f"sb={bits};N={n}"]]
        out["message"][f"60:{n}"] = [int(p) for p in message[f"sb=60;N={n}"]]
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
