"""Launch-shape autotuner for the gfx950 kernels, with a persisted cache.

The kernels' arithmetic is fixed; what varies is how a call is cut into
launches.  For the NTT the choice is the *schedule*: one launch over every
batch entry, or chunks of a few entries issued round-robin on two or more
streams, so that one chunk's tile pass overlaps the next chunk's column pass
and a chunk's intermediate can stay in the 256 MB Infinity Cache.  Which is
fastest depends on the shape (entries x limbs x N) and the device, so it is
measured, not guessed: ``tune_ntt`` times every candidate on synthetic
residues of the requested shape with HIP events, keeps the fastest, and
records every candidate's time in a JSON cache keyed by device, ring size,
limb count, entry count (rounded down to a power of two), direction and
arithmetic mode.  ``Gfx950NttBackend`` looks the schedule up per call; a
shape that was never tuned runs as one launch.

Every schedule computes the same values (each entry is transformed by the
same kernels), so tuning never changes results, only time.

The cache file is ``$FHELIUM_AMD_TUNING_CACHE`` or
``~/.cache/fhelium_amd/tuning.json``.
"""

from __future__ import annotations

import json
import os
import pathlib
import threading
import time
from dataclasses import dataclass

import torch

from fhelium_amd import _native as nat

CACHE_VERSION = 1
OUT_STANDARD_FORM = 2  # fh_ntt_inverse output form of the tuned inverse


@dataclass(frozen=True)
class NttSchedule:
    chunk_entries: int | None  # None: one launch over every entry
    streams: int = 1

    @property
    def label(self) -> str:
        return "one launch" if self.chunk_entries is None else f"{self.chunk_entries} entries x {self.streams} streams"


ONE_LAUNCH = NttSchedule(None, 1)
NTT_CANDIDATES = (ONE_LAUNCH, NttSchedule(16, 2), NttSchedule(8, 2), NttSchedule(4, 2), NttSchedule(8, 3),
                  NttSchedule(4, 3))


def default_cache_path() -> pathlib.Path:
    env = os.environ.get("FHELIUM_AMD_TUNING_CACHE")
    return pathlib.Path(env) if env else pathlib.Path.home() / ".cache" / "fhelium_amd" / "tuning.json"


def entry_bucket(entries: int) -> int:
    return 1 << (max(1, entries).bit_length() - 1)


class TuningCache:
    """JSON-backed map from a shape key to the chosen schedule."""

    def __init__(self, path: pathlib.Path | str | None = None) -> None:
        self.path = pathlib.Path(path) if path is not None else default_cache_path()
        self._lock = threading.Lock()
        self._entries: dict[str, dict] = {}
        if self.path.exists():
            try:
                body = json.loads(self.path.read_text())
                if body.get("version") == CACHE_VERSION:
                    self._entries = dict(body.get("entries", {}))
            except (OSError, json.JSONDecodeError):
                self._entries = {}

    def get(self, key: str) -> dict | None:
        with self._lock:
            return self._entries.get(key)

    def put(self, key: str, record: dict) -> None:
        with self._lock:
            self._entries[key] = record
            self.path.parent.mkdir(parents=True, exist_ok=True)
            tmp = self.path.with_suffix(".json.tmp")
            tmp.write_text(json.dumps({"version": CACHE_VERSION, "entries": self._entries}, indent=1,
                                      sort_keys=True) + "\n")
            tmp.replace(self.path)

    def __len__(self) -> int:
        return len(self._entries)


_device_tags: dict[int, str] = {}


def device_tag(device: torch.device) -> str:
    idx = device.index if device.index is not None else torch.cuda.current_device()
    tag = _device_tags.get(idx)
    if tag is None:
        with torch.cuda.device(idx):
            tag = nat.device_info().split(" mem=")[0]
        _device_tags[idx] = tag
    return tag


def ntt_key(device: torch.device, log_n: int, limbs: int, entries: int, forward: bool, mode: int) -> str:
    return (f"ntt|{device_tag(device)}|logN={log_n}|limbs={limbs}|entries={entry_bucket(entries)}|"
            f"{'forward' if forward else 'inverse'}|mode={mode}")


def schedule_from_record(record: dict | None) -> NttSchedule:
    if not record:
        return ONE_LAUNCH
    chunk, streams = record["schedule"]
    return NttSchedule(chunk, streams)


def tune_ntt(backend, entries: int, limbs: int, row_start: int = 0, *, forward: bool = True, reps: int = 5,
             candidates=NTT_CANDIDATES) -> dict:
    """Time every candidate schedule of a [entries, limbs, N] in-place
    transform on ``backend`` (its cache receives the winner) and return the
    record stored."""

    if backend.tuning is None:
        raise ValueError("the backend has no tuning cache attached")
    if backend.moduli is None:
        raise ValueError("the backend needs its moduli to build tuning inputs")
    t = backend.tables
    n = 1 << t.log_n
    dev = t.rows.device
    moduli = backend.moduli[row_start: row_start + limbs]
    q = torch.tensor(moduli, dtype=torch.int64, device=dev).view(1, -1, 1)
    x = torch.randint(0, 1 << 62, (entries, limbs, n), dtype=torch.int64, device=dev) % q
    times = {}
    for cand in candidates:
        if cand.chunk_entries is not None and cand.chunk_entries >= entries:
            continue
        y = x.clone()
        run = (lambda s=cand: backend._apply(y, row_start, forward, 1 if forward else OUT_STANDARD_FORM, s))
        run()  # warm-up
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
        for i in range(reps):
            evs[2 * i].record()
            run()
            evs[2 * i + 1].record()
        torch.cuda.synchronize(dev)
        ms = sorted(evs[2 * i].elapsed_time(evs[2 * i + 1]) for i in range(reps))[reps // 2]
        times[cand.label] = (ms, cand)
        del y
    best_label = min(times, key=lambda k: times[k][0])
    best = times[best_label][1]
    record = {"schedule": [best.chunk_entries, best.streams], "choice": best_label,
              "candidates_ms": {k: v[0] for k, v in times.items()}, "shape": [entries, limbs, n],
              "tuned_at": time.strftime("%Y-%m-%dT%H:%M:%S")}
    backend.tuning.put(ntt_key(dev, t.log_n, limbs, entries, forward, backend._mode), record)
    return record
