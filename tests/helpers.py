"""Shared fixtures: configs, seeded residues and device tables for parity tests."""

from __future__ import annotations

import functools

import numpy as np
import torch

from fhelium_amd import rns
from fhelium_amd.config import CkksConfig, Preset

L34 = Preset.slots32768_scale40_levels34_int64
L7 = Preset.slots8192_scale40_levels7_int64


@functools.lru_cache(maxsize=None)
def config(log_n: int = 16, preset: Preset = L34) -> CkksConfig:
    base = CkksConfig.from_preset(preset)
    return base if log_n == base.logN else base.with_ring_log(log_n)


@functools.lru_cache(maxsize=None)
def host_tables(log_n: int = 16, preset: Preset = L34) -> rns.HostTables:
    return rns.HostTables(config(log_n, preset))


def moduli_array(cfg: CkksConfig, row_start: int = 0, rows: int | None = None) -> np.ndarray:
    m = np.array(cfg.moduli, dtype=np.int64)
    return m[row_start: None if rows is None else row_start + rows]


def random_residues(rng: np.random.Generator, batch: int, moduli: np.ndarray, n: int,
                    lazy: bool = False) -> np.ndarray:
    """[batch, len(moduli), n] uniform in [0, q) (or [0, 2q) when lazy)."""

    hi = moduli * (2 if lazy else 1)
    out = np.empty((batch, len(moduli), n), dtype=np.int64)
    for j, h in enumerate(hi):
        out[:, j, :] = rng.integers(0, int(h), size=(batch, n), dtype=np.int64)
    return out


def to_dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def to_np(t: torch.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()
