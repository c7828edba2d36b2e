"""HIP-graph capture of fixed-shape engine work.

The reference replays recorded CUDA graphs for fixed-shape execution plans
(fhelium/execution/cuda_graph.py:58-560, with static buffers from
fhelium/execution/buffer.py:185-536).  On ROCm, ``torch.cuda.CUDAGraph``
records HIP graphs; the gfx950 kernels launch on PyTorch's current stream,
so every C-ABI launch of a captured call becomes a graph node, and the
workspaces the call allocates come from the graph's private memory pool
(fixed addresses across replays).

``CapturedStep(fn, *example_inputs)`` runs ``fn`` eagerly a few times on
static copies of the inputs (which fills every table and address cache, so
nothing is copied from host memory during capture), captures one call, and
then ``step(*inputs)`` copies new input payloads into the static buffers and
replays the graph.  Inputs are ``Ciphertext`` values (their level, scale,
domain and shape must match the captured ones) or device tensors.
"""

from __future__ import annotations

from dataclasses import replace
from typing import Any, Callable

import torch

from fhelium_amd.core import Ciphertext


def _static_copy(x: Any) -> Any:
    if isinstance(x, Ciphertext):
        return replace(x, data=x.data.clone())
    if isinstance(x, torch.Tensor):
        return x.clone()
    return x


def _payload(x: Any) -> torch.Tensor | None:
    if isinstance(x, Ciphertext):
        return x.data
    if isinstance(x, torch.Tensor):
        return x
    return None


def _same_state(a: Ciphertext, b: Ciphertext) -> bool:
    return (a.level == b.level and a.scale == b.scale and a.polynomial_domain == b.polynomial_domain
            and a.residue_representation == b.residue_representation and a.modulus_basis == b.modulus_basis
            and a.data.shape == b.data.shape and a.data.dtype == b.data.dtype)


class CapturedStep:
    """One captured call of ``fn`` replayable on new inputs of the same shape."""

    def __init__(self, fn: Callable[..., Any], *example_inputs: Any, warmup: int = 2) -> None:
        if not torch.cuda.is_available():
            raise RuntimeError("graph capture needs a ROCm device")
        self.fn = fn
        self.static_inputs = [_static_copy(x) for x in example_inputs]
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):
                fn(*self.static_inputs)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.static_output = fn(*self.static_inputs)

    def __call__(self, *inputs: Any) -> Any:
        """Copy the inputs' payloads into the static buffers, replay, and
        return the static output (overwritten by the next replay: clone it
        to keep it)."""

        if len(inputs) != len(self.static_inputs):
            raise ValueError("captured step called with a different number of inputs")
        for new, static in zip(inputs, self.static_inputs):
            if isinstance(static, Ciphertext):
                if not isinstance(new, Ciphertext) or not _same_state(new, static):
                    raise ValueError("replay input must match the captured ciphertext state and shape")
            src, dst = _payload(new), _payload(static)
            if dst is None:
                if new is not static and new != static:
                    raise ValueError("non-tensor inputs are fixed at capture time")
                continue
            if src is None or src.shape != dst.shape or src.dtype != dst.dtype:
                raise ValueError("replay input payload must match the captured shape and dtype")
            dst.copy_(src)
        self.graph.replay()
        return self.static_output
