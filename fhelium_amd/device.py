"""Python handles for the library's device runtime (csrc/runtime.hip).

* ``stream_copy``: the library's 16-byte-vector copy kernel, the byte
  reference for measured HBM ceilings;
* ``Arena``: one device allocation carved into aligned, fixed-address
  workspaces (``alloc`` returns int64 tensors viewing it) and reset as a
  whole;
* ``Event`` and ``NativeGraph``: timing events and stream-captured launch
  sequences for hosts that drive the C ABI without ``torch.cuda.CUDAGraph``.
"""

from __future__ import annotations

import ctypes

import torch

from fhelium_amd import _native as nat


def runtime_version() -> int:
    """HIP runtime version of this process (e.g. 70051831)."""

    v = ctypes.c_int(0)
    nat.call("fh_runtime_version", ctypes.byref(v))
    return int(v.value)


def stream_copy(dst: torch.Tensor, src: torch.Tensor) -> None:
    """dst <- src (same byte size, both contiguous on the device), on the
    current stream."""

    nat.require_device(dst, src)
    if not (dst.is_contiguous() and src.is_contiguous()):
        raise ValueError("stream_copy needs contiguous tensors")
    nbytes = src.numel() * src.element_size()
    if dst.numel() * dst.element_size() != nbytes:
        raise ValueError("stream_copy needs equal byte sizes")
    nat.call("fh_stream_copy", dst.data_ptr(), src.data_ptr(), nbytes, nat.stream_ptr(src.device))


class Arena:
    """A device arena on the current device.  ``alloc`` hands out int64
    tensors at fixed offsets; ``reset`` frees them all at once (the caller
    orders it after their last use)."""

    def __init__(self, nbytes: int, device: torch.device | str = "cuda") -> None:
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("an arena lives on a ROCm device")
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            nat.call("fh_arena_create", int(nbytes), ctypes.byref(handle))
        self._handle = handle
        self._storages: list = []

    def alloc(self, shape: tuple[int, ...], alignment: int = 256) -> torch.Tensor:
        numel = 1
        for s in shape:
            numel *= int(s)
        ptr = ctypes.c_void_p()
        nat.call("fh_arena_alloc", self._handle, numel * 8, int(alignment), ctypes.byref(ptr))
        return _tensor_at(int(ptr.value or 0), tuple(int(s) for s in shape), self.device, owner=self)

    def reset(self) -> None:
        nat.call("fh_arena_reset", self._handle)

    def stats(self) -> dict:
        cap, used, high = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        nat.call("fh_arena_stats", self._handle, ctypes.byref(cap), ctypes.byref(used), ctypes.byref(high))
        return {"capacity": cap.value, "used": used.value, "high_water": high.value}

    def close(self) -> None:
        if self._handle:
            torch.cuda.synchronize(self.device)
            nat.call("fh_arena_destroy", self._handle)
            self._handle = ctypes.c_void_p()

    def __enter__(self) -> "Arena":
        return self

    def __exit__(self, *exc) -> None:
        self.close()


class _DeviceBuffer:
    """__cuda_array_interface__ over a raw device address (no ownership)."""

    def __init__(self, ptr: int, shape: tuple[int, ...]) -> None:
        self.__cuda_array_interface__ = {"shape": shape, "typestr": "<i8", "data": (ptr, False), "version": 2,
                                         "strides": None}


def _tensor_at(ptr: int, shape: tuple[int, ...], device: torch.device, owner: object) -> torch.Tensor:
    t = torch.as_tensor(_DeviceBuffer(ptr, shape), device=device)
    if t.data_ptr() != ptr:
        raise RuntimeError("arena tensor was copied instead of viewed")
    t._fh_owner = owner  # keeps the arena alive while the view exists
    return t


class Event:
    def __init__(self) -> None:
        h = ctypes.c_void_p()
        nat.call("fh_event_create", ctypes.byref(h))
        self._h = h

    def record(self, stream: int | None = None) -> None:
        nat.call("fh_event_record", self._h, stream if stream is not None else torch.cuda.current_stream().cuda_stream)

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0.0)
        nat.call("fh_event_elapsed_ms", self._h, end._h, ctypes.byref(ms))
        return float(ms.value)

    def __del__(self) -> None:
        if getattr(self, "_h", None):
            nat.load().fh_event_destroy(self._h)


class NativeGraph:
    """Capture every library launch issued on ``stream`` between ``begin``
    and ``end`` (a non-default stream), then ``replay`` it."""

    def __init__(self) -> None:
        self._exec = ctypes.c_void_p()
        self._stream: int | None = None

    def begin(self, stream: torch.cuda.Stream) -> None:
        if stream.cuda_stream == 0:
            raise ValueError("capture needs a non-default stream")
        self._stream = stream.cuda_stream
        nat.call("fh_graph_begin", self._stream)

    def end(self) -> None:
        nat.call("fh_graph_end", self._stream, ctypes.byref(self._exec))

    def replay(self, stream: torch.cuda.Stream | None = None) -> None:
        s = stream.cuda_stream if stream is not None else self._stream
        nat.call("fh_graph_launch", self._exec, s)

    def __del__(self) -> None:
        if getattr(self, "_exec", None):
            nat.load().fh_graph_destroy(self._exec)
