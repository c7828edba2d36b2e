"""ctypes binding of the gfx950 C ABI (``include/fhelium_amd.h``).

The shared library is built in-tree by :mod:`fhelium_amd.build`.  There is no
CPU fallback: if the library is missing, cannot be loaded, or a device
operand is not on a ROCm device, every entry point raises.  ``import torch``
must precede loading so the library binds to PyTorch's HIP runtime (the same
``libamdhip64`` soname), which makes torch device pointers and streams valid
arguments.
"""

from __future__ import annotations

import ctypes
import pathlib
import threading

import torch

LIB_PATH = pathlib.Path(__file__).resolve().parent / "_lib" / "libfhelium_amd.so"
ABI_VERSION = 7


class NativeExtensionError(RuntimeError):
    """The gfx950 library is missing, stale or failed to load."""


class fh_view3(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("stride_batch", ctypes.c_int64),
        ("stride_limb", ctypes.c_int64),
        ("stride_coef", ctypes.c_int64),
    ]


class fh_table2(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("stride_row", ctypes.c_int64),
        ("stride_col", ctypes.c_int64),
    ]


class fh_vec(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("stride", ctypes.c_int64)]


_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_ptr = ctypes.c_void_p

# name -> argument types; every function returns int status.
SIGNATURES: dict[str, list] = {
    "fh_ntt_forward": [_ptr, _i64, _i64, _i64, _i64, _i64, _ptr, _ptr, _ptr, _i64, _i32, _i32, _ptr],
    "fh_ntt_inverse": [_ptr, _i64, _i64, _i64, _i64, _i64, _ptr, _ptr, _ptr, _i64, _i32, _i32, _ptr],
    "fh_ntt_forward_from": [_ptr, _ptr, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _ptr, _ptr, _ptr, _i64,
                            _i32, _i32, _ptr],
    "fh_ntt_inverse_from": [_ptr, _ptr, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _ptr, _ptr, _ptr, _i64,
                            _i32, _i32, _ptr],
    "fh_ntt_indexed_stage": [fh_view3, _i64, _i64, _ptr, _ptr, _i64, _ptr, _i64, _i64, fh_table2, _i64, _i32, _ptr],
    "fh_ntt_prepare_tables": [_ptr, _i64, _i64, _i64, _i64, _ptr, _i64, _ptr, _ptr, _ptr, _ptr],
    "fh_rns_binary": [_i32, fh_view3, fh_view3, fh_view3, fh_table2, _i64, _i64, _i64, _i64, _ptr],
    "fh_rns_unary_inplace": [_i32, fh_view3, fh_table2, _i64, _i64, _i64, _ptr],
    "fh_rns_row_scalars": [fh_view3, fh_view3, fh_vec, fh_table2, _i64, _i64, _i64, _i32, _ptr],
    "fh_rns_add_lazy_twice_modulus": [fh_view3, fh_view3, fh_view3, fh_vec, _i64, _i64, _i64, _ptr],
    "fh_rns_lift_centered": [fh_view3, fh_view3, fh_vec, _i64, _i64, _i64, _ptr],
    "fh_rns_mixed_radix_decompose": [
        fh_view3, fh_view3, _i64, _i64, _i64, fh_vec, fh_table2, fh_vec, fh_vec, fh_vec, fh_vec, _ptr,
    ],
    "fh_rns_basis_extend": [fh_view3, fh_view3, _i64, _i64, _i64, _i64, fh_table2, fh_table2, _ptr],
    "fh_rns_basis_extend_mfma": [fh_view3, fh_view3, _i64, _i64, _i64, _i64, fh_table2, fh_table2, _ptr],
    "fh_ckks_rescale": [fh_view3, fh_view3, fh_view3, fh_vec, fh_table2, _i64, _i64, _i64, _i64, _i32, _ptr],
    "fh_ckks_galois_coefficient": [fh_view3, fh_view3, _ptr, _i64, _ptr, _i64, fh_vec, _i64, _i64, _i64, _ptr],
    "fh_ckks_galois_ntt": [fh_view3, fh_view3, _ptr, _i64, _i64, _i64, _i64, _ptr],
    "fh_ckks_keyswitch_accumulate": [fh_view3, fh_view3, fh_view3, fh_view3, fh_table2, _i64, _i64, _i64, _i64, _ptr],
    "fh_ckks_moddown": [fh_view3, fh_view3, fh_view3, fh_table2, fh_table2, _i64, _i64, _i64, _i64, _ptr],
    "fh_ckks_add_plaintext": [_i32, fh_view3, fh_view3, fh_view3, _ptr, _i64, fh_table2, _i64, _i64, _i64, _i64, _ptr],
    "fh_ckks_tensor_product": [fh_view3] * 7 + [fh_table2, _i64, _i64, _i64, _ptr],
    "fh_ckks_scalar_combination": [_ptr, _i64, _i64, _i64, _ptr, _ptr, _ptr, _i64, _i64, _i64, _i64, _i64, _i64,
                                   fh_table2, _ptr],
    "fh_ckks_plaintext_inner_product": [_ptr, _i64, _i64, _ptr, _i64, _i64, _i64, _ptr, _i64, _i64, _i64, _i64,
                                        _i64, _i64, fh_table2, _ptr],
    "fh_ckks_keyswitch_accumulate_multi_key": [fh_view3, fh_view3, fh_view3, _ptr, _i64, _i64, _i64, _i64,
                                               fh_table2, _i64, _i64, _i64, _i64, _ptr],
    "fh_keyswitch_fused": [fh_view3, _i64, _i64, _i64, _i64, _i64, _i64, _ptr, _i64, _ptr, _ptr, fh_table2, _ptr,
                           _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _i64, _i64, _ptr, _ptr, _ptr, _ptr,
                           _ptr, fh_view3, fh_view3, fh_view3, fh_view3, _i32, _ptr, _ptr, _i64, _ptr, _ptr,
                           _ptr, _ptr],
    "fh_keyswitch_fused_static": [fh_view3, _i64, _i64, _i64, _i64, _i64, _i64, _ptr, _i64, _ptr, _ptr, fh_table2,
                                  _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _i64, _i64, _ptr, _ptr,
                                  fh_view3, fh_view3, fh_view3, fh_view3, _i32, _ptr, _ptr, _i64, _ptr, _ptr,
                                  _ptr, _ptr],
    "fh_keyswitch_hoisted": [fh_view3, _i64, _i64, _i64, _i64, _i64, _i64, _ptr, _i64, _ptr, _ptr, fh_table2,
                             _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _i64, _i64, _ptr, _ptr, _ptr, _ptr, _ptr,
                             _ptr, fh_view3, fh_view3, fh_view3, _i32, _ptr],
    "fh_runtime_version": [_ptr],
    "fh_stream_copy": [_ptr, _ptr, _i64, _ptr],
    "fh_arena_create": [_i64, _ptr],
    "fh_arena_alloc": [_ptr, _i64, _i64, _ptr],
    "fh_arena_reset": [_ptr],
    "fh_arena_stats": [_ptr, _ptr, _ptr, _ptr],
    "fh_arena_destroy": [_ptr],
    "fh_event_create": [_ptr],
    "fh_event_record": [_ptr, _ptr],
    "fh_event_elapsed_ms": [_ptr, _ptr, _ptr],
    "fh_event_destroy": [_ptr],
    "fh_graph_begin": [_ptr],
    "fh_graph_end": [_ptr, _ptr],
    "fh_graph_launch": [_ptr, _ptr],
    "fh_graph_destroy": [_ptr],
    "fh_chacha20_keystream": [_ptr, _ptr, ctypes.c_uint32, _i64, _ptr, _ptr],
    "fh_sample_uniform_mod": [_ptr, _ptr, ctypes.c_uint32, _ptr, _i64, _i64, _i64, _ptr, _ptr],
    "fh_sample_ternary": [_ptr, _ptr, ctypes.c_uint32, _i64, _ptr, _ptr],
    "fh_sample_gaussian": [_ptr, _ptr, ctypes.c_uint32, _i64, ctypes.c_double, _ptr, _ptr],
    "fh_ckks_encode": [_ptr, _i64, _i64, _ptr, ctypes.c_double, _ptr, _i64, _ptr, _ptr, _ptr],
    "fh_ckks_decode": [_ptr, _ptr, _i64, _i64, _i64, _i64, _i64, _i64, _ptr, ctypes.c_double, _ptr, _ptr, _ptr],
    "fh_ckks_hmult_tensor_coefficient": [_ptr, _ptr, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i64,
                                         _i64, _i64, fh_table2, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr,
                                         _i32, _ptr],
}

_lock = threading.Lock()
_lib: ctypes.CDLL | None = None


def load() -> ctypes.CDLL:
    """Load (once) and return the library; raise NativeExtensionError if absent."""

    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise NativeExtensionError(
                f"gfx950 library not built: {LIB_PATH} is missing; run "
                "`python -m fhelium_amd.build`"
            )
        from fhelium_amd import manifest

        mpath = manifest.manifest_path(LIB_PATH.parent)
        if not mpath.exists():
            raise NativeExtensionError(
                f"{LIB_PATH} has no build manifest ({mpath.name}); rebuild with `python -m fhelium_amd.build`")
        try:
            problems = manifest.compatibility_mismatches(manifest.read_manifest(mpath), lib_path=LIB_PATH,
                                                         library_abi=ABI_VERSION)
        except ValueError as exc:
            raise NativeExtensionError(str(exc)) from exc
        if problems:
            raise NativeExtensionError(f"{LIB_PATH} does not match this process: " + "; ".join(problems))
        try:
            lib = ctypes.CDLL(str(LIB_PATH))
        except OSError as exc:  # pragma: no cover - depends on the host
            raise NativeExtensionError(f"cannot load {LIB_PATH}: {exc}") from exc
        lib.fh_last_error.restype = ctypes.c_char_p
        lib.fh_last_error.argtypes = []
        lib.fh_abi_version.restype = ctypes.c_int
        lib.fh_device_info.argtypes = [ctypes.c_char_p, _i64]
        lib.fh_device_info.restype = ctypes.c_int
        if lib.fh_abi_version() != ABI_VERSION:
            raise NativeExtensionError(
                f"ABI mismatch: library reports {lib.fh_abi_version()}, "
                f"package expects {ABI_VERSION}"
            )
        for name, argtypes in SIGNATURES.items():
            fn = getattr(lib, name, None)
            if fn is None:
                raise NativeExtensionError(f"{LIB_PATH} does not export {name}")
            fn.argtypes = argtypes
            fn.restype = ctypes.c_int
        _lib = lib
        return lib


def exported_symbols() -> list[str]:
    """Names the package binds (every one must exist in the library)."""

    return sorted(SIGNATURES)


def call(name: str, *args) -> None:
    """Invoke an entry point and raise RuntimeError on a non-zero status."""

    lib = load()
    status = getattr(lib, name)(*args)
    if status != 0:
        message = lib.fh_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed ({status}): {message}")


def device_info() -> str:
    lib = load()
    buf = ctypes.create_string_buffer(256)
    status = lib.fh_device_info(buf, 256)
    if status != 0:
        raise RuntimeError(lib.fh_last_error().decode(errors="replace"))
    return buf.value.decode()


# ---------------------------------------------------------------------------
# Tensor -> ABI argument helpers.
# ---------------------------------------------------------------------------


def require_device(*tensors: torch.Tensor) -> None:
    for t in tensors:
        if t is None:
            continue
        if t.device.type != "cuda":
            raise RuntimeError(
                "the gfx950 backend requires ROCm device tensors "
                f"(got a tensor on {t.device})"
            )


def stream_ptr(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def batch_view(t: torch.Tensor, name: str) -> torch.Tensor:
    """[..., limb, N] -> metadata-only [B, limb, N] view (no copy)."""

    if t.dim() < 2:
        raise RuntimeError(f"{name} must have [..., limb, coefficient] layout; got rank {t.dim()}")
    if t.size(-2) <= 0 or t.size(-1) <= 0:
        raise RuntimeError(f"{name} must contain at least one limb and coefficient")
    if any(s <= 0 for s in t.shape[:-2]):
        raise RuntimeError(f"{name} batch dimensions must be nonzero")
    return t.view(-1, t.size(-2), t.size(-1))


def coef_batch_view(t: torch.Tensor, name: str) -> torch.Tensor:
    if t.dim() < 1 or t.size(-1) <= 0:
        raise RuntimeError(f"{name} must have [..., coefficient] layout")
    return t.view(-1, t.size(-1))


def v3(t: torch.Tensor, broadcast_batch: bool = False) -> fh_view3:
    s0, s1, s2 = t.stride()
    if broadcast_batch:
        s0 = 0
    return fh_view3(t.data_ptr(), s0, s1, s2)


def v3_from_2d(t: torch.Tensor) -> fh_view3:
    """[B, N] coefficient view as [B, 1, N] with a zero limb stride."""

    return fh_view3(t.data_ptr(), t.stride(0), 0, t.stride(1))


def t2(t: torch.Tensor | None) -> fh_table2:
    if t is None or t.numel() == 0:
        return fh_table2(None, 0, 0)
    if t.dim() != 2:
        raise RuntimeError("expected a rank-two table")
    return fh_table2(t.data_ptr(), t.stride(0), t.stride(1))


def vec(t: torch.Tensor | None) -> fh_vec:
    if t is None or t.numel() == 0:
        return fh_vec(None, 0)
    if t.dim() != 1:
        raise RuntimeError("expected a rank-one vector")
    return fh_vec(t.data_ptr(), t.stride(0))
