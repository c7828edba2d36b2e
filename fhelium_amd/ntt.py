"""gfx950 NTT backend.

Implements the reference's ``NttBackend`` protocol
(fhelium/engine/ntt/interface.py:23-115): representation-explicit forward /
inverse transforms over dense ``[*batch, limb, N]`` operands whose limbs are a
contiguous interval of the canonical QP rows starting at
``parameter_row_start``.

The backend owns its table layout (the protocol leaves layout to backends):
per QP row a 128-byte record (modulus, binary64 reciprocal, conversion
constants) plus standard-form twiddles in the compact bit-reversed order
``psi^{br(k)}``; rows of primes below 2^42 keep the twiddle as a binary64
bit pattern, larger primes keep ``(w, floor(w 2^64 / q))`` Shoup pairs.  The
tables are derived on the device from the reference-format inputs
(``rns_params`` and compact Montgomery twiddles) by ``fh_ntt_prepare_tables``.
"""

from __future__ import annotations

import torch

from fhelium_amd import _native as nat

NTT_ROW_WORDS = 16  # sizeof(NttRow) / 8

# inverse output forms (fh_ntt_inverse)
OUT_MONTGOMERY = 0
OUT_STANDARD_LAZY = 1
OUT_STANDARD = 2
OUT_CENTERED = 3


class NttTables:
    """Device tables for one direction pair over a set of QP rows."""

    def __init__(self, rns_params: torch.Tensor, forward_mont: torch.Tensor,
                 inverse_mont: torch.Tensor, log_n: int) -> None:
        nat.require_device(rns_params, forward_mont, inverse_mont)
        rows = rns_params.size(1)
        n = 1 << log_n
        if forward_mont.shape != (rows, n) or inverse_mont.shape != (rows, n):
            raise ValueError("twiddle tables must be [rows, N] matching rns_params")
        dev = rns_params.device
        self.log_n = log_n
        self.rows = torch.zeros((rows, NTT_ROW_WORDS), dtype=torch.int64, device=dev)
        self.fwd = torch.empty((rows, n), dtype=torch.int64, device=dev)
        self.fwd_shoup = torch.empty((rows, n), dtype=torch.int64, device=dev)
        self.inv = torch.empty((rows, n), dtype=torch.int64, device=dev)
        self.inv_shoup = torch.empty((rows, n), dtype=torch.int64, device=dev)
        s = nat.stream_ptr(dev)
        for tw, w, wp in ((forward_mont, self.fwd, self.fwd_shoup), (inverse_mont, self.inv, self.inv_shoup)):
            nat.call("fh_ntt_prepare_tables", rns_params.data_ptr(), rns_params.stride(0),
                     rns_params.stride(1), rows, log_n, tw.data_ptr(), tw.stride(0),
                     self.rows.data_ptr(), w.data_ptr(), wp.data_ptr(), s)


MODE_EXACT = 0
MODE_FAST = 1


class Gfx950NttBackend:
    """The ``NttBackend`` protocol on gfx950 (name ``gfx950_radix16_lds``).

    ``mode="exact"`` reproduces the reference's lazy forward outputs bit for
    bit; ``mode="fast"`` returns canonical forward outputs (equal to the
    reference's modulo q) at a fraction of the arithmetic.  Inverse outputs
    are identical in both modes.
    """

    name = "gfx950_radix16_lds"

    def __init__(self, tables: NttTables, mode: str = "fast", *, moduli: tuple[int, ...] | None = None,
                 tuning=None) -> None:
        if mode not in ("fast", "exact"):
            raise ValueError("mode must be 'fast' or 'exact'")
        self.tables = tables
        self.mode = mode
        self._mode = MODE_FAST if mode == "fast" else MODE_EXACT
        self.moduli = moduli
        # Optional autotune.TuningCache: per-shape launch schedules.  Without
        # one (or for an untuned shape) every transform is one launch.
        self.tuning = tuning
        self._side_streams: dict[int, list] = {}
        # Optional list receiving (start_event, end_event, algorithmic_bytes)
        # per transform, recorded on the operand's stream (roofline probes).
        self.recorder: list | None = None

    def _check(self, operand: torch.Tensor, row_start: int) -> torch.Tensor:
        nat.require_device(operand)
        if operand.dtype != torch.int64:
            raise TypeError("the gfx950 backend supports int64 residues")
        t = self.tables
        if operand.size(-1) != (1 << t.log_n):
            raise ValueError(f"operand ring dimension {operand.size(-1)} != {1 << t.log_n}")
        if row_start < 0 or row_start + operand.size(-2) > t.rows.size(0):
            raise ValueError(
                "NTT operand rows exceed the prepared canonical parameter range: "
                f"start={row_start}, rows={operand.size(-2)}, table_rows={t.rows.size(0)}"
            )
        view = nat.batch_view(operand, "residues")
        if view.stride(2) != 1:
            raise ValueError("NTT operands need a unit coefficient stride")
        return view

    def schedule_for(self, entries: int, limbs: int, forward: bool, device: torch.device):
        """The tuned launch schedule of this shape (one launch if untuned)."""

        from fhelium_amd import autotune

        if self.tuning is None or entries < 2:
            return autotune.ONE_LAUNCH
        key = autotune.ntt_key(device, self.tables.log_n, limbs, entries, forward, self._mode)
        return autotune.schedule_from_record(self.tuning.get(key))

    def _streams(self, device: torch.device, count: int) -> list:
        idx = device.index if device.index is not None else torch.cuda.current_device()
        pool = self._side_streams.setdefault(idx, [])
        while len(pool) < count:
            pool.append(torch.cuda.Stream(device=idx))
        return pool[:count]

    def _launch(self, fn: str, dst: torch.Tensor, src: torch.Tensor | None, row_start: int, form: int) -> None:
        t = self.tables
        fwd = fn.startswith("fh_ntt_forward")
        w = (t.fwd if fwd else t.inv).data_ptr()
        wp = (t.fwd_shoup if fwd else t.inv_shoup).data_ptr()
        s = nat.stream_ptr(dst.device)
        if src is None:
            nat.call(fn, dst.data_ptr(), dst.size(0), dst.size(1), t.log_n, dst.stride(0), dst.stride(1),
                     t.rows.data_ptr(), w, wp, row_start, form, self._mode, s)
        else:
            nat.call(fn, dst.data_ptr(), src.data_ptr(), src.size(0), src.size(1), t.log_n, dst.stride(0),
                     dst.stride(1), src.stride(0), src.stride(1), t.rows.data_ptr(), w, wp, row_start, form,
                     self._mode, s)

    def _apply(self, view: torch.Tensor, row_start: int, forward: bool, form: int, schedule=None,
               src: torch.Tensor | None = None) -> None:
        """Transform the [B, limb, N] ``view`` (in place, or from ``src``)
        with the given schedule: one launch, or chunks of entries issued
        round-robin on side streams forked from and joined to the current
        stream.  Results do not depend on the schedule."""

        fn = ("fh_ntt_forward" if forward else "fh_ntt_inverse") + ("_from" if src is not None else "")
        if schedule is None:
            schedule = self.schedule_for(view.size(0), view.size(1), forward, view.device)
        chunk = schedule.chunk_entries
        if chunk is None or chunk >= view.size(0):
            self._launch(fn, view, src, row_start, form)
            return
        cur = torch.cuda.current_stream(view.device)
        streams = self._streams(view.device, schedule.streams)
        for st in streams:
            st.wait_stream(cur)
        for i, b in enumerate(range(0, view.size(0), chunk)):
            with torch.cuda.stream(streams[i % len(streams)]):
                self._launch(fn, view[b: b + chunk], None if src is None else src[b: b + chunk], row_start, form)
        for st in streams:
            cur.wait_stream(st)

    def _run(self, fn: str, operand: torch.Tensor, row_start: int, mode: int, schedule=None) -> None:
        view = self._check(operand, row_start)
        rec = self.recorder
        if rec is not None:
            start = torch.cuda.Event(enable_timing=True)
            end = torch.cuda.Event(enable_timing=True)
            start.record()
        self._apply(view, row_start, fn == "fh_ntt_forward", mode, schedule)
        if rec is not None:
            end.record()
            # algorithmic bytes: every residue read once and written once
            rec.append((start, end, 2 * view.size(0) * view.size(1) * view.size(2) * 8))

    # -- protocol --------------------------------------------------------
    def forward_montgomery_(self, operand: torch.Tensor, parameter_row_start: int) -> None:
        self._run("fh_ntt_forward", operand, parameter_row_start, 0)

    def forward_to_montgomery_(self, operand: torch.Tensor, parameter_row_start: int) -> None:
        self._run("fh_ntt_forward", operand, parameter_row_start, 1)

    def forward_to_montgomery(self, operand: torch.Tensor, parameter_row_start: int) -> torch.Tensor:
        return self._run_from("fh_ntt_forward_from", operand, parameter_row_start, 1)

    def inverse_montgomery_(self, operand: torch.Tensor, parameter_row_start: int) -> None:
        self._run("fh_ntt_inverse", operand, parameter_row_start, OUT_MONTGOMERY)

    def inverse_to_standard_lazy_(self, operand: torch.Tensor, parameter_row_start: int) -> None:
        self._run("fh_ntt_inverse", operand, parameter_row_start, OUT_STANDARD_LAZY)

    def inverse_to_standard_(self, operand: torch.Tensor, parameter_row_start: int) -> None:
        self._run("fh_ntt_inverse", operand, parameter_row_start, OUT_STANDARD)

    def inverse_to_centered_(self, operand: torch.Tensor, parameter_row_start: int) -> None:
        self._run("fh_ntt_inverse", operand, parameter_row_start, OUT_CENTERED)

    def _run_from(self, fn: str, operand: torch.Tensor, row_start: int, form: int) -> torch.Tensor:
        """Out-of-place transform into a new contiguous tensor; ``operand``
        is left unchanged (the first pass reads it, so no clone is made)."""

        src = self._check(operand, row_start)
        out = torch.empty(operand.shape, dtype=operand.dtype, device=operand.device)
        dst = out.view(src.shape)
        rec = self.recorder
        if rec is not None:
            start = torch.cuda.Event(enable_timing=True)
            end = torch.cuda.Event(enable_timing=True)
            start.record()
        self._apply(dst, row_start, fn == "fh_ntt_forward_from", form, None, src=src)
        if rec is not None:
            end.record()
            rec.append((start, end, 2 * src.size(0) * src.size(1) * src.size(2) * 8))
        return out

    def inverse_to_standard(self, operand: torch.Tensor, parameter_row_start: int) -> torch.Tensor:
        """Out-of-place inverse to canonical standard residues."""

        return self._run_from("fh_ntt_inverse_from", operand, parameter_row_start, OUT_STANDARD)
