"""Continuous-batching serving loop for encrypted evaluations on one GPU.

The reference keeps values resident under an HBM budget
(fhelium/residency/manager.py:103-200, 642-720) and replays fixed-shape
plans (fhelium/execution/cuda_graph.py:58-560).  This module is the serving
tier built on ``CkksEngine``:

* A request is one ciphertext ``[2, Q_l, N]`` (coefficient/standard) and an
  arrival time.  Its work is a *program*: a fixed list of stages, each a few
  engine calls.
* **Admission.**  Arrived requests are admitted in arrival order while fewer
  than ``max_active`` are active and their estimated HBM footprint fits the
  memory budget.  The budget defaults to the device's free memory plus
  PyTorch's cached-but-unused blocks, minus a reserve, read when the loop
  starts.
* **Continuous batching.**  Every iteration groups the active requests by
  the program's batch key -- (stage, starting level) for most stages, one
  tensor shape per group; the rotate-and-add stages of every request at one
  level form one group whatever their step (per-entry rotation keys) -- and
  runs each group as batched engine calls of at most ``max_batch`` entries.  A
  request admitted while others are part-way through their programs joins
  the next iteration, so its stages run in the same iteration as theirs:
  batched with them where the stage and level agree, back to back on the
  stream otherwise.
* **Resident keys.**  The relinearization key and the rotation keys live in
  the engine and are shared by every request and batch entry.
* **Latency.**  A request completes at an event recorded on the stream after
  its last stage.  Events are anchored to the wall clock by an event
  recorded when the loop starts, so latency = completion - arrival on one
  clock.  The host waits for the device only to bound the iterations in
  flight (``max_inflight``).

Batching never changes a request's arithmetic: every kernel treats batch
entries independently, so a result is bit-identical alone and in any mix.
"""

from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Sequence

import numpy as np
import torch

from fhelium_amd.core import Ciphertext, Plaintext


# ---------------------------------------------------------------------------
# Batched helpers over lists of unbatched ciphertexts
# ---------------------------------------------------------------------------
def _check_batchable(cts: Sequence[Ciphertext], check_scale: bool = True) -> Ciphertext:
    first = cts[0]
    for c in cts:
        if (c.level, c.polynomial_domain, c.residue_representation, c.modulus_basis) != (
                first.level, first.polynomial_domain, first.residue_representation,
                first.modulus_basis) or c.data.shape != first.data.shape or (check_scale and c.scale != first.scale):
            raise ValueError("batched entries must share level, scale, state and shape")
    return first


def _uniform_view(cts: Sequence[Ciphertext]) -> torch.Tensor | None:
    """The [2, B, Q, N] view whose entries are exactly ``cts`` when they lie
    in one storage at a constant entry stride with equal strides (entries of
    one earlier batched result, in order), else None."""

    first = cts[0].data
    if first.dim() != 3:
        return None
    storage = first.untyped_storage().data_ptr()
    step = (cts[1].data.storage_offset() - first.storage_offset()) if len(cts) > 1 else first.size(1) * first.size(2)
    if step < first.size(1) * first.stride(1):
        return None
    for i, c in enumerate(cts):
        d = c.data
        if (d.untyped_storage().data_ptr() != storage or d.stride() != first.stride()
                or d.storage_offset() != first.storage_offset() + i * step):
            return None
    st0, st1, st2 = first.stride()
    return first.as_strided((first.size(0), len(cts), first.size(1), first.size(2)), (st0, step, st1, st2),
                            first.storage_offset())


def stack(cts: Sequence[Ciphertext], *, scale: float | None = None) -> Ciphertext:
    """[2, Q, N] x B -> one [2, B, Q, N] ciphertext (all states must agree;
    with ``scale`` given the entries may carry different scales and the
    batch is labelled with it).  Entries that already form one contiguous
    batched tensor (a whole earlier batched result, in order) are returned
    as that tensor; otherwise they are copied into a new one."""

    first = _check_batchable(cts, check_scale=scale is None)
    view = _uniform_view(cts)
    data = view if view is not None and view.is_contiguous() else torch.stack([c.data for c in cts], dim=1)
    return Ciphertext(data, first.level, first.scale if scale is None else float(scale), first.polynomial_domain,
                      first.residue_representation, first.modulus_basis)


def batch_view(cts: Sequence[Ciphertext]) -> Ciphertext:
    """Like ``stack`` but for read-only kernel operands: any constant entry
    stride in one storage is used in place (the result may be a
    non-contiguous view); other entries are copied."""

    first = _check_batchable(cts)
    view = _uniform_view(cts)
    data = view if view is not None else torch.stack([c.data for c in cts], dim=1)
    return Ciphertext(data, first.level, first.scale, first.polynomial_domain, first.residue_representation,
                      first.modulus_basis)


def unstack(ct: Ciphertext) -> list[Ciphertext]:
    return [Ciphertext(ct.data[:, b], ct.level, ct.scale, ct.polynomial_domain, ct.residue_representation,
                       ct.modulus_basis) for b in range(ct.data.size(1))]


def chebyshev_to_power(coefficients: Sequence[float]) -> np.ndarray:
    """a with sum_k c_k T_k(y) = sum_j a_j y^j (exact linear map, binary64)."""

    return np.polynomial.chebyshev.cheb2poly(np.asarray(coefficients, dtype=np.float64))


# ---------------------------------------------------------------------------
# Program: a polynomial followed by a window sum over slots
# ---------------------------------------------------------------------------
class PolynomialWindowSum:
    """p(x) = sum_{k<=7} c_k T_k(x / radius) for |x| <= radius, then
    s_j = sum_{i < window} p_{(j - i) mod S}.

    The Chebyshev series is evaluated in the power basis of y = x / radius
    (``chebyshev_to_power``) with the depth-optimal product tree
    y2 = y y, {y3 = y y2, y4 = y2 y2}, {y5 = y y4, y6 = y3 y3, y7 = y3 y4};
    each term a_j y^j is a plaintext-constant product whose encoding scale is
    chosen so that every product carries (default scale) x (the prime the
    next rescale drops): the seven products sum in one pass and share one
    rescale, landing on the engine's default scale.  The products of one
    stage run as one batched HMult.  The window sum is log2(window)
    rotate-and-add steps by 1, 2, 4, ...  A request at level l ends at
    level l + 5.

    Steps: 0 scale by 1/radius; 1 y2; 2 y3, y4; 3 y5, y6, y7; 4 the seven
    terms and a_0; 5 .. 5 + log2(window) - 1 the rotate-and-add steps.  The
    serving loop advances every active request by one *stage* per
    iteration; ``stage_plan`` groups the steps into stages:

      "fine"   one step per stage (the rotate-and-add stages of all
               requests at one level batch together whatever their step);
      "split"  the polynomial (steps 0-4) and the window sum (the rotation
               steps) as two stages; window sums at one level batch
               together across requests that entered at different times;
      "whole"  the whole program as one stage.

    Every request runs the same dependent steps, so a coarser plan takes
    fewer iterations per request at the same batch sizes (a stage's batch is
    the requests admitted together either way).  At 200 requests/s on one
    MI355X (logN=16, L34) the p99 latency over 600 and 2400 requests was
    41 and 58 ms (fine), 31 and 42 ms (split), 32 and 38 ms (whole, the
    default), with the same throughput and capacity.
    """

    PLANS = ("fine", "split", "whole")

    depth = 5

    def __init__(self, engine, chebyshev_coefficients: Sequence[float], *, radius: float = 4.0,
                 window: int = 64, stage_plan: str = "whole") -> None:
        c = np.asarray(chebyshev_coefficients, dtype=np.float64)
        if c.ndim != 1 or not 1 <= c.size <= 8:
            raise ValueError("need 1..8 Chebyshev coefficients (degree <= 7)")
        if window < 1 or window & (window - 1):
            raise ValueError("window must be a power of two")
        if window > engine.config.slots:
            raise ValueError("window exceeds the slot count")
        self.engine = engine
        self.chebyshev = c
        self.power = np.zeros(8)
        self.power[: c.size] = chebyshev_to_power(c)
        self.radius = float(radius)
        self.window = int(window)
        self.rotation_steps = [1 << k for k in range(int(math.log2(window)))]
        steps = list(range(5 + len(self.rotation_steps)))
        if stage_plan == "fine":
            self._plan = [[k] for k in steps]
        elif stage_plan == "split":
            self._plan = [steps[:5], steps[5:]] if len(steps) > 5 else [steps]
        elif stage_plan == "whole":
            self._plan = [steps]
        else:
            raise ValueError(f"stage_plan must be one of {self.PLANS}")
        self.stage_plan = stage_plan
        self.stage_count = len(self._plan)
        self._plaintexts: dict[tuple, Plaintext] = {}
        for s in self.rotation_steps:
            engine.rotation_key(s)
        if engine.relinearization_key is None:
            engine.generate_relinearization_key()

    # -- reference values --------------------------------------------------
    def expected(self, x: np.ndarray) -> np.ndarray:
        """Binary64 cleartext result for slot vector(s) x."""

        p = np.polynomial.chebyshev.chebval(np.asarray(x, dtype=np.float64) / self.radius, self.chebyshev)
        out = np.zeros_like(p)
        for i in range(self.window):
            out += np.roll(p, i, axis=-1)
        return out

    def max_start_level(self) -> int:
        return self.engine.config.public_level_count - 1 - self.depth

    def footprint_bytes(self, level: int) -> int:
        """HBM estimate of one active request at start level ``level``: ten
        live ciphertexts, plus its three entries of a stage's batched HMult
        (operand copies, key-switch workspaces and outputs)."""

        cfg = self.engine.config
        limb = cfg.N * 8
        lq = cfg.num_q_primes - level
        lqp = lq + cfg.num_p_primes
        dnum = len(self.engine.layout.digit_specs(level))
        return limb * (10 * 2 * lq + 3 * (2 * 2 * lq + lq + dnum * lqp + 2 * lqp + 2 * lq))

    # -- constants -----------------------------------------------------------
    def _dropped_prime(self, level: int) -> int:
        return self.engine.config.moduli[self.engine.runtime.row_start(level)]

    # -- batched building blocks --------------------------------------------
    def _hmult(self, lhs: list[Ciphertext], rhs: list[Ciphertext] | None) -> list[Ciphertext]:
        """Batched HMult + relinearize + rescale: the engine's fused HMult
        (bit-identical to NTT, multiply, relinearize, rescale in turn)."""

        # The residues of a product do not depend on the scales, so entries
        # of different scales share one batch (labelled 1.0) and each output
        # takes lhs.scale * rhs.scale / q, the scale the engine gives it.
        rhs = lhs if rhs is None else rhs
        a = stack(lhs, scale=1.0)
        b = a if rhs is lhs else stack(rhs, scale=1.0)
        outs = unstack(self.engine.hmult(a, b))
        q = float(self._dropped_prime(lhs[0].level))
        for o, x, y in zip(outs, lhs, rhs):
            o.scale = (x.scale * y.scale) / q
        return outs

    def _scaled_sum(self, terms: list[list[Ciphertext]], values: list[float], target_scale: float) -> Ciphertext:
        """Batched sum_t values[t] * terms[t], rescaled once, landing exactly
        on ``target_scale``.  Each constant is the plaintext constant whose
        encoding scale puts its product at target_scale * q (q the prime the
        rescale drops), so the products share a scale and sum before the
        rescale; the terms are read at the highest level among them
        (engine.multiply_constants_and_add: coefficient domain, no NTT)."""

        eng = self.engine
        batches = [batch_view(t) for t in terms]
        q = self._dropped_prime(max(b.level for b in batches))
        acc = eng.multiply_constants_and_add(batches, values, product_scale=target_scale * q)
        out = eng.rescale_to_next_level(acc)
        return eng.reinterpret_at_scale(out, target_scale, max_relative_change=1e-12)

    def _at(self, cts: list[Ciphertext], level: int) -> list[Ciphertext]:
        return [c if c.level == level else self.engine.mod_switch_to_level(c, level) for c in cts]

    # -- batching ---------------------------------------------------------------
    def batch_key(self, stage: int, start_level: int) -> tuple:
        """Requests whose current stages share a key run as one batch.  In
        the fine plan the rotate-and-add stages of every request at one level
        share a key whatever their step (``rotate_by_steps`` takes a step
        per entry); in the split plan the window-sum stage batches by level;
        other stages batch by (stage, start level)."""

        first = self._plan[stage][0]
        if self.stage_plan == "fine" and first >= 5:
            return ("rotate", start_level + self.depth)
        if self.stage_plan == "split" and first >= 5:
            return ("window", start_level + self.depth, stage)
        return (stage, start_level)

    def run_group(self, key: tuple, group: list["Request"]) -> None:
        if key[0] == "window":
            self.run_stage(key[2], group)
            return
        if key[0] != "rotate":
            self.run_stage(key[0], group)
            return
        eng = self.engine
        steps = [self.rotation_steps[r.stage - 5] for r in group]
        s = stack([r.registers["s"] for r in group])
        s = eng.add(s, eng.rotate_by_steps(s, steps), inplace=True)
        for r, v in zip(group, unstack(s)):
            r.registers["s"] = v

    # -- stages ----------------------------------------------------------------
    def run_stage(self, stage: int, group: list["Request"]) -> None:
        for step in self._plan[stage]:
            self._step(step, group)

    def _step(self, stage: int, group: list["Request"]) -> None:
        eng = self.engine
        regs = [r.registers for r in group]
        delta = eng.config.default_scale
        nreq = len(regs)
        if stage == 0:
            y = self._scaled_sum([[g["x"] for g in regs]], [1.0 / self.radius], delta)
            for g, v in zip(regs, unstack(y)):
                g["y1"] = v
                del g["x"]
        elif stage == 1:
            for g, v in zip(regs, self._hmult([g["y1"] for g in regs], None)):
                g["y2"] = v
        elif stage == 2:
            # y3 = y1 y2 and y4 = y2 y2 as one batched HMult of 2B entries
            lvl = regs[0]["y2"].level
            y1 = self._at([g["y1"] for g in regs], lvl)
            y2 = [g["y2"] for g in regs]
            out = self._hmult(y1 + y2, y2 + y2)
            for g, a, b in zip(regs, out[:nreq], out[nreq:]):
                g["y3"], g["y4"] = a, b
        elif stage == 3:
            # y5 = y1 y4, y6 = y3 y3 and y7 = y3 y4 as one HMult of 3B entries
            lvl = regs[0]["y4"].level
            y1 = self._at([g["y1"] for g in regs], lvl)
            y3 = [g["y3"] for g in regs]
            y4 = [g["y4"] for g in regs]
            out = self._hmult(y1 + y3 + y3, y4 + y3 + y4)
            for g, a, b, c in zip(regs, out[:nreq], out[nreq: 2 * nreq], out[2 * nreq:]):
                g["y5"], g["y6"], g["y7"] = a, b, c
        elif stage == 4:
            js = [j for j in range(1, 8) if self.power[j] != 0.0]
            if not js:
                raise ValueError("the polynomial has no non-constant term")
            acc = self._scaled_sum([[g[f"y{j}"] for g in regs] for j in js], [self.power[j] for j in js], delta)
            acc = eng.add_plaintext(acc, self._constant_coefficient(self.power[0], acc.level, delta),
                                    inplace=True)
            for g, v in zip(regs, unstack(acc)):
                g.clear()
                g["s"] = v
        else:
            step = self.rotation_steps[stage - 5]
            s = stack([g["s"] for g in regs])
            s = eng.add(s, eng.rotate_by_step(s, step), inplace=True)
            for g, v in zip(regs, unstack(s)):
                g["s"] = v

    def _constant_coefficient(self, value: float, level: int, scale: float) -> Plaintext:
        key = ("coefficient", float(value), level, float(scale))
        hit = self._plaintexts.get(key)
        if hit is None:
            eng = self.engine
            hit = eng.encode(np.full(eng.config.slots, value), level=level, scale=scale)
            self._plaintexts[key] = hit
        return hit

    def result(self, request: "Request") -> Ciphertext:
        return request.registers["s"]


# ---------------------------------------------------------------------------
# The loop
# ---------------------------------------------------------------------------
@dataclass
class Request:
    rid: int
    ciphertext: Ciphertext
    arrival: float
    start_level: int = field(init=False)
    stage: int = 0
    registers: dict = field(default_factory=dict)
    admitted: float | None = None
    admitted_iteration: int | None = None
    finished_iteration: int | None = None
    done_event: object = None
    latency: float | None = None
    output: Ciphertext | None = None

    def __post_init__(self) -> None:
        if self.ciphertext.component_count != 2 or self.ciphertext.batch_shape:
            raise ValueError("a request holds one unbatched two-component ciphertext")
        self.start_level = self.ciphertext.level


@dataclass
class ServingStats:
    completed: int
    wall_seconds: float
    iterations: int
    max_active: int
    max_batch_seen: int
    peak_committed_bytes: int
    budget_bytes: int
    latencies: list[float]

    @property
    def throughput(self) -> float:
        return self.completed / self.wall_seconds if self.wall_seconds > 0 else float("nan")

    def percentile(self, q: float) -> float:
        return float(np.percentile(np.asarray(self.latencies), q)) if self.latencies else float("nan")


class DeviceTimeline:
    """Stream events for completions, anchored to an event recorded at the
    start of a run."""

    def __init__(self, device: torch.device) -> None:
        self.device = device
        self._start: torch.cuda.Event | None = None

    def start(self) -> None:
        torch.cuda.synchronize(self.device)
        self._start = torch.cuda.Event(enable_timing=True)
        self._start.record()

    def mark(self) -> torch.cuda.Event:
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        return ev

    def wait(self, mark: torch.cuda.Event) -> None:
        mark.synchronize()

    def finish(self) -> None:
        torch.cuda.synchronize(self.device)

    def seconds(self, mark: torch.cuda.Event) -> float:
        return self._start.elapsed_time(mark) * 1e-3

    def budget(self, reserve: int) -> int:
        free, _ = torch.cuda.mem_get_info(self.device)
        cached = torch.cuda.memory_reserved(self.device) - torch.cuda.memory_allocated(self.device)
        return max(0, free + cached - reserve)


class ServingLoop:
    def __init__(self, program: PolynomialWindowSum, *, max_batch: int = 16, max_active: int = 64,
                 memory_budget_bytes: int | None = None, memory_reserve_bytes: int = 4 << 30,
                 max_inflight: int = 2, timeline=None) -> None:
        if max_batch < 1 or max_active < 1 or max_inflight < 1:
            raise ValueError("max_batch, max_active and max_inflight must be positive")
        self.program = program
        self.max_batch = max_batch
        self.max_active = max_active
        self.memory_budget_bytes = memory_budget_bytes
        self.memory_reserve_bytes = memory_reserve_bytes
        self.max_inflight = max_inflight
        self.timeline = timeline if timeline is not None else DeviceTimeline(program.engine.device)
        self.trace: list[tuple[int, int, int, tuple[int, ...]]] = []

    def run(self, requests: Sequence[Request], *, clock: Callable[[], float] | None = None,
            sleep: Callable[[float], None] | None = None) -> ServingStats:
        """Serve ``requests`` (arrival times in seconds after the start) to
        completion.  ``clock``/``sleep`` default to the wall clock; a test may
        pass a simulated clock.  ``self.trace`` records every batched stage
        call as (iteration, stage, start level, request ids)."""

        prog = self.program
        tl = self.timeline
        max_level = prog.max_start_level()
        for r in requests:
            if r.start_level > max_level:
                raise ValueError(f"request {r.rid} starts at level {r.start_level} > {max_level}")
        clock = clock or time.perf_counter
        sleep = sleep or time.sleep
        budget = (int(self.memory_budget_bytes) if self.memory_budget_bytes is not None
                  else tl.budget(self.memory_reserve_bytes))
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        for r in pending:
            if prog.footprint_bytes(r.start_level) > budget:
                raise MemoryError(f"request {r.rid} needs more HBM than the whole budget")
        active: list[Request] = []
        committed = 0
        inflight: deque = deque()
        self.trace = []
        tl.start()
        t0 = clock()
        iterations = max_active_seen = max_batch_seen = peak_committed = 0
        finished: list[Request] = []
        while pending or active:
            now = clock() - t0
            while (pending and pending[0].arrival <= now and len(active) < self.max_active
                   and committed + prog.footprint_bytes(pending[0].start_level) <= budget):
                r = pending.popleft()
                r.admitted, r.admitted_iteration = now, iterations
                r.stage, r.registers = 0, {"x": r.ciphertext}
                committed += prog.footprint_bytes(r.start_level)
                active.append(r)
            if not active:
                sleep(max(0.0, pending[0].arrival - now))
                continue
            max_active_seen = max(max_active_seen, len(active))
            peak_committed = max(peak_committed, committed)
            # batching groups: the program's own keys where it defines them
            # (its rotate-and-add stages batch across requests' stages),
            # otherwise one group per (stage, start level)
            key_of = getattr(prog, "batch_key", None)
            run_group = getattr(prog, "run_group", None)
            groups: dict[tuple, list[Request]] = {}
            for r in active:
                key = key_of(r.stage, r.start_level) if key_of else (r.stage, r.start_level)
                groups.setdefault(key, []).append(r)

            def order(key: tuple) -> tuple:
                return (-key[0] if isinstance(key[0], int) else -prog.stage_count, key[1])

            for key, members in sorted(groups.items(), key=lambda kv: order(kv[0])):
                for i in range(0, len(members), self.max_batch):
                    chunk = members[i: i + self.max_batch]
                    max_batch_seen = max(max_batch_seen, len(chunk))
                    self.trace.append((iterations, key[0], key[1], tuple(r.rid for r in chunk)))
                    if run_group is not None:
                        run_group(key, chunk)
                    else:
                        prog.run_stage(key[0], chunk)
            still: list[Request] = []
            for r in active:
                r.stage += 1
                if r.stage == prog.stage_count:
                    r.output = prog.result(r)
                    r.registers = {}
                    r.done_event = tl.mark()
                    r.finished_iteration = iterations
                    committed -= prog.footprint_bytes(r.start_level)
                    finished.append(r)
                else:
                    still.append(r)
            active = still
            iterations += 1
            inflight.append(tl.mark())
            if len(inflight) > self.max_inflight:
                tl.wait(inflight.popleft())
        tl.finish()
        wall_end = 0.0
        for r in finished:
            done = tl.seconds(r.done_event)
            r.latency = done - r.arrival
            wall_end = max(wall_end, done)
        first = min((r.arrival for r in finished), default=0.0)
        return ServingStats(completed=len(finished), wall_seconds=wall_end - first, iterations=iterations,
                            max_active=max_active_seen, max_batch_seen=max_batch_seen,
                            peak_committed_bytes=peak_committed, budget_bytes=budget,
                            latencies=[r.latency for r in finished])


def logistic_chebyshev(seed: int, *, degree: int = 7, radius: float = 4.0, jitter: float = 0.05) -> np.ndarray:
    """Seeded sigmoid-like coefficients: the degree-``degree`` Chebyshev
    interpolant (in y = x / radius) of 1 / (1 + e^-x) on [-radius, radius],
    each coefficient scaled by 1 + jitter * U[-1, 1]."""

    k = np.arange(degree + 1)
    nodes = np.cos(np.pi * (k + 0.5) / (degree + 1))
    c = np.polynomial.chebyshev.chebfit(nodes, 1.0 / (1.0 + np.exp(-radius * nodes)), degree)
    return c * (1.0 + jitter * np.random.default_rng(seed).uniform(-1.0, 1.0, degree + 1))


def poisson_arrivals(count: int, rate_per_s: float, rng: np.random.Generator) -> np.ndarray:
    """Arrival times (s) of a Poisson process: cumulative exponential gaps."""

    return np.cumsum(rng.exponential(1.0 / rate_per_s, size=count))
