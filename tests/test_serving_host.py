"""Host logic of the serving loop on the CPU: admission by memory and slot
limits, continuous batching of requests that arrive mid-flight, stage order,
and the Chebyshev-to-power conversion the polynomial program relies on."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest
import torch

from fhelium_amd.core import Ciphertext
from fhelium_amd.serving import Request, ServingLoop, chebyshev_to_power, poisson_arrivals


class FakeProgram:
    stage_count = 4
    depth = 1

    def __init__(self):
        self.engine = SimpleNamespace(device=torch.device("cpu"))
        self.calls = []

    def max_start_level(self):
        return 10

    def footprint_bytes(self, level):
        return 100 - level

    def run_stage(self, stage, group):
        self.calls.append((stage, tuple(r.rid for r in group)))
        for r in group:
            r.registers.setdefault("history", []).append(stage)

    def result(self, r):
        return r.registers["history"]


class FakeTimeline:
    def __init__(self):
        self.t = 0

    def start(self):
        self.t = 0

    def mark(self):
        self.t += 1
        return self.t

    def wait(self, mark):
        pass

    def finish(self):
        pass

    def seconds(self, mark):
        return float(mark)

    def budget(self, reserve):
        return 10**9


def request(rid, arrival, level=0):
    return Request(rid, Ciphertext(torch.zeros(2, 3, 4, dtype=torch.int64), level, 1.0), arrival)


def unit_clock():
    t = [0.0]

    def clock():
        now = t[0]
        t[0] += 1.0
        return now

    def sleep(dt):
        t[0] += dt

    return clock, sleep


def test_late_request_joins_in_flight_and_batches_by_stage_and_level():
    prog = FakeProgram()
    loop = ServingLoop(prog, max_batch=2, timeline=FakeTimeline())
    reqs = [request(0, 0.0), request(1, 0.0), request(2, 0.0, level=3), request(3, 0.0), request(4, 2.5)]
    clock, sleep = unit_clock()
    stats = loop.run(reqs, clock=clock, sleep=sleep)
    assert stats.completed == 5
    for r in reqs:
        assert r.output == [0, 1, 2, 3]
    late = reqs[4]
    assert late.admitted_iteration > 0 and late.admitted_iteration < reqs[0].finished_iteration
    # the late request's stages run in iterations shared with earlier requests
    shared = {it for it, _, _, ids in loop.trace if 4 in ids} & {it for it, _, _, ids in loop.trace if 0 in ids}
    assert shared
    # every batched call is one (stage, level) group of at most max_batch entries
    by_id = {r.rid: r for r in reqs}
    for it, stage, level, ids in loop.trace:
        assert 1 <= len(ids) <= 2
        assert all(by_id[i].start_level == level for i in ids)
    assert stats.max_batch_seen == 2


def test_admission_respects_memory_budget_and_fifo():
    prog = FakeProgram()
    loop = ServingLoop(prog, max_batch=8, memory_budget_bytes=200, timeline=FakeTimeline())
    reqs = [request(i, 0.0) for i in range(5)]
    clock, sleep = unit_clock()
    stats = loop.run(reqs, clock=clock, sleep=sleep)
    assert stats.completed == 5 and stats.max_active == 2 and stats.peak_committed_bytes == 200
    order = sorted(reqs, key=lambda r: (r.admitted_iteration, r.rid))
    assert [r.rid for r in order] == [0, 1, 2, 3, 4]
    assert all(r.output == [0, 1, 2, 3] for r in reqs)


def test_slot_limit_and_idle_wait():
    prog = FakeProgram()
    loop = ServingLoop(prog, max_active=1, timeline=FakeTimeline())
    reqs = [request(0, 5.0), request(1, 5.0)]
    clock, sleep = unit_clock()
    stats = loop.run(reqs, clock=clock, sleep=sleep)
    assert stats.completed == 2 and stats.max_active == 1
    assert reqs[1].admitted_iteration >= reqs[0].finished_iteration + 1


def test_rejections():
    prog = FakeProgram()
    with pytest.raises(ValueError):
        ServingLoop(prog, timeline=FakeTimeline()).run([request(0, 0.0, level=11)])
    with pytest.raises(MemoryError):
        ServingLoop(prog, memory_budget_bytes=50, timeline=FakeTimeline()).run([request(0, 0.0)])
    with pytest.raises(ValueError):
        Request(0, Ciphertext(torch.zeros(2, 2, 3, 4, dtype=torch.int64), 0, 1.0), 0.0)


def test_chebyshev_to_power():
    # T7(y) = 64 y^7 - 112 y^5 + 56 y^3 - 7 y
    np.testing.assert_array_equal(chebyshev_to_power([0] * 7 + [1]), [0, -7, 0, 56, 0, -112, 0, 64])
    c = np.random.default_rng(1).uniform(-1, 1, 8)
    y = np.linspace(-1, 1, 101)
    np.testing.assert_allclose(np.polyval(chebyshev_to_power(c)[::-1], y),
                               np.polynomial.chebyshev.chebval(y, c), atol=1e-13)


def test_poisson_arrivals():
    t = poisson_arrivals(20000, 200.0, np.random.default_rng(2))
    assert np.all(np.diff(t) > 0)
    assert abs(t[-1] / 20000 - 1 / 200.0) < 0.05 / 200.0


class KeyedProgram(FakeProgram):
    """Stages 2 and 3 share a batch key (as the window sum's rotate-and-add
    stages do), so requests at different stages batch together there."""

    def batch_key(self, stage, start_level):
        return ("late", start_level) if stage >= 2 else (stage, start_level)

    def run_group(self, key, group):
        self.calls.append((key, tuple((r.rid, r.stage) for r in group)))
        for r in group:
            r.registers.setdefault("history", []).append(r.stage)


def test_program_batch_keys_merge_stages():
    prog = KeyedProgram()
    loop = ServingLoop(prog, max_batch=8, timeline=FakeTimeline())
    reqs = [request(0, 0.0), request(1, 1.5)]
    clock, sleep = unit_clock()
    stats = loop.run(reqs, clock=clock, sleep=sleep)
    assert stats.completed == 2 and all(r.output == [0, 1, 2, 3] for r in reqs)
    mixed = [members for key, members in prog.calls if key[0] == "late" and len({s for _, s in members}) > 1]
    assert mixed, prog.calls
