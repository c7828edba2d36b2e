"""The serving loop on gfx950: each request's result is bit-identical when
served alone and when served in a mix of requests that arrive while others
are on the GPU (at two start levels, under two admission budgets), and each
decrypts to the binary64 polynomial-plus-window-sum within
window x 2e-5 (the reference's per-operation tolerance,
tests/test_ckks_operation_correctness.py:48-51, once per summed slot)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fhelium_amd.engine import CkksEngine
from fhelium_amd.serving import PolynomialWindowSum, Request, ServingLoop, logistic_chebyshev
from tests import helpers as H

pytestmark = pytest.mark.gpu

WINDOW = 64
TOL = WINDOW * 2e-5


@pytest.fixture(scope="module")
def setup():
    eng = CkksEngine(H.config(12), rng_seed=41)
    eng.generate_secret_key()
    eng.generate_relinearization_key()
    prog = PolynomialWindowSum(eng, logistic_chebyshev(7), radius=4.0, window=WINDOW)
    rng = np.random.default_rng(5)
    levels = [0, 0, 16, 0, 16, 3]
    xs = [rng.uniform(-4, 4, eng.config.slots) for _ in levels]
    cts = [eng.encrypt(x, level=lvl) for x, lvl in zip(xs, levels)]
    return eng, prog, xs, cts


def unit_clock():
    t = [0.0]

    def clock():
        now = t[0]
        t[0] += 1.0
        return now

    def sleep(dt):
        t[0] += dt

    return clock, sleep


def serve(prog, cts, arrivals, **kw):
    reqs = [Request(i, ct.clone(), a) for i, (ct, a) in enumerate(zip(cts, arrivals))]
    loop = ServingLoop(prog, **kw)
    clock, sleep = unit_clock()
    stats = loop.run(reqs, clock=clock, sleep=sleep)
    torch.cuda.synchronize()
    return reqs, stats, loop


@pytest.mark.parametrize("plan", PolynomialWindowSum.PLANS)
def test_same_alone_and_in_mix(setup, plan):
    eng, _, xs, cts = setup
    prog = PolynomialWindowSum(eng, logistic_chebyshev(7), radius=4.0, window=WINDOW, stage_plan=plan)
    alone = []
    for i in range(len(cts)):
        reqs, stats, _ = serve(prog, [cts[i]], [0.0])
        assert stats.completed == 1
        alone.append(reqs[0].output.data.clone())
        err = np.abs(eng.decrypt(reqs[0].output).real - prog.expected(xs[i])).max()
        assert err < TOL, (i, err)
    # arrivals spread so that later requests join while earlier ones are mid-program
    arrivals = [0.0, 0.0, 1.5, 3.0, 3.2, 6.5]
    reqs, stats, loop = serve(prog, cts, arrivals, max_batch=4)
    assert stats.completed == len(cts)
    assert any(len(ids) > 1 for _, _, _, ids in loop.trace)
    if plan == "fine":
        assert reqs[3].admitted_iteration < reqs[0].finished_iteration
    if plan != "whole":
        # some iteration ran requests at different stages: later arrivals
        # joined while earlier ones were part-way through the program
        stages_by_iteration = {}
        for it, stage, _, _ in loop.trace:
            stages_by_iteration.setdefault(it, set()).add(stage)
        assert any(len(v) > 1 for v in stages_by_iteration.values())
    for r in reqs:
        assert torch.equal(r.output.data, alone[r.rid]), f"request {r.rid} differs in the mix"
    # a budget of two level-0 footprints holds back admissions; same bits
    budget = 2 * prog.footprint_bytes(0)
    reqs2, stats2, _ = serve(prog, cts, arrivals, max_batch=4, memory_budget_bytes=budget)
    assert stats2.completed == len(cts)
    assert stats2.peak_committed_bytes <= budget and stats2.max_active <= stats.max_active
    if plan == "fine":
        assert stats2.max_active < stats.max_active
    for r in reqs2:
        assert torch.equal(r.output.data, alone[r.rid]), f"request {r.rid} differs under the budget"
    assert all(r.output.level == cts[r.rid].level + prog.depth for r in reqs2)


def test_wall_clock_run_reports_latency(setup):
    eng, prog, xs, cts = setup
    arrivals = np.cumsum(np.random.default_rng(9).exponential(0.002, size=len(cts)))
    reqs = [Request(i, ct.clone(), float(a)) for i, (ct, a) in enumerate(zip(cts, arrivals))]
    stats = ServingLoop(prog).run(reqs)
    assert stats.completed == len(cts)
    assert all(r.latency is not None and r.latency > 0 for r in reqs)
    assert stats.percentile(50) <= stats.percentile(99)
    assert stats.throughput > 0
