"""Independent batched calls issued on several streams.

A step of the batching tier is a list of independent tasks (for instance
one batched HMult and rotation per level group).  The tasks differ in size
by more than an order of magnitude: a level-0 group fills the GPU while a
level-32 group's launches are short and latency-bound.  ``run_on_streams``
issues them on side streams forked from the current stream and joins them
back, so that small tasks run beside large ones.  With two streams one
stream walks the tasks from the first (largest, when the caller orders them
by size) and the other from the last, issued alternately; with more
streams the tasks go round robin.  Every task must only read its own
inputs: tasks on different streams run concurrently.
"""

from __future__ import annotations

from typing import Any, Callable, Sequence

import torch

_STREAMS: dict[int, list[torch.cuda.Stream]] = {}


def issue_order(n: int, streams: int) -> list[tuple[int, int]]:
    """(task index, stream index) pairs in issue order."""

    if n < 0 or streams < 1:
        raise ValueError("need n >= 0 tasks and at least one stream")
    if streams == 1:
        return [(i, 0) for i in range(n)]
    if streams == 2:
        plan, lo, hi, turn = [], 0, n - 1, 0
        while lo <= hi:
            if turn == 0:
                plan.append((lo, 0))
                lo += 1
            else:
                plan.append((hi, 1))
                hi -= 1
            turn ^= 1
        return plan
    return [(i, i % streams) for i in range(n)]


def side_streams(count: int, device: torch.device | None = None) -> list[torch.cuda.Stream]:
    """``count`` default-priority side streams of ``device``, created once."""

    dev = torch.cuda.current_device() if device is None else torch.device(device).index
    pool = _STREAMS.setdefault(dev, [])
    while len(pool) < count:
        pool.append(torch.cuda.Stream(device=dev))
    return pool[:count]


def run_on_streams(tasks: Sequence[Callable[[], Any]], streams: int = 2) -> list[Any]:
    """Run every task (a callable issuing GPU work on the current stream)
    with its own stream current, ``streams`` streams in all, and return the
    results in task order.  The side streams wait for the caller's stream
    before the first task and the caller's stream waits for all of them
    after the last, so the call orders like one launch on the caller's
    stream.  With ``streams == 1`` the tasks run in order on the caller's
    stream."""

    if streams <= 1:
        return [task() for task in tasks]
    main = torch.cuda.current_stream()
    side = side_streams(streams, main.device)
    fork = torch.cuda.Event()
    fork.record(main)
    for st in side:
        st.wait_event(fork)
    out: list[Any] = [None] * len(tasks)
    try:
        for i, si in issue_order(len(tasks), streams):
            with torch.cuda.stream(side[si]):
                out[i] = tasks[i]()
    finally:
        # joined even when a task raises, so later work on the caller's
        # stream stays ordered after everything already issued
        for st in side:
            join = torch.cuda.Event()
            join.record(st)
            main.wait_event(join)
    return out
