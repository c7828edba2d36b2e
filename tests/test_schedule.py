"""Issue order of fhelium_amd.schedule (host logic, no GPU)."""

from __future__ import annotations

import pytest

from fhelium_amd.schedule import issue_order


@pytest.mark.parametrize("n", [0, 1, 2, 5, 33])
@pytest.mark.parametrize("streams", [1, 2, 3, 4])
def test_every_task_once(n, streams):
    plan = issue_order(n, streams)
    assert sorted(i for i, _ in plan) == list(range(n))
    assert all(0 <= s < streams for _, s in plan)


def test_two_streams_walk_from_both_ends():
    plan = issue_order(7, 2)
    assert plan == [(0, 0), (6, 1), (1, 0), (5, 1), (2, 0), (4, 1), (3, 0)]
    # each stream runs its tasks in its own direction
    assert [i for i, s in plan if s == 0] == [0, 1, 2, 3]
    assert [i for i, s in plan if s == 1] == [6, 5, 4]


def test_one_stream_keeps_order_and_bad_arguments_raise():
    assert issue_order(4, 1) == [(0, 0), (1, 0), (2, 0), (3, 0)]
    with pytest.raises(ValueError):
        issue_order(3, 0)
