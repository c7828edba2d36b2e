"""Rotation-step planning (fhelium/core/rotation.py:1-110) on the CPU.

The breadth-first composition is pinned three ways: walks found by hand from
the reference's rule (smallest signed step per residue, ascending edge
order, first walk to reach the target wins), walk length against an
exhaustive shortest-walk search over small inventories, and the sum of the
walk modulo S.  The power-of-two decompositions are pinned by their closed
forms (binary digits; non-adjacent form has no two adjacent non-zero digits
and the minimum weight).
"""

from __future__ import annotations

import itertools

import pytest

from fhelium_amd.rotation import (decompose_power_of_two_rotation, decompose_rotation_step,
                                  decompose_signed_power_of_two_rotation, power_of_two_key_steps)


def _shortest_len(target, steps, slots):
    """Exhaustive: smallest k such that some k-multiset of steps sums to target mod S."""
    target %= slots
    if target == 0:
        return 0
    for k in range(1, slots + 1):
        for combo in itertools.combinations_with_replacement(steps, k):
            if sum(combo) % slots == target:
                return k
    return None


@pytest.mark.parametrize("target, keys, want", [
    # S=16, keys {1, 2, -1}: edges -1, 1, 2; 0 reaches 15, 1, 2; then 1 + 2 = 3
    (3, {1: 0, 2: 0, -1: 0}, [1, 2]),
    # the smallest signed step represents a residue: -3 and 13 are both 13 mod 16
    (13, {-3: 0, 13: 0, 5: 0}, [-3]),
    # 0 reaches 13 (-3) and 5 first; 13 + 13 = 10 comes before 5 + 5 = 10
    (10, {-3: 0, 5: 0}, [-3, -3]),
    # direct key present: one hop
    (7, {7: 0, 1: 0}, [7]),
    # step 0 needs no hop
    (16, {1: 0}, []),
])
def test_bfs_walks_by_hand(target, keys, want):
    assert decompose_rotation_step(target, 16, keys) == want


def test_bfs_is_shortest_and_sums_to_target():
    slots = 16
    inventories = [(1,), (3,), (-1, 4), (2, 5), (1, 2, 4, 8), (-3, 6), (4, -4)]
    for inv in inventories:
        for target in range(slots):
            best = _shortest_len(target, inv, slots)
            if best is None:
                with pytest.raises(KeyError):
                    decompose_rotation_step(target, slots, inv)
                continue
            walk = decompose_rotation_step(target, slots, inv)
            assert len(walk) == best, (inv, target, walk)
            assert sum(walk) % slots == target % slots
            assert all(s in inv for s in walk)


def test_no_keys_or_unreachable():
    with pytest.raises(KeyError):
        decompose_rotation_step(3, 16, {})
    with pytest.raises(KeyError):
        decompose_rotation_step(3, 16, {2: 0, 4: 0})  # even steps never reach an odd target
    with pytest.raises(ValueError):
        decompose_rotation_step(3, 12, {1: 0})


def test_power_of_two_decompositions():
    slots = 64
    for r in range(-100, 100):
        bits = decompose_power_of_two_rotation(r, slots)
        assert sum(bits) == r % slots
        assert all(b & (b - 1) == 0 and b > 0 for b in bits)
        naf = decompose_signed_power_of_two_rotation(r, slots)
        assert sum(naf) % slots == r % slots
        mags = sorted(abs(s) for s in naf)
        # distinct powers of two, never adjacent (non-adjacent form)
        assert all(m & (m - 1) == 0 for m in mags)
        assert all(b > 2 * a for a, b in zip(mags, mags[1:])) or len(mags) < 2
    assert decompose_signed_power_of_two_rotation(7, 64) == [-1, 8]
    assert decompose_signed_power_of_two_rotation(-7, 64) == [1, -8]
    assert decompose_power_of_two_rotation(-1, 16) == [1, 2, 4, 8]


def test_power_of_two_inventory_reaches_everything_in_log_hops():
    slots = 256
    keys = power_of_two_key_steps(slots)
    assert 0 not in keys and all(-slots // 2 <= k < slots // 2 for k in keys)
    for r in range(slots):
        walk = decompose_rotation_step(r, slots, keys)
        assert sum(walk) % slots == r
        assert len(walk) <= 5  # ceil(log2(256) / 2) + 1 with signed digits
