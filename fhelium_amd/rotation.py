"""Rotation-step planning: how a slot rotation is realised from installed keys.

Counterpart of fhelium/core/rotation.py:1-110.  A rotation by r slots is a
walk in the cyclic group Z/S (S = N/2 slots); every installed rotation key is
one edge length.  When no key for r itself is installed and the engine may
not generate one, the rotation is composed from installed keys along a
shortest walk.  The walk the reference picks is the one every noise-exact
comparison depends on (two different compositions of the same total step
decrypt to the same slots but leave different residues), so the choice rule
is kept exactly:

* edges are the distinct non-zero residues of the installed steps; the step
  recorded for a residue is the smallest installed signed step with that
  residue (``decompose_rotation_step``, rotation.py:71-77);
* the search is breadth first from 0, expanding walk ends in the order they
  were reached and, from each, the edges in ascending order of their signed
  step (rotation.py:79-88); the first walk that reaches r wins.
"""

from __future__ import annotations

from collections.abc import Iterable, Mapping


def _check_slots(num_slots: int) -> None:
    if num_slots <= 0 or num_slots & (num_slots - 1):
        raise ValueError("num_slots must be a positive power of two")


def decompose_power_of_two_rotation(rotation_step: int, num_slots: int) -> list[int]:
    """The set bits of r mod S as ascending positive powers of two
    (rotation.py:9-22)."""

    _check_slots(num_slots)
    r = int(rotation_step) % num_slots
    out = []
    bit = 1
    while bit < num_slots:
        if r & bit:
            out.append(bit)
        bit <<= 1
    return out


def decompose_signed_power_of_two_rotation(rotation_step: int, num_slots: int) -> list[int]:
    """Non-adjacent form of the canonical step in [-S/2, S/2): signed powers
    of two, least significant first, each reported as a canonical step
    (rotation.py:25-54).  The walk has the minimum number of non-zero signed
    binary digits."""

    _check_slots(num_slots)
    half = num_slots // 2

    def canon(v: int) -> int:
        return (v + half) % num_slots - half

    value = canon(int(rotation_step))
    weight = 1
    out: list[int] = []
    while value != 0:
        if value % 2:
            # digit +1 when value = 1 (mod 4), -1 when value = 3 (mod 4)
            digit = 1 if value % 4 == 1 else -1
            out.append(canon(digit * weight))
            value -= digit
        value //= 2
        weight *= 2
    return out


def _edges(installed: Iterable[int], num_slots: int) -> list[tuple[int, int]]:
    """(residue, signed step) per distinct non-zero residue, the smallest
    signed step per residue, in ascending signed-step order."""

    chosen: dict[int, int] = {}
    for step in sorted(int(s) for s in installed):
        residue = step % num_slots
        if residue and residue not in chosen:
            chosen[residue] = step
    return sorted(((r, s) for r, s in chosen.items()), key=lambda rs: rs[1])


def decompose_rotation_step(rotation_step: int, num_slots: int,
                            rotation_keys: Mapping[int, object] | Iterable[int]) -> list[int]:
    """Shortest sequence of installed steps whose sum is r modulo S, chosen
    by the reference's breadth-first rule (rotation.py:57-110).  Raises
    KeyError when the installed steps cannot reach r."""

    _check_slots(num_slots)
    target = int(rotation_step) % num_slots
    if target == 0:
        return []
    edges = _edges(rotation_keys, num_slots)
    if not edges:
        raise KeyError(f"No installed rotation-key path can realize step {rotation_step}")
    # parent[v] = (previous walk end, step taken); walk ends are expanded in
    # the order they were first reached
    parent: dict[int, tuple[int, int] | None] = {0: None}
    frontier = [0]
    head = 0
    while head < len(frontier) and target not in parent:
        here = frontier[head]
        head += 1
        for residue, step in edges:
            nxt = (here + residue) % num_slots
            if nxt not in parent:
                parent[nxt] = (here, step)
                frontier.append(nxt)
    if target not in parent:
        raise KeyError(f"No installed rotation-key path can realize step {rotation_step}; "
                       f"installed steps are {tuple(s for _, s in edges)}")
    walk: list[int] = []
    node = target
    while node != 0:
        prev, step = parent[node]
        walk.append(step)
        node = prev
    walk.reverse()
    return walk


def power_of_two_key_steps(num_slots: int, *, signed: bool = True) -> list[int]:
    """Canonical steps of a power-of-two key inventory: +-2^k (or +2^k) for
    2^k < S/2, plus -S/2.  Any step is then at most log2(S) hops."""

    _check_slots(num_slots)
    half = num_slots // 2
    steps = set()
    k = 1
    while k < num_slots:
        for v in ((k, -k) if signed else (k,)):
            steps.add((v + half) % num_slots - half)
        k <<= 1
    steps.discard(0)
    return sorted(steps)
