"""CPU check of the held-column bookkeeping of the LDS-resident kernels.

The FP64 class kernel (compact layout) and the FP32 stage of the mixed
pipeline keep the top column of each circle position in a lane group's
registers and move only bottom columns through LDS: half + 2 slots, the
bottom column at position P in global round g in slot (g + P) mod (half + 2).
A group that leaves position half - 1 stores its top into slot g - 2 (it
becomes bottom[half - 1]); one that arrives at position 1 loads the previous
bottom[0] from slot g - 1.  This model replays that bookkeeping (the same
integer updates as the kernels' round loop) while tracking which column each
slot and each group's registers hold, and checks every round against the
reference circle (the oracle's schedule, orientation included): each group
must hold its pair's top column, read its bottom column from the slot it
addresses, and the pairs of a round must be the reference round's pairs.
After the last round every column must be held exactly once (top in
registers or bottom in its slot), which is what the epilogues write out.
"""
import pytest

import oracle


def _replay(n, sweeps):
    even = n + (n & 1)
    half = even // 2
    ring = even - 1
    ns = half + 2
    rounds = even - 1 if n >= 2 else 0
    slots = [None] * ns
    # Top row staged through slots 0 .. half-1 into registers, then the
    # bottom row (position P holds column even-1-P) in slots 0 .. half-1.
    for sl in range(half):
        slots[sl] = sl
    reg = [slots[k] if k < n else None for k in range(half)]
    for sl in range(half):
        slots[sl] = even - 1 - sl
    pos = list(range(half))
    ct = list(range(half))
    cb = [even - 1 - k for k in range(half)]
    fresh = [0] * half
    z, gs = even - 1, 0
    schedule = oracle.schedule(n)
    for _ in range(sweeps):
        for r in range(rounds):
            z1 = ring if z - 1 < 1 else z - 1
            z2 = z - 2 + ring if z - 2 < 1 else z - 2
            ls = ns - 1 if gs == 0 else gs - 1
            ws = gs + ns - 2 if gs < 2 else gs - 2
            got = set()
            for k in range(half):
                p, q = ct[k], cb[k]
                if fresh[k] and p < n:
                    reg[k] = slots[ls]
                bsl = 0 if half == 1 else (gs + pos[k] - ns if gs + pos[k] >= ns else gs + pos[k])
                if p < n:
                    assert reg[k] == p, (n, r, k, "top register")
                if q < n:
                    assert slots[bsl] == q, (n, r, k, "bottom slot")
                if p < n and q < n:
                    got.add((p, q))
                wrap = k > 0 and pos[k] == half - 1
                if wrap and p < n:
                    assert ws not in [(gs + P) % ns for P in range(half)] or half == 1
                    slots[ws] = reg[k]
                q1 = ring if q - 1 < 1 else q - 1
                q2 = q - 2 + ring if q - 2 < 1 else q - 2
                cb[k] = q1 if k == 0 else (z2 if wrap else q2)
                ct[k] = z if wrap else p
                pos[k] = 1 if wrap else (pos[k] + 1 if k > 0 else 0)
                fresh[k] = int(wrap)
            z = z1
            gs = 0 if gs + 1 == ns else gs + 1
            assert got == set(schedule[r]), (n, r, "pairs of the round")
    # Epilogue: pending top loads, then every column held exactly once.
    ls = ns - 1 if gs == 0 else gs - 1
    held = []
    for k in range(half):
        if fresh[k] and ct[k] < n:
            reg[k] = slots[ls]
        bsl = 0 if half == 1 else (gs + pos[k]) % ns
        if ct[k] < n:
            assert reg[k] == ct[k]
            held.append(ct[k])
        if cb[k] < n:
            assert slots[bsl] == cb[k]
            held.append(cb[k])
    assert sorted(held) == list(range(n)), (n, "epilogue columns")


@pytest.mark.parametrize("n", list(range(1, 91)) + [127, 128, 129, 255, 256])
def test_held_column_bookkeeping_follows_reference_circle(n):
    _replay(n, sweeps=3)
