"""Hand-worked pins of the science-lambda block fix-up (AtmosphericTurbulence.c:3627-4216).

The reference compares the peaks of neighbouring (Nout/20)^2 blocks in twelve
unrolled loops (AT.c:3667-3903 for float images, 3937-4173 for double).  Read
by hand, those loops say:
  * every pair of distinct blocks whose offset has Chebyshev length 1 or 2 is
    compared once (8 near and 16 far neighbours of an interior block; the
    loop ranges xs - xm, ys - ym cover exactly the pairs inside the grid);
  * when the two peaks differ by more than pi, the block with the larger peak
    gets -c and the other +c, c = 0.2 for length 1 (AT.c:3679) and
    pcoeff2 = 0.15 for length 2 (AT.c:1596, 3761);
  * plim = max(max |count| - 0.001, 0.5) (AT.c:3905-3912); a block whose count
    exceeds plim moves by +2 pi, below -plim by -2 pi (AT.c:3916-3930);
  * a round that moves no block with 1 < ii2 < xs - 2 and 1 < jj2 < ys - 2
    adds one to chcnt0cnt, any other round resets it; the loop runs while
    plim > 0.51, fewer than 1000 rounds and chcnt0cnt < 5 (AT.c:3627, 4209-4213).
The tests below state those facts in forms of their own (a brute-force pair
enumeration by distance, explicit count maps, oscillating configurations whose
round count follows from the interior rule) so that a wrong entry in the
oracle's relation table, weight, range or termination rule fails one of them.
"""
import itertools

import numpy as np
import pytest


def _pairs_by_distance(ys, xs):
    """All unordered pairs of blocks at Chebyshev distance 1 or 2 with their weights."""
    out = []
    cells = [(j, i) for j in range(ys) for i in range(xs)]
    for a, b in itertools.combinations(cells, 2):
        d = max(abs(a[0] - b[0]), abs(a[1] - b[1]))
        if d == 1:
            out.append((a, b, 0.2))
        elif d == 2:
            out.append((a, b, 0.15))
    return out


def _brute_counts(peaks):
    ys, xs = peaks.shape
    ch = np.zeros_like(peaks)
    for a, b, c in _pairs_by_distance(ys, xs):
        pa, pb = peaks[a], peaks[b]
        if abs(pa - pb) > np.pi:
            hi, lo = (a, b) if pa > pb else (b, a)
            ch[hi] -= c
            ch[lo] += c
    return ch


def test_every_offset_weight_and_grid_edge(oracle):
    # one pair A (+2.5), B (-2.5) on a zero background, at every position of a
    # 7 x 7 grid where both fit: only A-B differ by more than pi
    n = 7
    offsets = [(dx, dy) for dx in range(-3, 4) for dy in range(-3, 4) if (dx, dy) != (0, 0)]
    checked = 0
    for dx, dy in offsets:
        dist = max(abs(dx), abs(dy))
        c = {1: 0.2, 2: 0.15}.get(dist, 0.0)
        for jy in range(n):
            for ix in range(n):
                ib, jb = ix + dx, jy + dy
                if not (0 <= ib < n and 0 <= jb < n):
                    continue
                peaks = np.zeros((n, n))
                peaks[jy, ix] = 2.5
                peaks[jb, ib] = -2.5
                out, ch, it = oracle.block_fix(peaks)
                want = np.zeros((n, n))
                want[jy, ix] = -c
                want[jb, ib] = c
                assert np.array_equal(ch, want), (dx, dy, ix, jy)
                # |count| <= 0.2 < plim = 0.5: nothing moves, one round (plim <= 0.51)
                assert it == 1 and np.array_equal(out, peaks)
                checked += 1
    assert checked == sum((n - abs(dx)) * (n - abs(dy)) for dx, dy in offsets)


@pytest.mark.parametrize("shape", [(3, 3), (5, 4), (4, 7), (8, 8), (9, 6)])
def test_first_round_counts_match_pair_enumeration(oracle, shape):
    rng = np.random.default_rng(sum(shape))
    for _ in range(20):
        peaks = rng.choice([-6.0, -3.3, -1.7, 0.0, 1.7, 3.3, 6.0], size=shape) + rng.uniform(-0.1, 0.1, shape)
        _, ch, _ = oracle.block_fix(peaks)
        assert np.allclose(ch, _brute_counts(peaks), atol=1e-12, rtol=0)


def test_plim_picks_the_largest_count_near_weights_beat_far(oracle):
    # A (-2.5) has three +2.5 near neighbours: count +0.6; B (-2.5) has three
    # +2.5 neighbours at distance 2: count +0.45; the groups are 3 apart.
    # plim = 0.599: A moves by +2 pi after round one, B does not.  With the
    # weights swapped B would move instead.
    n = 9
    peaks = np.zeros((n, n))
    A, B = (2, 2), (6, 6)
    peaks[A] = peaks[B] = -2.5
    near = [(2, 3), (3, 2), (3, 3)]
    far = [(6, 8), (8, 6), (8, 8)]
    for p in near + far:
        peaks[p] = 2.5
    _, ch, _ = oracle.block_fix(peaks)
    want = np.zeros((n, n))
    want[A], want[B] = 0.6, 0.45
    for p in near:
        want[p] = -0.2
    for p in far:
        want[p] = -0.15
    assert np.allclose(ch, want, atol=1e-12)
    one, _, it = oracle.block_fix(peaks, max_rounds=1)
    assert it == 1
    moved = np.argwhere(one != peaks)
    assert moved.tolist() == [list(A)] and np.isclose(one[A], -2.5 + 2 * np.pi)


@pytest.mark.parametrize("count,moves", [(0.45, False), (0.5, False), (0.55, True), (0.6, True)])
def test_plim_floor_half(oracle, count, moves):
    # corner block (0, 0) of a 6 x 6 grid below some +2.5 neighbours:
    # 0.45 = 3 far, 0.5 = 1 near + 2 far, 0.55 = 2 near + 1 far, 0.6 = 3 near.
    # plim = max(count - 0.001, 0.5): counts up to 0.5 leave it in place and
    # end the loop after one round (plim <= 0.51)
    sets = {0.45: [(0, 2), (2, 0), (2, 2)], 0.5: [(1, 1), (0, 2), (2, 0)],
            0.55: [(0, 1), (1, 0), (2, 2)], 0.6: [(0, 1), (1, 0), (1, 1)]}
    peaks = np.zeros((6, 6))
    peaks[0, 0] = -2.5
    for p in sets[count]:
        peaks[p] = 2.5
    _, ch, it = oracle.block_fix(peaks)
    assert np.isclose(ch[0, 0], count, atol=1e-12)
    one, _, _ = oracle.block_fix(peaks, max_rounds=1)
    assert (one[0, 0] != peaks[0, 0]) == moves
    if not moves:
        assert it == 1


def _oscillator(n, pos):
    # block at pos (-2.5) below three +2.5 near neighbours, zeros elsewhere:
    # round 1 moves it to 3.78 (count +0.6), which then exceeds every zero
    # neighbour by more than pi, so round 2 moves it back; nothing else moves
    peaks = np.zeros((n, n))
    jy, ix = pos
    peaks[jy, ix] = -2.5
    sx = 1 if ix + 1 < n else -1
    sy = 1 if jy + 1 < n else -1
    for p in [(jy, ix + sx), (jy + sy, ix), (jy + sy, ix + sx)]:
        peaks[p] = 2.5
    return peaks


@pytest.mark.parametrize("pos,rounds", [
    ((4, 1), 5), ((4, 2), 1000), ((4, 6), 1000), ((4, 7), 5),   # x: interior is 2 <= ii2 <= 6
    ((1, 4), 5), ((2, 4), 1000), ((6, 4), 1000), ((7, 4), 5),   # y: interior is 2 <= jj2 <= 6
    ((0, 0), 5), ((8, 8), 5), ((4, 4), 1000),
])
def test_interior_rule_and_five_quiet_rounds(oracle, pos, rounds):
    # an oscillating block outside the interior moves every round without
    # counting, so the loop stops after five rounds; inside, it counts every
    # round and the loop runs to the 1000-round cap
    out, ch, it = oracle.block_fix(_oscillator(9, pos))
    assert it == rounds
    assert np.isclose(ch[pos], 0.6, atol=1e-12)


def test_isolated_interior_block_count_is_four(oracle):
    # an interior block 2 pi off a uniform background: 8 x 0.2 + 16 x 0.15 = 4
    # against it; each neighbour gets the opposite single weight
    n = 9
    peaks = np.full((n, n), 2.9)
    peaks[4, 4] = 2.9 + 0.3 - 2 * np.pi
    out, ch, it = oracle.block_fix(peaks)
    assert np.isclose(ch[4, 4], 4.0, atol=1e-12)
    d = np.maximum(np.abs(np.arange(n)[:, None] - 4), np.abs(np.arange(n)[None, :] - 4))
    want = np.where(d == 1, -0.2, np.where(d == 2, -0.15, 0.0))
    want[4, 4] = 4.0
    assert np.allclose(ch, want, atol=1e-12)
    assert it == 2 and np.isclose(out[4, 4], 3.2)


def _reference_order(peaks, b):
    """The +-c steps that reach block b in the reference's order: relation by
    relation (AT.c:3667-3903), and within one relation by the loop's (ii2, jj2)
    order of the pair's base block."""
    ys, xs = peaks.shape
    jb, ib = b
    rels = [  # (block 1, block 2) offsets from the base as written in the loops, weight
        ((0, 0), (1, 0), 0.2), ((0, 0), (0, 1), 0.2), ((0, 0), (1, 1), 0.2), ((0, 1), (1, 0), 0.2),
        ((0, 0), (2, 0), 0.15), ((0, 2), (0, 0), 0.15), ((1, 2), (0, 0), 0.15), ((0, 2), (1, 0), 0.15),
        ((2, 2), (0, 0), 0.15), ((0, 2), (2, 0), 0.15), ((0, 1), (2, 0), 0.15), ((2, 1), (0, 0), 0.15)]
    steps = []
    for d1, d2, c in rels:
        touches = []
        for role, (dx, dy) in ((1, d1), (2, d2)):
            bx, by = ib - dx, jb - dy           # base (ii2, jj2) of the pair
            ox, oy = (d2 if role == 1 else d1)  # the other block of the pair
            px, py = bx + ox, by + oy
            if 0 <= bx and 0 <= by and 0 <= px < xs and 0 <= py < ys and bx + max(d1[0], d2[0]) < xs \
                    and by + max(d1[1], d2[1]) < ys:
                touches.append(((bx, by), role, peaks[py, px]))
        for _, role, pv_other in sorted(touches, key=lambda t: t[0]):
            pv_self = peaks[jb, ib]
            pv1, pv2 = (pv_self, pv_other) if role == 1 else (pv_other, pv_self)
            if pv2 > pv1 + np.pi:
                steps.append(c if role == 1 else -c)
            if pv1 > pv2 + np.pi:
                steps.append(-c if role == 1 else c)
    return steps


def test_float_counts_follow_the_reference_order(oracle):
    # WFprecision = 0: every += rounds to float, so the order of the steps
    # shows in the last bits (e.g. -0.2 +0.2 -0.2 x4 +0.15 +0.15 -0.15 +0.15
    # +0.15 -0.15 gives -0.50000006, not -0.5)
    seq = [-0.2, 0.2, -0.2, -0.2, -0.2, -0.2, 0.15, 0.15, -0.15, 0.15, 0.15, -0.15]
    acc = np.float32(0)
    for v in seq:
        acc = np.float32(np.float64(acc) + v)
    assert acc == np.float32(-0.50000006)
    rng = np.random.default_rng(7)
    differs = 0
    for _ in range(30):
        peaks = rng.choice([-6.5, -3.3, 0.0, 3.3, 6.5], size=(7, 8)).astype(np.float32).astype(np.float64)
        _, ch32, _ = oracle.block_fix(peaks, prec_f32=True)
        _, ch64, _ = oracle.block_fix(peaks)
        for jb in range(7):
            for ib in range(8):
                acc = np.float32(0)
                for v in _reference_order(peaks, (jb, ib)):
                    acc = np.float32(np.float64(acc) + v)
                assert ch32[jb, ib] == acc, (jb, ib)
                differs += ch32[jb, ib] != ch64[jb, ib]
    assert differs > 0  # the float mode is visibly a different accumulation


def test_plim_margin_moves_ties_not_the_next_count(oracle):
    # counts 0.6, 0.6 and 0.55 on separate corners of a 9 x 9 grid: plim =
    # 0.6 - 0.001, so both 0.6 blocks move after round one and the 0.55 block
    # (one step of 0.05 below) does not
    n = 9
    peaks = np.zeros((n, n))
    groups = {(0, 0): [(0, 1), (1, 0), (1, 1)], (8, 8): [(8, 7), (7, 8), (7, 7)],
              (0, 8): [(0, 7), (1, 8), (2, 6)]}
    for blk, nb in groups.items():
        peaks[blk] = -2.5
        for p in nb:
            peaks[p] = 2.5
    _, ch, _ = oracle.block_fix(peaks)
    assert np.allclose([ch[0, 0], ch[8, 8], ch[0, 8]], [0.6, 0.6, 0.55], atol=1e-12)
    one, _, _ = oracle.block_fix(peaks, max_rounds=1)
    moved = sorted(map(tuple, np.argwhere(one != peaks).tolist()))
    assert moved == [(0, 0), (8, 8)]
