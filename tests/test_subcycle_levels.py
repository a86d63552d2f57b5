"""The subcycle kernel's stage areas (pcom_amd/csrc/trop.hip, ring_level).

k_trop_tiles runs each stage of a substep only on the region points a later
stage of the round reads, ordered so that whole wavefronts skip the rest.
These CPU checks pin that bookkeeping independently of the GPU parity tests:

- the thread order is a permutation of the region sorted by level, and the
  points of level <= L are the (TW + L) x (TH + L) rectangle the kernel
  assumes;
- propagating "holds this substep's value" through the stages, with the
  reads each stage makes (mod_int.f90:184-273 as the kernel evaluates it),
  leaves every owned point's ch and velocity valid after every substep and
  round, for the kernel's own stage levels (read from trop.hip) and halo of
  2S points;
- a halo one point narrower, or any stage area one level smaller, loses a
  value an owned point needs (so the check can see a missing point).
"""
import os
import re

import pytest

SRC = os.path.join(os.path.dirname(__file__), "..", "pcom_amd", "csrc", "trop.hip")


def kernel_stage_levels():
    m = re.search(r"kStageLevel\[5\]\s*=\s*\{([^}]*)\}", open(SRC).read())
    return [int(v) for v in m.group(1).split(",")]


def ring_level(d, t):
    if d < 0:
        return -2 * d - 1
    if d > t - 1:
        return 2 * (d - t + 1)
    return 0


def ring_before(m, bx, by, h, tw, th):
    if m < 0:
        return 0
    lo = h - (m + 1) // 2
    w, ht = tw + m, th + m
    rows = min(max(by - lo, 0), ht)
    cols = min(max(bx - lo, 0), w) if lo <= by < lo + ht else 0
    return rows * w + cols


def order(tw, th, s):
    """The kernel's persistent-mode thread -> region point map."""
    h = 2 * s
    rw, rh = tw + 2 * h, th + 2 * h
    out = [None] * (rw * rh)
    for t in range(rw * rh):
        rx, ry = t % rw, t // rw
        lv = max(ring_level(rx - h, tw), ring_level(ry - h, th))
        pos = ((tw + lv - 1) * (th + lv - 1) if lv > 0 else 0) + ring_before(lv, rx, ry, h, tw, th) \
            - ring_before(lv - 1, rx, ry, h, tw, th)
        assert out[pos] is None
        out[pos] = (rx, ry, lv)
    return out


@pytest.mark.parametrize("s", [1, 2, 3, 4])
def test_thread_order_is_a_level_sorted_permutation(s):
    for tw in range(1, 30):
        for th in range(1, 30):
            if (tw + 4 * s) * (th + 4 * s) > 1024:
                continue
            o = order(tw, th, s)
            levels = [p[2] for p in o]
            assert None not in o and levels == sorted(levels)
            for lv in range(4 * s + 1):
                assert sum(1 for x in levels if x <= lv) == (tw + lv) * (th + lv)


def run_rounds(tw, th, s, levels, halo, rounds=2):
    """Validity of each array through `rounds` rounds of s substeps; raises
    AssertionError when an owned point's ch or velocity is not this
    substep's value."""
    rw, rh = tw + 2 * halo, th + 2 * halo
    pts = [(x, y) for y in range(rh) for x in range(rw)]
    lvl = {(x, y): max(ring_level(x - halo, tw), ring_level(y - halo, th)) for x, y in pts}
    owned = {(x, y) for x, y in pts if halo <= x < halo + tw and halo <= y < halo + th}
    chp = dict.fromkeys(pts, True)
    uv = dict.fromkeys(pts, True)

    def v(arr, p, dx, dy):  # a read outside the region is a clamped (garbage) read
        q = (p[0] + dx, p[1] + dy)
        return q in arr and arr[q]

    quad = [(0, 0), (1, 0), (0, 1), (1, 1)]
    south_west = [(0, -1), (0, 0), (-1, -1), (-1, 0)]
    nt1 = s * rounds
    for nt in range(1, nt1 + 1):
        k4 = 4 * (min(-(-nt // s) * s, nt1) - nt)
        on = [lambda p, j=j: lvl[p] <= k4 + levels[j] for j in range(5)]
        sf = {p: on[0](p) and all(v(chp, p, a, b) for a, b in quad) and v(uv, p, 0, 0) for p in pts}
        pr = {p: on[1](p) and all(v(sf, p, a, b) for a, b in south_west) and v(chp, p, 0, 0) for p in pts}
        wa = {p: on[2](p) and v(uv, p, 0, 0) and v(uv, p, -1, 0) for p in pts}
        wb = {p: on[2](p) and v(uv, p, 0, 0) and v(uv, p, 0, -1) for p in pts}
        gx = {p: on[2](p) and v(pr, p, 0, 0) and v(pr, p, 1, 0) for p in pts}
        gy = {p: on[2](p) and v(pr, p, 0, 0) and v(pr, p, 0, 1) for p in pts}
        sch = {p: on[2](p) and all(v(pr, p, a, b) for a, b in quad) for p in pts}
        mom = {p: on[3](p) and sch[p] and v(uv, p, 0, 0) and v(wa, p, 0, 0) and v(wa, p, 1, 0)
               and v(wb, p, 0, 0) and v(wb, p, 0, 1) and v(gx, p, 0, 0) and v(gx, p, 0, 1)
               and v(gy, p, 0, 0) and v(gy, p, 1, 0) and all(v(pr, p, a, b) for a, b in quad) for p in pts}
        cor = {p: on[4](p) and all(v(mom, p, a, b) for a, b in south_west) and v(chp, p, 0, 0) for p in pts}
        for p in owned:
            assert mom[p] and cor[p], f"tile {tw}x{th}x{s}, substep {nt}, point {p}"
        chp, uv = cor, mom
        if nt % s == 0 and nt < nt1:  # the exchange reloads every point the tile does not own
            chp = {p: chp[p] if p in owned else True for p in pts}
            uv = {p: uv[p] if p in owned else True for p in pts}


TILES = [(4, 4), (5, 3), (16, 17), (7, 8), (24, 11), (3, 9), (1, 1), (2, 5)]


@pytest.mark.parametrize("s", [1, 2, 3, 4])
def test_stage_areas_cover_what_owned_points_read(s):
    levels = kernel_stage_levels()
    for tw, th in TILES:
        run_rounds(tw, th, s, levels, 2 * s)


def test_narrower_areas_lose_needed_points():
    levels = kernel_stage_levels()
    with pytest.raises(AssertionError):
        run_rounds(8, 8, 2, levels, 2 * 2 - 1)
    for j in range(5):
        if levels[j] == 0:
            continue
        smaller = list(levels)
        smaller[j] -= 1
        with pytest.raises(AssertionError):
            run_rounds(8, 8, 2, smaller, 4)
