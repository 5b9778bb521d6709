"""Tiles of the two-step pass with two x-cells a lane (k_yee2p), on the CPU.

The layout is the engine's own (iris_pair_layout: the host arithmetic a
handle uses, no device).  Each lane's place in a tile is restated from the
kernel (iris_yee2p.h: segment = lane / sl, row = wave * ns + segment, cell a
at x = x0 + bx * ox - 1 + 2 * lane-in-row, y = by * (12 ns - 3) + row; lanes
past the last segment stand in for the first lane of the last row; own
lanes are 1 .. ox/2 of rows 1 .. 12 ns - 3 inside the grid).  Checked for
every layout: each cell of the plane is owned by exactly one lane, every
lane's footprint (and the periodic source it reads) lies within one period
of the grid, and every LDS entry a lane writes or reads is inside the
kernel's 12 x 64 entries.  The GPU tests check the kernel's results bit for
bit against single steps on these layouts.
"""
import pytest

from pyiris.capi import pair_layout

R = 12  # waves a workgroup


def _parts(L):
    out = []
    if L["tx"]:
        out.append((L["tx"], L["ty"], 0, L["ox"], L["sl"], L["ns"]))
    if L["txn"]:
        out.append((L["txn"], L["tyn"], L["x0n"], L["oxn"], L["sln"], L["nsn"]))
    return out


def _check(nx, ny, L):
    owned = set()
    for tx, ty, x0, ox, sl, ns in _parts(L):
        assert ox % 2 == 0 and sl == ox // 2 + 2 and 1 <= ns <= 64 // sl
        ry = R * ns
        oy = ry - 3
        for bx in range(tx):
            for by in range(ty):
                for w in range(R):
                    for lane in range(64):
                        seg = lane // sl
                        idle = seg >= ns
                        cx = 0 if idle else lane - seg * sl
                        cy = ry - 1 if idle else w * ns + seg
                        xg = x0 + bx * ox - 1 + 2 * cx
                        yg = by * oy + cy
                        # single periodic wrap: the source cell is within one period
                        assert 1 - nx <= xg and xg + 1 <= 2 * nx, (nx, ny, bx, lane)
                        assert 1 - ny <= yg <= 2 * ny, (nx, ny, by, w, lane)
                        li, lp, lm = cy * sl + cx, min(cy + 1, ry - 1) * sl + cx, max(cy - 1, 0) * sl + cx
                        assert 0 <= min(li, lp, lm) and max(li, lp, lm) < R * 64
                        if (not idle) and 1 <= cx <= ox // 2 and 1 <= cy <= oy and xg <= nx and yg <= ny:
                            assert xg % 2 == 1 and xg >= 1 and yg >= 1
                            for x in (xg, xg + 1):
                                assert (x, yg) not in owned, (nx, ny, x, yg)
                                owned.add((x, yg))
    assert len(owned) == nx * ny


@pytest.mark.parametrize("ny", [1, 3, 5, 8, 9, 21, 40, 97])
def test_every_cell_owned_once_within_one_period(ny):
    seen = 0
    for nx in list(range(2, 264, 2)) + [496, 512]:
        for only in (0, 124, 60, 28):
            L = pair_layout(nx, ny, only)
            if L is None:
                continue
            seen += 1
            _check(nx, ny, L)
    # a 12-row tile needs its footprint rows within two periods of y
    assert seen > 100 if ny >= 6 else seen == 0


def test_layout_choices():
    # 512 x 512: 8 x 25 main tiles of 60 cells a row (two rows a wave) and 16
    # narrow tiles of 32 cells over the last 32 columns
    L = pair_layout(512, 512)
    assert (L["ox"], L["sl"], L["ns"], L["tx"], L["ty"]) == (60, 32, 2, 8, 25)
    assert (L["oxn"], L["txn"], L["tyn"], L["x0n"]) == (32, 1, 16, 480)
    # narrower than every main tile: narrow tiles only
    L = pair_layout(40, 21)
    assert L["tx"] == 0 and L["txn"] >= 1
    assert pair_layout(41, 21) is None          # odd nx: pairs would straddle the period
    assert pair_layout(512, 512, 124)["ox"] == 124
