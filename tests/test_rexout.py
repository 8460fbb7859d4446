"""rex point series (rexout style 0) written by the product's writer
(slg_pnt_*), against the reference's rexout (sorc/rexout.cd, built from its
sources into oracle/_ref) reading a rex file of the same frames.

The frames are the oracle's encoding (pinned to the reference's rex writer) of
seeded fields with dry cells, clipped values, negative heights and long runs;
the file around them is the product's rex writer (byte-equal to the
reference's).  The values the product's writer prints are the rex value rule
applied to the same fields; rexout decodes them from the frames.  Dates come
from model clocks with a fractional second and across a month boundary.
"""
import os

import numpy as np
import pytest

import rexutil
import sloshgpu

pytestmark = pytest.mark.skipif(not rexutil.have_rexout(), reason="reference rexout (oracle/_ref) not built")


def _values(hb, zb, xy):
    """The rex frame's value at squares (I, J): tenths of a foot, 999 where
    HB + ZB == 0, clipped to -320..700 (rex.c:470-490)."""
    out = []
    for x, y in xy:
        h, z = float(hb[y - 1, x - 1]), float(zb[y - 1, x - 1])
        v = 999 if h + z == 0.0 else int(h * 10 + .5)
        out.append(max(-320, v) if v == 999 or v <= 700 else 700)
    return np.array(out, dtype=np.int16)


@pytest.mark.parametrize("abbrev", ["tny", " ab ", "lake"])
def test_point_series_equals_reference_rexout(tmp_path, abbrev):
    rng = np.random.default_rng(sum(abbrev.encode()))
    im, jm = 37, 29
    zb = None
    frames, fields = [], []
    clock0 = 1_756_684_740.0 - 3 * 600 + 0.5          # 2025-08-31 23:29:00.5
    for k, kind in enumerate(["mixed", "mixed", "dry", "flat", "mixed"]):
        hb, zb = rexutil.frame_fields(rng, im, jm, kind, zb)
        t6 = rng.normal(0.0, 50.0, size=6)
        frames.append((t6, clock0 + 600 * k, rexutil.oracle_frame(hb, zb)))
        fields.append(hb)
    xy = [(1, 1), (im - 1, jm - 1), (1, jm - 1), (im - 1, 1), (5, 7), (20, 3), (30, 25), (12, 12)]
    xy += [(int(x), int(y)) for x, y in zip(rng.integers(1, im, 20), rng.integers(1, jm, 20))]
    for h in (85.0, -41.0):   # squares whose values are clipped (to 70.0 and -32.0 ft)
        j, i = np.argwhere(fields[0][: jm - 1, : im - 1] == h)[0]
        xy.append((int(i) + 1, int(j) + 1))
    root = str(tmp_path)
    rex = os.path.join(root, "s.rex")
    comment = "FileVersion=2013"
    trk = os.path.join(root, "s.trk")
    rexutil.trk_2013(trk, 6, rng)
    rexutil.product_rex(rex, zb, frames, trk, rexutil.oracle_frame(fields[-1], zb), comment=comment, abbrev=abbrev)
    ref = rexutil.reference_rexout(rex, rexutil.rexout_abbrev(abbrev), xy, root)
    mine_path = os.path.join(root, "pts.txt")
    P = sloshgpu.PointSeries(mine_path, abbrev, comment, rex, xy)
    for (t6, clock, _), hb in zip(frames, fields):
        P.row(clock, _values(hb, zb, xy))
    P.close()
    mine = open(mine_path).read()
    rexutil.compare_series(mine, ref, comment)
    # the values span the cases the rule has
    v = rexutil.rexout_values(ref)
    assert (v == 999).any() and (v == 700).any() and (v == -320).any() and (v < 0).any()


def test_point_series_planted_mistakes(tmp_path):
    """A value rounded the other way, a date off by one second or a point
    swapped each make the texts differ."""
    rng = np.random.default_rng(3)
    im, jm = 21, 17
    hb, zb = rexutil.frame_fields(rng, im, jm, "mixed")
    clock = 1_700_000_000.0
    root = str(tmp_path)
    rex = os.path.join(root, "s.rex")
    trk = os.path.join(root, "s.trk")
    rexutil.trk_2013(trk, 6, rng)
    rexutil.product_rex(rex, zb, [(np.zeros(6), clock, rexutil.oracle_frame(hb, zb))], trk,
                        rexutil.oracle_frame(hb, zb))
    xy = [(2, 3), (4, 5), (10, 9)]
    ref = rexutil.reference_rexout(rex, "tny", xy, root)

    def series(c, vals, pts):
        path = os.path.join(root, "m.txt")
        P = sloshgpu.PointSeries(path, "tny", "FileVersion=2013", rex, pts)
        P.row(c, vals)
        P.close()
        return open(path).read()

    good = _values(hb, zb, xy)
    rexutil.compare_series(series(clock, good, xy), ref, "FileVersion=2013")
    bad = good.copy()
    bad[0] += 1
    for text in [series(clock, bad, xy), series(clock + 1, good, xy), series(clock, good, xy[::-1])]:
        with pytest.raises(AssertionError):
            rexutil.compare_series(text, ref, "FileVersion=2013")
