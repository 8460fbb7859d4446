"""MEOW track sets in memory (slg_meow_seeds / slg_meow_tracks; CPU only).

Seed points: the reference places them in its Tcl GUI (conMeowGenCmd,
sorc/rexout.cd/convert.c:1181-1288), which cannot be built here (no Tcl
headers).  oracle/_ref/meowref walks the same steps over the reference's own
compiled projection code (convert.c load_bsndta / cl2cnf / sll2xy / sxy2ll /
cnf2cl / pq2xy); the product's points must equal its output bit for bit, on
polar, elliptic and hyperbolic basins, along the plane and along great
circles.  Tracks through the seeds: CreateMEOWTrk (gui/tclsrc/meow.tcl:290-344)
steps its points with halo_DistCompute, whose source is not in the reference;
the product steps them with Earth_Jump (convert.c:1127-1151).  They are pinned
by what a great-circle step fixes: the seed at the landfall hour, and
consecutive points speed miles apart on the heading.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import helpers
from helpers import sloshgpu, synth

MEOWREF = os.path.join(helpers.REPO, "oracle", "_ref", "meowref")


def _basins():
    return [synth.config1()[0], synth.config2()[0], synth.config4_basins()[4]]


BASINS = _basins()


@pytest.mark.skipif(not os.path.exists(MEOWREF), reason="oracle/_ref not built")
@pytest.mark.parametrize("xy", [0, 1])
@pytest.mark.parametrize("bi", range(len(BASINS)), ids=[b.gtype.replace(" ", "p") for b in BASINS])
def test_meow_seeds_equal_reference_projection(tmp_path, bi, xy):
    b = BASINS[bi]
    case = synth.write_case(str(tmp_path), b, synth.straight_track(b))
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    for d, apart, num, r34 in ((0.0, 7.5, -1, 20.0), (45.0, 12.0, -1, 0.0), (300.0, 5.0, 9, 35.0),
                               (181.5, 9.0, 4, 60.0)):
        m = sloshgpu.meow(xyflag=xy, lat=b.alto + 0.2, lon=b.alno - 0.3, dir=d, apart=apart, r34=r34, num=num)
        idx, la, lo = sloshgpu.meow_seeds(B, m)
        out = subprocess.run([MEOWREF, case["bnt"], b.name, str(xy), repr(m.lat), repr(m.lon), repr(d),
                              repr(apart), str(num), repr(r34)], capture_output=True, text=True, check=True)
        ref = [line.split() for line in out.stdout.splitlines() if line.strip()]
        assert len(ref) == len(idx), (d, apart, num)
        for r, i, a, o in zip(ref, idx, la, lo):
            assert int(r[0]) == i and float(r[1]) == a and float(r[2]) == o, (d, r, i, a, o)
        if num != -1:
            assert len(idx) <= num
    B.close()


def _gc_dist_mi(la1, lo1, la2, lo2):
    """great-circle distance on Earth_Jump's sphere (60 nm per degree, 1.151 mi per nm)"""
    p1, p2 = math.radians(la1), math.radians(la2)
    dl = math.radians(lo2 - lo1)
    c = math.sin(p1) * math.sin(p2) + math.cos(p1) * math.cos(p2) * math.cos(dl)
    return math.degrees(math.acos(max(-1.0, min(1.0, c)))) * 60.0 * 1.151


def _bearing(la1, lo1, la2, lo2):
    p1, p2 = math.radians(la1), math.radians(la2)
    dl = math.radians(lo2 - lo1)
    y = math.sin(dl) * math.cos(p2)
    x = math.cos(p1) * math.sin(p2) - math.sin(p1) * math.cos(p2) * math.cos(dl)
    return math.degrees(math.atan2(y, x)) % 360.0


@pytest.mark.parametrize("xy", [0, 1])
def test_meow_tracks_geometry(tmp_path, xy):
    b = BASINS[1]
    case = synth.write_case(str(tmp_path), b, synth.straight_track(b))
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    m = sloshgpu.meow(xyflag=xy, lat=b.alto, lon=b.alno, dir=330.0, apart=15.0, r34=10.0, speed=13.8,
                      delp=70.0, rmw=25.0, landfall_hour=40, begin_hour=10, end_hour=60)
    idx, la, lo = sloshgpu.meow_seeds(B, m)
    tracks = sloshgpu.meow_tracks(B, m)
    assert len(tracks) == len(idx) > 3
    for t, a, o in zip(tracks, la, lo):
        assert list(t["hour"]) == list(range(1, sloshgpu.MEOW_HOURS + 1))
        # the seed at the landfall hour (the track's longitude east positive)
        assert abs(t["lat"][39] - a) < 1e-9 and abs(t["lon"][39] + o) < 1e-9
        for k in range(0, 99, 7):
            d = _gc_dist_mi(t["lat"][k], t["lon"][k], t["lat"][k + 1], t["lon"][k + 1])
            assert abs(d - 13.8) < 1e-6
        assert abs((_bearing(t["lat"][39], t["lon"][39], t["lat"][40], t["lon"][40]) - 330.0 + 180) % 360 - 180) < 1e-3
        assert np.all(t["delp"] == 70.0) and np.all(t["rmw"] == 25.0)
    # the set opens as a batch from memory, schedules on the host as a check
    X = sloshgpu.Batch(B, precision=64, sched_host=True, tracks=tracks[:4])
    assert all(n == int((60 - 10) * 3600 / 30.0 + 1 / 30.0) for n in X.nsteps)
    X.close()
    B.close()


def test_meow_refuses_bad_parameters(tmp_path):
    b = BASINS[0]
    case = synth.write_case(str(tmp_path), b, synth.straight_track(b))
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    with pytest.raises(sloshgpu.SloshError, match="apart"):
        sloshgpu.meow_seeds(B, sloshgpu.meow(lat=b.alto, lon=b.alno, apart=0.0))
    with pytest.raises(sloshgpu.SloshError, match="room"):
        sloshgpu.meow_seeds(B, sloshgpu.meow(lat=b.alto, lon=b.alno, apart=1.0, r34=50.0), cap=3)
    B.close()
