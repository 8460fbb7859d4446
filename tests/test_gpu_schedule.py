"""Per-step storm values generated on the device (k_sched) against the host's.

The host walks each storm once (INPUTN2, STERON, SETCMP, then TTIMER and
STMVAL every step with FORCE1's table refreshes: ttimer.f:33-42,
stmval.f:95-236, force1.f:55-183) and keeps a storm program; the device expands
it into the SL_NPAR values of every step.  The host schedule (SLG_SCHED_HOST)
is pinned to the reference build's storm dumps (test_host_golden); here the
device expansion must equal it bit for bit, on straight, spline, MEOW, uniform
random and forecast tracks (the last refresh FORCE1's tables many times a
run), on a basin whose time step changes mid-run (DLTIN(2) != DLTIN(1):
stmval.f:108-135), and in both precisions.
"""
import os

import numpy as np
import pytest

import helpers
from helpers import sloshgpu, synth

pytestmark = pytest.mark.gpu


def _cases():
    out = []
    b1, t1 = synth.config1()
    out.append(("c1_straight", b1, [t1]))
    b2, t2 = synth.config2()
    out.append(("c2_spline", b2, [t2]))
    b3 = synth.config3_basin()
    out.append(("c3_meow", b3, synth.meow_tracks(b3, n=1024)[::32]))
    rng = np.random.default_rng(41)
    b4 = synth.config4_basins()[1]
    out.append(("c4_random", b4, [synth.random_track(b4, rng) for _ in range(24)]))
    out.append(("c5_forecast", b4, [synth.forecast_track(b4, rng) for _ in range(24)]))
    bd = synth.make_basin("dlt", " ", 60, 50, seed=6, dt=60.0)
    bd.dltin = (60.0, 30.0, 30.0)
    out.append(("dltin_switch", bd, [synth.straight_track(bd, hours=24, landfall=12),
                                     synth.straight_track(bd, 8.0, 10.0, 90.0, 15.0, hours=30, landfall=14)]))
    return out


CASES = _cases()


def _write(root, b, tracks):
    case = synth.write_case(root, b, tracks[0])
    paths = [case["trk"]]
    for k, t in enumerate(tracks[1:]):
        p = os.path.join(root, "t%03d.trk" % k)
        with open(p, "w") as fh:
            fh.write(synth.trk_text(t))
        paths.append(p)
    return case, paths


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("name,b,tracks", CASES, ids=[c[0] for c in CASES])
def test_device_schedule_equals_host(tmp_path, name, b, tracks, prec):
    print("schedule case %s, FP%d, %d storms" % (name, prec, len(tracks)), flush=True)
    case, paths = _write(str(tmp_path), b, tracks)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    H = sloshgpu.Batch(B, paths, precision=prec, sched_host=True)
    D = sloshgpu.Batch(B, paths, precision=prec)
    H.upload()
    D.upload()
    assert H.nsteps == D.nsteps
    events = 0
    for s in range(len(paths)):
        # INTLHT's initial water: the host's (rounded to R) and the device's
        for f in ("HB", "HMX", "UB", "VB"):
            assert np.array_equal(H.get(s, f), D.get(s, f)), "%s storm %d initial %s" % (name, s, f)
        host = H.schedule(s, from_device=False)
        assert np.array_equal(H.schedule(s, from_device=True), host)
        dev = D.schedule(s, from_device=True)
        bad = np.argwhere(dev != host)
        assert bad.size == 0, "%s storm %d: first difference at step %d value %d" % (name, s, bad[0][0] + 1,
                                                                                    bad[0][1])
        # the table set in force changes wherever FORCE1 refreshed the tables
        events += int(np.count_nonzero(np.diff(host[:, 21])))
    if name == "c5_forecast":
        assert events > 24, "forecast tracks refresh FORCE1's tables during their runs"
    if name == "dltin_switch":
        assert len(np.unique(H.schedule(0, from_device=False)[:, 1])) == 2, "DELT changes mid-run"
    H.close()
    D.close()
    B.close()


@pytest.mark.parametrize("prec", [32, 64])
def test_memory_tracks_run_equal_files(tmp_path, prec):
    """Runs opened from tracks in memory (device schedule) step to the same
    fields as runs opened from the same tracks' trk files (host schedule)."""
    b = synth.config3_basin()
    tracks = synth.meow_tracks(b, n=1024)[5::128]
    case, paths = _write(str(tmp_path), b, tracks)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    F = sloshgpu.Batch(B, paths, precision=prec, sched_host=True)
    M = sloshgpu.Batch(B, precision=prec, tracks=[synth.track_dict(t) for t in tracks])
    for X in (F, M):
        X.upload()
        X.set_path(sloshgpu.PATH_FUSED, 32)
        X.step(300)
        X.sync()
    for s in range(len(tracks)):
        for f in ("HB", "UB", "VB", "HMX"):
            assert np.array_equal(F.get(s, f), M.get(s, f)), "storm %d field %s" % (s, f)
    F.close()
    M.close()
    B.close()


@pytest.mark.parametrize("prec", [32, 64])
def test_meow_set_from_memory(tmp_path, prec):
    """A MEOW track set generated in memory (slg_meow_tracks) opens and steps
    with the device schedule exactly as with the host's."""
    b = synth.config3_basin()
    case = synth.write_case(str(tmp_path), b, synth.straight_track(b))
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    m = sloshgpu.meow(xyflag=1, lat=b.alto, lon=b.alno, dir=315.0, apart=12.0, r34=15.0, speed=11.5, delp=60.0,
                      rmw=20.0, landfall_hour=40, begin_hour=16, end_hour=64)
    tracks = sloshgpu.meow_tracks(B, m)
    print("MEOW set of %d tracks, FP%d" % (len(tracks), prec), flush=True)
    assert len(tracks) > 8
    H = sloshgpu.Batch(B, precision=prec, sched_host=True, tracks=tracks)
    D = sloshgpu.Batch(B, precision=prec, tracks=tracks)
    for X in (H, D):
        X.upload()
        X.set_path(sloshgpu.PATH_FUSED, 32)
    for s in range(len(tracks)):
        assert np.array_equal(D.schedule(s, True), H.schedule(s, False)), "storm %d" % s
    for X in (H, D):
        X.step(240)
        X.sync()
    for s in range(0, len(tracks), 3):
        for f in ("HB", "UB", "VB", "HMX"):
            assert np.array_equal(H.get(s, f), D.get(s, f)), "storm %d field %s" % (s, f)
    H.close()
    D.close()
    B.close()
