"""Tracks in memory and the storm program (CPU: no device needed).

A track handed over in memory (slg_batch_open_tracks) must give the schedule a
trk file of the same values gives (inputn2.f:110-605 reads the file; the
library reads either into the same rows).  The per-step values compared are the
host schedule (SLG_SCHED_HOST), itself pinned to the reference build's storm
dumps by test_host_golden.
"""
import os

import numpy as np
import pytest

import helpers
from helpers import sloshgpu, synth


def _open_pair(tmp_path, b, tracks, prec=64):
    case = synth.write_case(str(tmp_path / "c"), b, tracks[0])
    paths = [case["trk"]]
    for k, t in enumerate(tracks[1:]):
        p = str(tmp_path / ("t%d.trk" % k))
        with open(p, "w") as fh:
            fh.write(synth.trk_text(t))
        paths.append(p)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    F = sloshgpu.Batch(B, paths, precision=prec, sched_host=True)
    M = sloshgpu.Batch(B, precision=prec, sched_host=True, tracks=[synth.track_dict(t) for t in tracks])
    return B, F, M


@pytest.mark.parametrize("prec", [32, 64])
def test_memory_tracks_equal_files(tmp_path, prec):
    b = synth.make_basin("mem", " ", 40, 34, seed=3, dt=60.0)
    rng = np.random.default_rng(7)
    tracks = [synth.straight_track(b, hours=12, landfall=6),
              synth.spline_track(b, seed=4, hours=30),
              synth.forecast_track(b, rng, hours=24)]
    B, F, M = _open_pair(tmp_path, b, tracks, prec)
    assert F.nsteps == M.nsteps
    for s in range(len(tracks)):
        a, m = F.schedule(s, from_device=False), M.schedule(s, from_device=False)
        assert a.shape == m.shape and a.shape[0] == F.nsteps[s]
        assert np.array_equal(a, m), "storm %d" % s
    F.close()
    M.close()
    B.close()


def test_memory_track_rows_checked(tmp_path):
    b = synth.make_basin("mem", " ", 24, 20, seed=3, dt=60.0)
    t = synth.track_dict(synth.straight_track(b, hours=6, landfall=3))
    case = synth.write_case(str(tmp_path / "c"), b, synth.straight_track(b, hours=6, landfall=3))
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    bad = dict(t)
    bad["delp"] = list(t["delp"])
    bad["delp"][2] = 300.0          # above INPUTN2's 213 mb bound
    with pytest.raises(sloshgpu.SloshError, match="out of bounds"):
        sloshgpu.Batch(B, tracks=[bad])
    short = dict(t, end_hour=t["hour"][-1] + 5)
    with pytest.raises(sloshgpu.SloshError, match="does not cover"):
        sloshgpu.Batch(B, tracks=[short])
    # a 999 row ends the track as it ends a file
    cut = {k: (list(v) + [999.0] + list(v[:2]) if isinstance(v, list) else v) for k, v in t.items()}
    X = sloshgpu.Batch(B, sched_host=True, tracks=[cut])
    Y = sloshgpu.Batch(B, sched_host=True, tracks=[t])
    assert np.array_equal(X.schedule(0, False), Y.schedule(0, False))
    X.close()
    Y.close()
    B.close()


def test_host_values_refused_without_host_schedule(tmp_path):
    b = synth.make_basin("mem", " ", 24, 20, seed=3, dt=60.0)
    case = synth.write_case(str(tmp_path / "c"), b, synth.straight_track(b, hours=6, landfall=3))
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, [case["trk"]])
    with pytest.raises(sloshgpu.SloshError, match="SLG_SCHED_HOST"):
        X.schedule(0, from_device=False)
    X.close()
    B.close()
