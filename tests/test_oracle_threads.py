"""The oracle may step several runs at once, one per thread (the full-size GPU
tests check sampled storms against it that way): eight storms stepped in
threads give exactly the fields of the same storms stepped one after
another.  (Its SMPT2G corner-smoothing lists are per thread.)"""
import os

import numpy as np

import helpers
from helpers import synth


def test_threaded_oracle_equals_sequential(tmp_path):
    root = str(tmp_path)
    b, t = synth.tiny(seed=31, ni=40, nj=34, hours=8)
    case = helpers.make_case(root, b, t)
    tracks = synth.meow_tracks(b, n=1024, hours=8, landfall=4)[::97][:8]
    jobs = []
    for k, tr in enumerate(tracks):
        p = os.path.join(root, "t%d.trk" % k)
        with open(p, "w") as fh:
            fh.write(synth.trk_text(tr))
        jobs.append((helpers.track_case(case, p), "64", [100, 400], os.path.join(root, "o%d" % k)))
    seq = [helpers.oracle_checkpoints(*j) for j in jobs]
    par = helpers.oracle_many(jobs, workers=8)
    for a, c in zip(seq, par):
        for st in (100, 400):
            for f in helpers.FIELDS:
                assert np.array_equal(a[st][f], c[st][f]), (st, f)
