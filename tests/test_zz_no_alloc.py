"""slg_batch_envelopes_noalloc, the allocation-free variant of
slg_batch_envelopes (its device staging is a fixed scratch array of the
library), gives the same per-run envelopes and MOM as the allocating entry
point, through the C ABI: on a batch of 300 storms on the 250x300 basin (two
chunks of the scratch, the MOM carried across them) and then, in the same
process, on a smaller batch (the scratch keeps the first call's contents, so
this checks that everything read is rewritten first), at both precisions."""
import ctypes

import numpy as np
import pytest

import helpers
from helpers import synth
import sloshgpu

pytestmark = pytest.mark.gpu


def _noalloc(X):
    L = sloshgpu.lib()
    L.slg_batch_envelopes_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p,
                                              ctypes.c_int]
    nj, ni = X.basin.jmxb, X.basin.imxb
    env = np.full((X.n, nj, ni), -7, dtype=np.int16)
    mom = np.full((nj, ni), -7, dtype=np.int16)
    err = ctypes.create_string_buffer(512)
    rc = L.slg_batch_envelopes_noalloc(X.h, env.ctypes.data_as(ctypes.c_void_p), mom.ctypes.data_as(ctypes.c_void_p),
                                       err, 512)
    assert rc == 0, err.value
    return env, mom


def _batch(tmp, b, tracks, prec, steps):
    case = helpers.make_case(tmp, b, tracks[0])
    trks = [case["trk"]]
    for k, t in enumerate(tracks[1:]):
        p = "%s/n%04d.trk" % (tmp, k)
        open(p, "w").write(synth.trk_text(t))
        trks.append(p)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, trks, precision=prec)
    X.set_path(sloshgpu.PATH_FUSED, 32)
    X.upload()
    X.step(steps)
    X.sync()
    return B, X


@pytest.mark.parametrize("precision", [32, 64])
def test_envelopes_noalloc_equal_allocating(precision, tmp_path):
    big = synth.config3_basin()
    tracks = synth.meow_tracks(big, n=1024)[::3][:300]
    print("noalloc: 300 storms on 250x300, 1200 steps, FP%d" % precision, flush=True)
    B, X = _batch(str(tmp_path / "big"), big, tracks, precision, 1200)
    e1, m1 = X.envelopes()
    e2, m2 = _noalloc(X)
    assert np.array_equal(e1, e2) and np.array_equal(m1, m2)
    assert (e1 != 999).any() and (m1 != 999).any()
    X.close()
    B.close()
    small, t = synth.tiny(seed=8, ni=30, nj=26, hours=6)
    tr = synth.meow_tracks(small, n=1024, hours=6, landfall=3)[::97][:11]
    print("noalloc: 11 storms on 30x26, FP%d" % precision, flush=True)
    B, X = _batch(str(tmp_path / "small"), small, [t] + tr, precision, 300)
    e1, m1 = X.envelopes()
    e2, m2 = _noalloc(X)
    assert np.array_equal(e1, e2) and np.array_equal(m1, m2)
    X.close()
    B.close()
