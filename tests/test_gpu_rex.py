"""rex v2 frames encoded on the device (slg_batch_rex_frames), through the C ABI.

* Every storm's frame of HHB = min(HB, HMX) (tmstpc2.f:74-85) and of HMX
  equals the oracle's encoding (oracle/orc_rex.c, itself pinned to the
  reference's rex.c / pack.c) of the same fields, at FP32 and FP64, on a small
  basin and on the 250 x 300 polar basin (75 000 cells: many scan chunks).
* HHB is refused after a step that filtered on the fused path (HB before
  FILTER is gone), as slg_batch_get("HBS") refuses it.
* End to end: an FP64 batch run written with the product's rex writer, frames
  every 10 model minutes as PerformRun saves them (slosh2.c:1346-1366,
  2366-2386, 2520-2580) and the envelope frame at the end (slosh2.c:2713),
  equals the reference CLI's rex file byte for byte.
"""
import datetime
import os

import numpy as np
import pytest

import helpers
import rexutil
import sloshgpu
import synth

pytestmark = pytest.mark.gpu


def _hhb(X, s):
    hmx, hbs = X.get(s, "HMX"), X.get(s, "HBS")
    return np.where(hmx < hbs, hmx, hbs), hmx


def _check(X, zb, storms):
    hhb = X.rex_frames("HHB")
    env = X.rex_frames("HMX")
    assert len(hhb) == len(env) == X.n
    for s in storms:
        h, m = _hhb(X, s)
        assert hhb[s] == rexutil.oracle_frame(h, zb), ("HHB", s)
        assert env[s] == rexutil.oracle_frame(m, zb), ("HMX", s)


@pytest.mark.parametrize("prec", [32, 64])
def test_rex_frames_match_oracle_small(tmp_path, prec):
    b, t = synth.tiny(seed=71, ni=40, nj=34, hours=8)
    root = str(tmp_path)
    case = helpers.make_case(root, b, t)
    rng = np.random.default_rng(71)
    trks = [case["trk"]]
    for k in range(5):
        p = os.path.join(root, "r%d.trk" % k)
        with open(p, "w") as fh:
            fh.write(synth.trk_text(synth.random_track(b, rng, hours_range=(6, 8))))
        trks.append(p)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    zb = B.field("ZB")
    X = sloshgpu.Batch(B, trks, precision=prec)
    X.upload()
    X.set_path(sloshgpu.PATH_STAGED, 16)
    done = 0
    for n in (1, 37, 250):
        X.step(n - done)
        done = n
        X.sync()
        print("rex small prec %d step %d" % (prec, n), flush=True)
        _check(X, zb, range(X.n))
    X.close()
    B.close()


def test_rex_frames_match_oracle_250x300(tmp_path):
    b = synth.config3_basin()
    tr = synth.meow_tracks(b, n=1024)
    root = str(tmp_path)
    case = helpers.make_case(root, b, tr[0])
    trks = [case["trk"]]
    for k in (200, 480, 777):
        p = os.path.join(root, "m%d.trk" % k)
        with open(p, "w") as fh:
            fh.write(synth.trk_text(tr[k]))
        trks.append(p)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    zb = B.field("ZB")
    X = sloshgpu.Batch(B, trks)
    X.upload()
    X.set_path(sloshgpu.PATH_STAGED, 16)
    X.step(40)
    X.sync()
    print("rex 250x300 step 40", flush=True)
    _check(X, zb, range(X.n))
    X.close()
    B.close()


def test_rex_hhb_refused_after_fused_filter(tmp_path):
    b, t = synth.tiny(seed=72, ni=30, nj=26, hours=6)
    case = helpers.make_case(str(tmp_path), b, t)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, [case["trk"]])
    X.upload()
    X.set_path(sloshgpu.PATH_FUSED, 16)
    X.step(19)   # ITIME 19: (ITIME+1) mod INCSM == 0, a filter step (tmstpc2.f:100-118)
    X.sync()
    with pytest.raises(sloshgpu.SloshError):
        X.rex_frames("HHB")
    env = X.rex_frames("HMX")
    assert env[0] == rexutil.oracle_frame(X.get(0, "HMX"), B.field("ZB"))
    X.close()
    B.close()


@pytest.mark.skipif(not os.path.exists(helpers.REF_CLI), reason="reference CLI (oracle/_ref) not built")
@pytest.mark.parametrize("cid", [0, 1])
def test_batched_rex_file_equals_reference_cli(tmp_path, cid):
    b, t = [synth.tiny(seed=5, ni=30, nj=26, hours=6), synth.tiny(seed=9, ni=28, nj=24, hours=6)][cid]
    root = str(tmp_path)
    case = helpers.make_case(root, b, t)
    rexutil.trk_canonical(t, case["trk"])
    _, rex_ref = helpers.run_cli(helpers.REF_CLI, root, case["basin"], "ref")
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, [case["trk"]], precision=64)
    X.upload()
    y, mo, d, hr, mi = X.start_date(0)
    clock = (datetime.date(y, mo, d) - datetime.date(1970, 1, 1)).days * 86400.0
    clock = clock + hr * 3600.
    clock = clock + mi * 60
    path = os.path.join(root, "batch.rex")
    R = sloshgpu.Rex(path, "FileVersion=2013", rexutil.rex_abbrev(case["basin"]), B.imxb, B.jmxb)
    rex_clock, del_t, hhb = clock, 0.0, None
    for n in range(1, X.nsteps[0] + 1):
        f_pass = n == 1 or clock + del_t >= rex_clock
        X.set_path(sloshgpu.PATH_STAGED if f_pass else sloshgpu.PATH_FUSED, 16)
        X.step(1)
        if f_pass:
            hhb = X.rex_frames("HHB")[0]
        tr = X.track(0, n)
        del_t = tr[2]
        clock = clock + del_t
        if clock >= rex_clock:
            R.frame([tr[3], -tr[4], tr[7], tr[8], tr[5], tr[6]], clock, hhb)
            rex_clock += 10 * 60
    R.env(case["trk"], X.rex_frames("HMX")[0])
    R.close()
    X.close()
    B.close()
    mine, ref = open(path, "rb").read(), open(rex_ref, "rb").read()
    assert len(mine) == len(ref)
    assert mine == ref
