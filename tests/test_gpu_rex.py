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


@pytest.mark.skipif(not (os.path.exists(helpers.REF_CLI) and rexutil.have_rexout()),
                    reason="reference CLI or rexout (oracle/_ref) not built")
def test_batched_point_series_equals_reference_rexout(tmp_path):
    """The rex point series of a batched FP64 run (slg_batch_rex_points at
    every rex save, written by slg_pnt_*) equals the reference's rexout over
    the reference CLI's own rex file of the same run."""
    b, t = synth.tiny(seed=5, ni=30, nj=26, hours=6)
    root = str(tmp_path)
    case = helpers.make_case(root, b, t)
    rexutil.trk_canonical(t, case["trk"])
    _, rex_ref = helpers.run_cli(helpers.REF_CLI, root, case["basin"], "ref")
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, [case["trk"]], precision=64)
    X.upload()
    y, mo, d, hr, mi = X.start_date(0)
    clock = (datetime.date(y, mo, d) - datetime.date(1970, 1, 1)).days * 86400.0 + hr * 3600. + mi * 60
    rng = np.random.default_rng(11)
    xy = [(1, 1), (B.imxb - 1, B.jmxb - 1), (1, B.jmxb - 1), (B.imxb - 1, 1)]
    xy += [(int(x), int(y)) for x, y in zip(rng.integers(1, B.imxb, 60), rng.integers(1, B.jmxb, 60))]
    abbrev = rexutil.rex_abbrev(case["basin"])
    series = os.path.join(root, "pts.txt")
    P = sloshgpu.PointSeries(series, abbrev, "FileVersion=2013", rex_ref, xy)
    rex_clock, del_t, vals = clock, 0.0, None
    for n in range(1, X.nsteps[0] + 1):
        f_pass = n == 1 or clock + del_t >= rex_clock
        X.set_path(sloshgpu.PATH_STAGED if f_pass else sloshgpu.PATH_FUSED, 16)
        X.step(1)
        if f_pass:
            vals = X.rex_points(xy, "HHB")[0]
        del_t = X.track(0, n)[2]
        clock = clock + del_t
        if clock >= rex_clock:
            P.row(clock, vals)
            rex_clock += 10 * 60
    P.close()
    X.close()
    B.close()
    ref = rexutil.reference_rexout(rex_ref, rexutil.rexout_abbrev(abbrev), xy, root)
    rexutil.compare_series(open(series).read(), ref, "FileVersion=2013")
    v = rexutil.rexout_values(ref)
    assert len(v) > 30 and (v == 999).any() and (v != 999).any()


@pytest.mark.skipif(not rexutil.have_rexout(), reason="reference rexout (oracle/_ref) not built")
@pytest.mark.parametrize("prec", [32, 64])
def test_rex_points_equal_rexout_of_device_frames(tmp_path, prec):
    """slg_batch_rex_points of HHB, HMX and HBMAX for every storm of a batch
    equals what rexout reads back from the device's own rex frames of the same
    fields; points outside the frame are refused."""
    root = str(tmp_path)
    rng = np.random.default_rng(prec)
    b, t = synth.tiny(seed=7, ni=41, nj=33, hours=8)
    case = helpers.make_case(root, b, t)
    trks = [case["trk"]]
    for k in range(2):
        p = os.path.join(root, "s%d.trk" % k)
        with open(p, "w") as fh:
            fh.write(synth.trk_text(synth.random_track(b, rng, hours_range=(6, 8))))
        trks.append(p)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, trks, precision=prec)
    X.upload()
    X.set_maxstat(1)
    X.set_path(sloshgpu.PATH_STAGED, 16)
    xy = [(1, 1), (B.imxb - 1, B.jmxb - 1)]
    xy += [(int(x), int(y)) for x, y in zip(rng.integers(1, B.imxb, 50), rng.integers(1, B.jmxb, 50))]
    abbrev = "tny"
    frames = {f: [[] for _ in trks] for f in ("HHB", "HMX", "HBMAX")}
    pts = {f: [] for f in frames}
    done = 0
    for n in (1, 60, 200):
        X.step(n - done)
        done = n
        print("rex points prec %d step %d" % (prec, n), flush=True)
        for f in frames:
            for s, bits in enumerate(X.rex_frames(f)):
                frames[f][s].append(bits)
            pts[f].append(X.rex_points(xy, f))
    zb = B.field("ZB")
    for f in frames:
        for s in range(len(trks)):
            rex = os.path.join(root, "%s%d.rex" % (f, s))
            rexutil.product_rex(rex, zb, [(np.zeros(6), 1.7e9 + 600 * k, bits) for k, bits in
                                          enumerate(frames[f][s])], trks[s], frames[f][s][-1], abbrev=abbrev)
            got = rexutil.rexout_values(rexutil.reference_rexout(rex, abbrev, xy, root))
            want = np.array([p[s] for p in pts[f]], dtype=np.int64)
            assert np.array_equal(got, want), (f, s)
    with pytest.raises(sloshgpu.SloshError):
        X.rex_points([(B.imxb, 1)], "HMX")
    with pytest.raises(sloshgpu.SloshError):
        X.rex_points([(1, 0)], "HMX")
    X.close()
    B.close()


def test_maxstat_values(tmp_path):
    """Windowed maxima on the device: -ZB at the start and after a reset, then
    the maximum of HB after each step, on every square (MaxStatInitGrid /
    MaxStatUpdateGrid, slosh2.c:1227-1316); FP32 runs keep them in double."""
    b, t = synth.tiny(seed=73, ni=30, nj=26, hours=6)
    case = helpers.make_case(str(tmp_path), b, t)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    zb = B.field("ZB")
    X = sloshgpu.Batch(B, [case["trk"]])
    X.upload()
    X.set_maxstat(True)
    sq = (slice(0, B.jmxb - 1), slice(0, B.imxb - 1))
    want = -1 * zb
    assert np.array_equal(X.get(0, "HBMAX")[sq], want[sq])
    for n in range(40):
        X.step(1)
        want = np.maximum(want, X.get(0, "HB"))
        if n == 25:
            X.maxstat_reset()
            want = -1 * zb
    got = X.get(0, "HBMAX")
    assert np.array_equal(got[sq], want[sq])
    X.close()
    B.close()


@pytest.mark.skipif(not os.path.exists(helpers.REF_CLI), reason="reference CLI (oracle/_ref) not built")
def test_batched_f_stat_windows_equal_reference_cli(tmp_path):
    """The reference's -f_stat mode from a batch: windows of 10 model minutes
    as PerformRun keeps them (slosh2.c:2366-2386, 2420-2520, the trailing
    window 2597-2640); every window's env file (ExtendEnvName, slosh2.c:
    1998-2011) and the rex file of the window maxima equal the reference
    CLI's."""
    import subprocess
    b, t = synth.tiny(seed=5, ni=30, nj=26, hours=6)
    root = str(tmp_path / "ref")
    case = helpers.make_case(root, b, t)
    rexutil.trk_canonical(t, case["trk"])
    r = subprocess.run([helpers.REF_CLI, "-basin", case["basin"], "-rootDir", root, "-trk", case["trk"],
                        "-rex", os.path.join(root, "out.rex"), "-env", os.path.join(root, "out.env"), "-verbose",
                        "0", "-f_stat", "1"], cwd=root, capture_output=True, text=True, timeout=600,
                       preexec_fn=helpers._big_stack)
    assert r.returncode == 0, r.stderr[-400:]
    mine = str(tmp_path / "mine")
    os.makedirs(mine)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, [case["trk"]], precision=64)
    X.upload()
    X.set_maxstat(True)
    ht1, ht2 = sloshgpu.track_datums(case["trk"])
    y, mo, d, hr, mi = X.start_date(0)
    start = (datetime.date(y, mo, d) - datetime.date(1970, 1, 1)).days * 86400.0
    start = start + hr * 3600.
    start = start + mi * 60
    clock = start
    rex_clock = 0.0
    while rex_clock < start:
        rex_clock += 10 * 60
    R = sloshgpu.Rex(os.path.join(mine, "out.rex"), "FileVersion=2013", rexutil.rex_abbrev(case["basin"]),
                     B.imxb, B.jmxb)
    names = set()

    def save(t6, at):
        R.frame(t6, at, X.rex_frames("HBMAX")[0])
        name = "out%03d.env" % int((at - start) / 3600. + .5)
        sloshgpu.write_env(os.path.join(mine, name), X.maxstat_env()[0], ht1, ht2)
        names.add(name)

    t6 = None
    for n in range(1, X.nsteps[0] + 1):
        X.step(1)
        tr = X.track(0, n)
        clock = clock + tr[2]
        t6 = [tr[3], -tr[4], tr[7], tr[8], tr[5], tr[6]]
        if clock >= rex_clock:
            save(t6, clock)
            rex_clock += 10 * 60
            X.maxstat_reset()
    if rex_clock != clock + 10 * 60:
        save(t6, rex_clock)
    R.env(case["trk"], X.rex_frames("HMX")[0])
    R.close()
    X.close()
    B.close()
    ref_names = sorted(f for f in os.listdir(root) if f.startswith("out") and f.endswith(".env"))
    assert sorted(names) == ref_names and len(ref_names) >= 3
    for f in ref_names:
        a, c = helpers.read_env(os.path.join(mine, f)), helpers.read_env(os.path.join(root, f))
        assert a[:2] == c[:2] and a[4] == c[4], f
        assert np.array_equal(a[3], c[3]), f
    assert open(os.path.join(mine, "out.rex"), "rb").read() == open(os.path.join(root, "out.rex"), "rb").read()


def test_batched_envsave2_windows_equal_reference_cli(tmp_path):
    """-f_stat with -envSave2 60 over a batch of two storms: the first window
    restarts at every 10-minute rex save, the second only when an hourly
    envSave2 save falls due (MaxStatInitGrid flags 1 and 3, slosh2.c:
    2514-2520); every envDir file (ExtendEnvName) and envDir2 file
    (ExtendEnvName2, slosh2.c:2014-2027), the trailing ones included
    (slosh2.c:2597-2668), equals the reference CLI's for each storm."""
    import subprocess
    b, t = synth.tiny(seed=5, ni=30, nj=26, hours=6)
    root = str(tmp_path / "ref")
    case = helpers.make_case(root, b, t)
    rexutil.trk_canonical(t, case["trk"])
    t2 = synth.straight_track(b, 9.0, 300.0, 70.0, 20.0, hours=6, landfall=3)
    trk2 = os.path.join(root, "second.trk")
    open(trk2, "w").write(synth.trk_text(t2))
    rexutil.trk_canonical(t2, trk2)
    trks = [case["trk"], trk2]
    sub = b.name.lower()   # the basin's folder under -envDir (setup.c:241-251)
    for trk in trks:
        for d in ("e1", "e2"):
            os.makedirs(os.path.join(root, d, sub), exist_ok=True)
        r = subprocess.run([helpers.REF_CLI, "-basin", case["basin"], "-rootDir", root, "-trk", trk, "-env", "yes",
                            "-envDir", os.path.join(root, "e1"), "-envDir2", os.path.join(root, "e2"),
                            "-envSave2", "60", "-verbose", "0", "-f_stat", "1"], cwd=root, capture_output=True,
                           text=True, timeout=600, preexec_fn=helpers._big_stack)
        assert r.returncode == 0, r.stderr[-400:]
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, trks, precision=64)
    X.upload()
    X.set_maxstat(True)
    X.set_maxstat2(True)
    assert X.nsteps[0] == X.nsteps[1] and X.start_date(0) == X.start_date(1)
    ht = [sloshgpu.track_datums(p) for p in trks]
    y, mo, d, hr, mi = X.start_date(0)
    start = (datetime.date(y, mo, d) - datetime.date(1970, 1, 1)).days * 86400.0
    start = start + hr * 3600.
    start = start + mi * 60
    clock, rex_clock, env2_clock = start, 0.0, 0.0
    while rex_clock < start:
        rex_clock += 10 * 60
    while env2_clock < start:
        env2_clock += 60 * 60
    mine = {"e1": {}, "e2": {}}

    def save(which, at):
        env = X.maxstat_env() if which == "e1" else X.maxstat2_env()
        for s in range(2):
            name = "%s%03d.env" % (os.path.basename(trks[s])[:-4], int((at - start) / 3600. + .5))
            mine[which][name] = env[s]

    for n in range(1, X.nsteps[0] + 1):
        X.step(1)
        clock = clock + X.track(0, n)[2]
        if clock >= rex_clock:
            save("e1", clock)
            due2 = clock >= env2_clock
            if due2:
                save("e2", clock)
            rex_clock += 10 * 60
            X.maxstat_reset()
            if due2:
                env2_clock += 60 * 60
                X.maxstat2_reset()
    if rex_clock != clock + 10 * 60:
        save("e1", rex_clock)
    if env2_clock != clock + 60 * 60:
        save("e2", env2_clock)
    X.close()
    B.close()
    for which in ("e1", "e2"):
        d = os.path.join(root, which, sub)
        ref_names = sorted(os.listdir(d))
        assert sorted(mine[which]) == ref_names and len(ref_names) >= 10, which
        for f in ref_names:
            s = 0 if f.startswith("storm") else 1
            path = os.path.join(str(tmp_path), which + f)
            sloshgpu.write_env(path, mine[which][f], *ht[s])
            a, c = helpers.read_env(path), helpers.read_env(os.path.join(d, f))
            assert a[:2] == c[:2] and a[4] == c[4], (which, f)
            assert np.array_equal(a[3], c[3]), (which, f)


def test_maxstat_streams_and_graphs(tmp_path):
    """The windowed maxima (both windows) are updated on every launch path:
    the batch in 3 parts on 3 streams and the step loop replayed as hipGraphs
    give the one-stream values, storm by storm."""
    b, t = synth.tiny(seed=74, ni=30, nj=26, hours=8)
    root = str(tmp_path)
    case = helpers.make_case(root, b, t)
    rng = np.random.default_rng(74)
    trks = [case["trk"]]
    for k in range(6):
        p = os.path.join(root, "m%d.trk" % k)
        with open(p, "w") as fh:
            fh.write(synth.trk_text(synth.random_track(b, rng, hours_range=(6, 8))))
        trks.append(p)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    res = []
    for mode in ("one", "parts", "graphs"):
        X = sloshgpu.Batch(B, trks)
        X.upload()
        X.set_path(sloshgpu.PATH_FUSED, 8)
        if mode == "parts":
            X.set_streams(3)
        if mode == "graphs":
            X.set_graph(20)
        X.set_maxstat(True)
        X.set_maxstat2(True)
        X.step(130)
        X.maxstat_reset()
        X.step(170)
        X.sync()
        res.append(([X.get(s, "HBMAX") for s in range(X.n)], X.maxstat2_env()))
        X.close()
    B.close()
    for r in res[1:]:
        for s in range(len(trks)):
            assert np.array_equal(res[0][0][s], r[0][s]), s
        assert np.array_equal(res[0][1], r[1])
