"""The gridded tide inside the batch (slg_batch_set_tide): no driver between
steps; the device adds the tide after every step as the reference's -f_tide V2
coupling does (RunInit, slosh2.c:1789-1869; CalcTideGrid2, :813-1027;
RunLoopStep, :1553-1582).

- FP64: a batched run's env file is byte-identical to the reference CLI's with
  the same -f_tide (the reference's default build), for V2 on polar and
  hyperbolic grids, V2.1 with squares deeper than 290 ft (STATICHT + tide),
  V2.2 / V2.4 on a basin with narrowed 1-D sides (ISSUBGRID) and the VDEF
  basin's flavour, with and without a datum-adjustment file.
- FP32: byte-identical to the reference's driver over the FP32 oracle (the
  driver's COMMON copy of HB is double there too).
- One batch of runs with different start dates, years and lengths (several
  clock groups and year tables, runs that end before others) gives every run
  the env the reference CLI gives it alone."""
import os

import numpy as np
import pytest

import helpers
import sloshgpu
from helpers import synth

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (os.path.exists(helpers.REF_CLI) and os.path.exists(helpers.ORC_CLI["32"])),
                                 reason="reference CLI builds not present")]


def _tide_files(case):
    d = case["tidedir"]
    adj = os.path.join(d, case["basin"] + ".adj")
    return os.path.join(d, "ft03.dta"), os.path.join(d, case["basin"] + ".bhc"), adj if os.path.exists(adj) else None


def _batched_envs(case, trks, flag, precision, root, tag, sched_host=False, streams=1):
    mode, thresh = sloshgpu.parse_tide(flag)
    ft03, bhc, adj = _tide_files(case)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, trks, precision=precision, sched_host=sched_host)
    X.set_tide(ft03, bhc, adj, mode, thresh)
    X.upload()
    if streams > 1:   # some steps on one stream, the rest in parts on several
        X.step(100)
        X.set_streams(streams)
    X.step(max(X.nsteps))
    env, _ = X.envelopes()
    X.close()
    B.close()
    out = []
    for k, trk in enumerate(trks):
        ht1, ht2 = sloshgpu.track_datums(trk, mode)
        p = os.path.join(root, "%s_%d.env" % (tag, k))
        sloshgpu.write_env(p, env[k], ht1, ht2)
        out.append(p)
    return out


def _same_env(a_path, b_path):
    a, b = helpers.read_env(a_path), helpers.read_env(b_path)
    assert a[:3] == b[:3]
    bad = np.argwhere(a[3] != b[3])
    assert bad.size == 0, "%d squares differ, first (j, i) %s: %d vs %d" % (
        len(bad), tuple(bad[0]), a[3][tuple(bad[0])], b[3][tuple(bad[0])])
    assert a[4] == b[4]


@pytest.mark.parametrize("idx", range(len(helpers.TIDE_CASES)), ids=[c[0] for c in helpers.TIDE_CASES])
def test_batch_tide_env_equals_reference_cli(idx, tmp_path):
    root = str(tmp_path)
    case, flag = helpers.tide_case(root, idx, hours=12, adj=idx % 2 == 0)
    print("batch tide", helpers.TIDE_CASES[idx], "adj", idx % 2 == 0, flush=True)
    bsn = case["basin"]
    mflag = "V2.2.-3" if flag == "VDEF" else flag    # the flavour file's line
    env_r, _ = helpers.run_cli(helpers.REF_CLI, root, bsn, "ref", extra=["-f_tide", flag])
    (env_g,) = _batched_envs(case, [case["trk"]], mflag, 64, root, "b64")
    _same_env(env_g, env_r)
    env_o, _ = helpers.run_cli(helpers.ORC_CLI["32"], root, bsn, "orc32", extra=["-f_tide", flag])
    (env_h,) = _batched_envs(case, [case["trk"]], mflag, 32, root, "b32", sched_host=idx % 2 == 1)
    _same_env(env_h, env_o)


# start dates (year, month, day, hour, minute) and lengths (hours) of the runs
# of one batch: two runs share a clock, one starts on a leap day, one runs
# into the next year
_RUNS = [((2025, 9, 1, 0, 0), 6), ((2025, 9, 1, 0, 0), 6), ((2024, 2, 29, 13, 30), 8), ((2031, 12, 31, 22, 45), 10)]


@pytest.mark.parametrize("precision,streams", [(64, 1), (32, 1), (64, 3)])
def test_batch_tide_runs_of_several_clocks(precision, streams, tmp_path):
    root = str(tmp_path)
    print("batch tide, several clocks, FP%d, %d streams" % (precision, streams), flush=True)
    b, _ = helpers.tide_basin("polar")
    rng = np.random.default_rng(7)
    cases = []
    for k, (date, hours) in enumerate(_RUNS):
        t = synth.straight_track(b, float(rng.uniform(8, 18)), float(rng.uniform(280, 360)), float(rng.uniform(40, 80)),
                                 float(rng.uniform(15, 30)), hours=hours, landfall=hours // 2,
                                 offset_mi=float(rng.uniform(-8, 8)))
        t.date = date
        r = os.path.join(root, "run%d" % k)
        case = helpers.make_case(r, b, t)
        case["tidedir"] = synth.write_tide_db(r, b, case["basin"], seed=50, adj=True)
        cases.append(case)
    envs = _batched_envs(cases[0], [c["trk"] for c in cases], "V2", precision, root, "multi%d" % precision,
                         streams=streams)
    cli = helpers.REF_CLI if precision == 64 else helpers.ORC_CLI["32"]
    for k, case in enumerate(cases):
        r = os.path.join(root, "run%d" % k)
        env_r, _ = helpers.run_cli(cli, r, case["basin"], "cli", extra=["-f_tide", "V2"])
        _same_env(envs[k], env_r)


def test_batch_tide_across_a_delt_switch(tmp_path):
    """A basin whose time step halves mid-run (DLTIN): the model clock moves by
    each step's DELT, as TMSTPM hands it to the driver."""
    root = str(tmp_path)
    print("batch tide, DELT switch", flush=True)
    bd = synth.make_basin("dlt", " ", 60, 50, seed=6, dt=60.0)
    bd.dltin = (60.0, 30.0, 30.0)
    t = synth.straight_track(bd, hours=24, landfall=12)
    case = helpers.make_case(root, bd, t)
    case["tidedir"] = synth.write_tide_db(root, bd, case["basin"], seed=57, adj=True)
    env_r, _ = helpers.run_cli(helpers.REF_CLI, root, case["basin"], "ref", extra=["-f_tide", "V2.1.-5"])
    (env_g,) = _batched_envs(case, [case["trk"]], "V2.1.-5", 64, root, "b64")
    _same_env(env_g, env_r)


def test_batch_tide_hsf_six_minute_cadence(tmp_path):
    """The hsf basins call the tide only every 6 model minutes (RunLoopStep's
    tidetime and PerformRun's tideClock, slosh2.c:1565-1577, 2583-2586): the
    old tide is the previous call's and TIDESH holds between calls."""
    root = str(tmp_path)
    print("batch tide, hsf cadence", flush=True)
    # a hyperbolic basin 'sfa' goes by 'hsfa' (setup.c:41-64)
    b = synth.make_basin("sfa", "+", 30, 26, seed=45, land_frac=0.25, n_barriers=12, dt=40.0,
                         size_km=(1.0, 3.0), corr=4.0)
    t = synth.straight_track(b, 12.0, 330.0, 60.0, 15.0, hours=12, landfall=6)
    case = helpers.make_case(root, b, t)
    case["tidedir"] = synth.write_tide_db(root, b, case["basin"], seed=58, adj=True)
    env_r, _ = helpers.run_cli(helpers.REF_CLI, root, case["basin"], "ref", extra=["-f_tide", "V2"])
    (env_g,) = _batched_envs(case, [case["trk"]], "V2", 64, root, "b64")
    _same_env(env_g, env_r)
    env_o, _ = helpers.run_cli(helpers.ORC_CLI["32"], root, case["basin"], "orc32", extra=["-f_tide", "V2"])
    (env_h,) = _batched_envs(case, [case["trk"]], "V2", 32, root, "b32")
    _same_env(env_h, env_o)
