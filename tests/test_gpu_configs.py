"""The named workloads at their full size on the GPU, through the C ABI.

* Drop-in, one storm (elliptic 250x300, 72 h spline track, 8640 steps): the
  FP32 step bit-identical to the FP32 oracle through the whole run, and the
  FP64 drop-in CLI's env and rex files equal to the reference CLI's.
* Batched MEOW ensemble (1024 storms, polar 250x300, 48 h, FP64): storms
  sampled across headings, speeds, deficits, sizes and offsets bit-identical
  to the FP64 oracle at landfall and at the end; members equal to the same
  storm run alone; the MOM equal to the maximum of the envelopes.
* Uneven ensemble (4096 storms over six basins, FP32): one storm per basin
  bit-identical to the FP32 oracle over its whole run; the batch run to
  completion with a member per basin equal to the same storm alone.
* Served stream (Poisson arrivals of 72 h forecast runs over the six basins,
  FP64): sampled runs' envelopes equal to the same run alone.
"""
import os

import numpy as np
import pytest

import helpers
from helpers import synth
import sloshgpu

pytestmark = pytest.mark.gpu


def _identical(a, b, label):
    for f in helpers.FIELDS:
        bad = np.flatnonzero(a[f].ravel() != b[f].ravel())
        assert bad.size == 0, "%s %s: %d cells differ" % (label, f, bad.size)


def _write_tracks(root, tracks, prefix):
    out = []
    for k, t in enumerate(tracks):
        p = os.path.join(root, "%s%04d.trk" % (prefix, k))
        with open(p, "w") as fh:
            fh.write(synth.trk_text(t))
        out.append(p)
    return out


def _alone(B, trk, prec, path=sloshgpu.PATH_FUSED):
    X = sloshgpu.Batch(B, [trk], precision=prec)
    X.set_path(path, 32)
    X.upload()
    X.step(X.nsteps[0])
    X.sync()
    out = {f: X.get(0, f) for f in helpers.FIELDS}
    env, _ = X.envelopes(want_mom=False)
    X.close()
    return out, env[0]


def test_c2_full_run_fp32_bitwise(tmp_path):
    b, t = synth.config2()
    case = helpers.make_case(str(tmp_path), b, t)
    checks = [1, 2000, 4320, 6480, 8640]
    print("C2 elliptic 250x300, 8640 steps, FP32 vs oracle", flush=True)
    out, events = helpers.gpu_vs_oracle(case, 8640, checks, tile_rows=32)
    assert max(out) == 8640
    for step, fields in sorted(out.items()):
        for f, (g, o) in fields.items():
            bad = np.flatnonzero(g.ravel() != o.ravel())
            assert bad.size == 0, "step %d %s: %d cells differ" % (step, f, bad.size)
    assert events > 100


def test_c2_full_run_fp64_cli_matches_reference(tmp_path):
    from concurrent.futures import ThreadPoolExecutor
    b, t = synth.config2()
    root = str(tmp_path)
    bsn = synth.write_case(root, b, t)["basin"]
    print("C2 drop-in CLI FP64 vs reference CLI, 8640 steps", flush=True)
    with ThreadPoolExecutor(2) as ex:
        ref = ex.submit(helpers.run_cli, helpers.REF_CLI, root, bsn, "ref")
        old = os.environ.get("SLG_PRECISION")
        os.environ["SLG_PRECISION"] = "64"
        try:
            env_g, rex_g = helpers.run_cli(helpers.GPU_CLI, root, bsn, "gpu64")
        finally:
            if old is None:
                del os.environ["SLG_PRECISION"]
            else:
                os.environ["SLG_PRECISION"] = old
        env_r, rex_r = ref.result()
    assert helpers.same_cli_outputs(env_g, rex_g, env_r, rex_r) == []


def test_c3_meow_1024_fp64_sampled_bitwise(tmp_path):
    root = str(tmp_path)
    b = synth.config3_basin()
    tracks = synth.meow_tracks(b, n=1024)
    case = synth.write_case(root, b, tracks[0], trk_name="s0000.trk")
    trks = [case["trk"]] + _write_tracks(root, tracks[1:], "s")
    # across headings (0-7), speeds (8s), deficits (32s), sizes (128s), offsets (256s)
    sample = [0, 3 + 8 * 1 + 32 * 2, 6 + 8 * 3 + 32 * 3 + 128, 5 + 8 * 2 + 256 * 2, 1 + 32 * 1 + 128 + 256 * 3,
              7 + 8 * 3 + 32 * 3 + 128 + 256 * 3]
    landfall, end = 2880, 5760
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, trks, precision=64)
    X.upload()
    rows, path, ms = X.autotune(20)
    X.set_streams(3)
    print("C3 1024 storms FP64: tuned %s %s, to landfall" % (rows, path), flush=True)
    X.step(landfall - X.current_step)
    X.sync()
    got = {landfall: {s: {f: X.get(s, f) for f in helpers.FIELDS} for s in sample}}
    print("C3: to the end", flush=True)
    X.step(end - X.current_step)
    X.sync()
    got[end] = {s: {f: X.get(s, f) for f in helpers.FIELDS} for s in sample}
    env, mom = X.envelopes()
    X.close()
    print("C3: oracle for %d storms" % len(sample), flush=True)
    jobs = [(helpers.track_case(case, trks[s]), "64", [landfall, end], os.path.join(root, "o%d" % s))
            for s in sample]
    snaps = helpers.oracle_many(jobs, workers=len(sample))
    for s, sn in zip(sample, snaps):
        for c in (landfall, end):
            _identical(got[c][s], sn[c], "storm %d step %d" % (s, c))
    # members equal the same storm alone; the MOM is the maximum of the envelopes
    for s in sample[:2]:
        alone, env_a = _alone(B, trks[s], 64)
        _identical(got[end][s], alone, "storm %d alone" % s)
        assert np.array_equal(env[s], env_a)
    v = env.astype(np.int32)
    ok = v < 500
    want = np.where(ok.any(axis=0), np.where(ok, v, -10 ** 6).max(axis=0), 999)
    assert np.array_equal(mom.astype(np.int32), want)
    assert (env[:, :-1, :-1] != 999).any()
    B.close()


def test_c4_uneven_4096_fp32(tmp_path):
    root = str(tmp_path)
    basins = synth.config4_basins()
    cells = np.array([(b.imxb - 1) * (b.jmxb - 1) for b in basins], dtype=float)
    share = np.floor(4096 * cells / cells.sum()).astype(int)
    share[-1] += 4096 - share.sum()
    rng = np.random.default_rng(4096)
    batches, jobs, picks = [], [], []
    for b, n in zip(basins, share):
        d = os.path.join(root, b.name)
        tracks = [synth.random_track(b, rng) for _ in range(n)]
        case = synth.write_case(d, b, tracks[0])
        trks = [case["trk"]] + _write_tracks(d, tracks[1:], "t")
        # the storm with the shortest run goes to the oracle
        k = int(np.argmin([t.end - t.begin for t in tracks]))
        picks.append(k)
        jobs.append((helpers.track_case(case, trks[k]), "32", [], os.path.join(d, "orc")))
        B = sloshgpu.Basin(case["dta"], case["bntline"])
        X = sloshgpu.Batch(B, trks)
        X.upload()
        X.autotune(10)
        X.set_graph(20)
        batches.append((B, X, trks))
    print("C4: %d storms over 6 basins, to the end of every run" % share.sum(), flush=True)
    while True:
        live = [X for B, X, _ in batches if X.current_step < max(X.nsteps)]
        if not live:
            break
        for X in live:
            X.step(min(20, max(X.nsteps) - X.current_step))
    print("C4: oracle, one storm per basin", flush=True)
    snaps = helpers.oracle_many(jobs, workers=6)
    for (B, X, trks), k, sn in zip(batches, picks, snaps):
        X.sync()
        got = {f: X.get(k, f) for f in helpers.FIELDS}
        _identical(got, sn["final"], "basin %s storm %d vs oracle" % (B.imxb, k))
        other = (k + 1) % X.n
        alone, _ = _alone(B, trks[other], 32)
        _identical({f: X.get(other, f) for f in helpers.FIELDS}, alone, "basin %s storm %d alone" % (B.imxb, other))
        X.close()
        B.close()


def test_c5_served_stream_fp64_equals_alone(tmp_path):
    root = str(tmp_path)
    basins = synth.config4_basins()
    cells = np.array([(b.imxb - 1) * (b.jmxb - 1) for b in basins], dtype=float)
    rng = np.random.default_rng(505)
    S = sloshgpu.Server(precision=64, hbm_fraction=0.3)
    ids, cases = [], []
    for b in basins:
        d = os.path.join(root, b.name)
        case = synth.write_case(d, b, synth.forecast_track(b, rng, hours=72))
        ids.append(S.add_basin(case["dta"], case["bntline"], max_slots=32, max_steps=72 * 120 + 400,
                               tables_per_slot=160))
        cases.append(case)
    n = 60
    which = rng.choice(len(basins), size=n, p=cells / cells.sum())
    which[:6] = np.arange(6)          # every basin takes part
    runs = []
    for k in range(n):
        p = os.path.join(root, "a%03d.trk" % k)
        with open(p, "w") as fh:
            fh.write(synth.trk_text(synth.forecast_track(basins[which[k]], rng, hours=72)))
        runs.append((p, int(which[k])))
    print("C5: %d forecast runs over 6 basins, arriving over the pumps" % n, flush=True)
    ids_run, pending = [], list(runs)
    while pending or any(S.result(r, ids[w])[0] != 2 for r, w, _ in ids_run):
        for _ in range(min(len(pending), 3)):
            p, w = pending.pop(0)
            ids_run.append((S.submit(ids[w], p), w, p))
        S.pump(64)
    served = {p: S.result(r, ids[w])[1] for r, w, p in ids_run}
    print("C5: tiles %s; runs alone" % S.tiles(), flush=True)
    S.close()
    sample = [runs[k] for k in (0, 1, 2, 3, 4, 5, 17, 33, 59)]
    for p, w in sample:
        B = sloshgpu.Basin(cases[w]["dta"], cases[w]["bntline"])
        _, env = _alone(B, p, 64)
        B.close()
        assert np.array_equal(served[p], env), (p, w)
