"""Batched products and the serving loop on the GPU.

- The device envelope of a run is the env file the reference's own writer
  produces for it: the product's env file equals (values, datums, header) the
  one the drop-in CLI writes through EnvSave for the same storm.
- The device MOM is the maximum of the MEOWs by the rule of
  gui/tclsrc/meowgen.tcl:34-110, recomputed here from the per-run envelopes.
- Serving: a run's envelope is the same when it is served alone, in a mix of
  runs on two basins admitted at different times, and through the batch API."""
import os

import numpy as np
import pytest

import helpers
import sloshgpu
from helpers import synth

pytestmark = pytest.mark.gpu


def _tracks(b, root, n, hours=6, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        tr = synth.straight_track(b, float(rng.uniform(6, 20)), float(rng.uniform(0, 360)),
                                  float(rng.uniform(30, 90)), float(rng.uniform(12, 30)), hours=hours,
                                  landfall=hours // 2, offset_mi=float(rng.uniform(-10, 10)))
        p = os.path.join(root, "r%03d_%d.trk" % (k, seed))
        open(p, "w").write(synth.trk_text(tr))
        out.append(p)
    return out


@pytest.mark.skipif(not os.path.exists(helpers.GPU_CLI), reason="drop-in CLI not built")
def test_env_file_equals_dropin_cli(tmp_path):
    b, t = synth.tiny(seed=5)
    root = str(tmp_path)
    case = synth.write_case(root, b, t)
    env_cli, _ = helpers.run_cli(helpers.GPU_CLI, root, case["basin"], "cli", env_extra={"SLG_PRECISION": "32"})
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, [case["trk"]])
    X.upload()
    X.step(X.nsteps[0])
    env, mom = X.envelopes()
    ht1, ht2 = sloshgpu.track_datums(case["trk"])
    mine = os.path.join(root, "mine.env")
    sloshgpu.write_env(mine, env[0], ht1, ht2)
    X.close()
    B.close()
    a, c = helpers.read_env(mine), helpers.read_env(env_cli)
    assert a[:3] == c[:3]
    assert np.array_equal(a[3], c[3])
    assert a[4] == c[4]


def test_mom_is_max_of_meows(tmp_path):
    b = synth.tiny(seed=11, ni=32, nj=28, hours=6)[0]
    root = str(tmp_path)
    trks = _tracks(b, root, 12, seed=1)
    case = synth.write_case(root, b, synth.tiny(seed=11, ni=32, nj=28, hours=6)[1])
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, trks)
    X.upload()
    X.step(max(X.nsteps))
    env, mom = X.envelopes()
    X.close()
    B.close()
    v = env.astype(np.int32)
    ok = v < 500                       # meowgen.tcl: value/10 < 50 ft counts
    best = np.where(ok, v, -10 ** 6).max(axis=0)
    want = np.where(ok.any(axis=0), best, 999)
    assert np.array_equal(mom.astype(np.int32), want)
    assert (env == 999).any() and (env != 999).any()


@pytest.mark.parametrize("precision", [32, 64])
def test_served_run_same_alone_and_in_mix(precision, tmp_path):
    root = str(tmp_path)
    b1, t1 = synth.tiny(seed=21, ni=30, nj=26, hours=6)
    b2, t2 = synth.tiny(seed=22, gtype="$", ni=26, nj=22, hours=6)
    b2.name = "tnz"
    c1 = synth.write_case(os.path.join(root, "b1"), b1, t1)
    c2 = synth.write_case(os.path.join(root, "b2"), b2, t2)
    trk1 = _tracks(b1, root, 9, seed=3)
    trk2 = _tracks(b2, root, 5, seed=4)
    probe = trk1[4]
    print("serve precision", precision, flush=True)

    # alone
    S = sloshgpu.Server(precision=precision)
    bid = S.add_basin(c1["dta"], c1["bntline"], max_slots=4, max_steps=800)
    rid = S.submit(bid, probe)
    while S.result(rid, bid)[0] != 2:
        S.pump(50)
    alone = S.result(rid, bid)[1]
    S.close()

    # in a mix: two basins, 4 slots each (fewer than runs), arrivals spread over
    # the pumps so runs join pools that are already stepping
    S = sloshgpu.Server(precision=precision)
    b_1 = S.add_basin(c1["dta"], c1["bntline"], max_slots=4, max_steps=800)
    b_2 = S.add_basin(c2["dta"], c2["bntline"], max_slots=4, max_steps=800)
    ids = []
    arrivals = [(b_1, p) for p in trk1] + [(b_2, p) for p in trk2]
    rng = np.random.default_rng(7)
    rng.shuffle(arrivals)
    pending = list(arrivals)
    while pending or any(S.result(r, bb)[0] != 2 for r, bb, _ in ids):
        for _ in range(min(len(pending), 2)):
            bb, p = pending.pop()
            ids.append((S.submit(bb, p), bb, p))
        S.pump(37)
    st = S.stats()
    assert st["done"] == len(arrivals) and st["queued"] == 0
    mix = {p: S.result(r, bb)[1] for r, bb, p in ids}
    S.close()
    assert np.array_equal(mix[probe], alone)

    # and the batch API on the same run
    B = sloshgpu.Basin(c1["dta"], c1["bntline"])
    X = sloshgpu.Batch(B, [probe], precision=precision)
    X.upload()
    X.step(X.nsteps[0])
    env, _ = X.envelopes(want_mom=False)
    X.close()
    B.close()
    assert np.array_equal(env[0], alone)


@pytest.mark.parametrize("precision", [32, 64])
def test_served_side_basin_runs_equal_staged_alone(precision, tmp_path):
    """A basin with 1-D flow sides and channels is served beside an ordinary
    basin (slots fewer than runs, arrivals spread over the pumps): every
    served envelope equals the same storm run alone on the staged path."""
    root = str(tmp_path)
    bs, ts = helpers.edge_basin(5, hours=6)     # sides with NOF1D '+' and a channel
    bs.name = "tsd"
    be, te = helpers.edge_basin(3, hours=6)     # sides, banks, weirs, cuts, 3 channels
    be.name = "tse"
    cs = synth.write_case(os.path.join(root, "bs"), bs, ts)
    ce = synth.write_case(os.path.join(root, "be"), be, te)
    trs = _tracks(bs, root, 5, seed=11)
    tre = _tracks(be, root, 6, seed=12)
    print("served sides precision", precision, flush=True)
    S = sloshgpu.Server(precision=precision)
    b_s = S.add_basin(cs["dta"], cs["bntline"], max_slots=3, max_steps=800)
    b_e = S.add_basin(ce["dta"], ce["bntline"], max_slots=3, max_steps=800)
    arrivals = [(b_s, p, cs) for p in trs] + [(b_e, p, ce) for p in tre]
    rng = np.random.default_rng(13)
    rng.shuffle(arrivals)
    pending, ids = list(arrivals), []
    while pending or any(S.result(r, bb)[0] != 2 for r, bb, _, _ in ids):
        for _ in range(min(len(pending), 2)):
            bb, p, c = pending.pop()
            ids.append((S.submit(bb, p), bb, p, c))
        S.pump(29)
    served = [(S.result(r, bb)[1], p, c) for r, bb, p, c in ids]
    S.close()
    for env, p, c in served:
        B = sloshgpu.Basin(c["dta"], c["bntline"])
        X = sloshgpu.Batch(B, [p], precision=precision)
        X.set_path(sloshgpu.PATH_STAGED, 16)
        X.upload()
        X.step(X.nsteps[0])
        alone, _ = X.envelopes(want_mom=False)
        X.close()
        B.close()
        assert np.array_equal(env, alone[0]), p


def test_server_hardest_first_admission(tmp_path):
    """With the hardest-first order and one slot, queued runs are admitted by
    their work, largest first (the reference's job sort), and each run's
    envelope is the one the default submission order gives."""
    import time
    root = str(tmp_path)
    b, t = synth.tiny(seed=31, ni=30, nj=26, hours=6)
    case = synth.write_case(os.path.join(root, "b"), b, t)
    hours = [6, 12, 9, 7]
    paths = []
    for k, h in enumerate(hours):
        tr = synth.straight_track(b, 10.0 + k, 300.0 + 10 * k, 60.0, 20.0, hours=h, landfall=h // 2)
        p = os.path.join(root, "h%02d.trk" % h)
        open(p, "w").write(synth.trk_text(tr))
        paths.append(p)
    envs = []
    for hardest in (False, True):
        S = sloshgpu.Server(precision=64)
        bid = S.add_basin(case["dta"], case["bntline"], max_slots=1, max_steps=1500)
        S.set_order(hardest)
        ids = [S.submit(bid, p) for p in paths]
        time.sleep(1.0)   # every schedule prepared before the first admission
        done = []
        while len(done) < len(ids):
            S.pump(40)
            for r in ids:
                if r not in done and S.result(r, bid)[0] == 2:
                    done.append(r)
        envs.append([S.result(r, bid)[1] for r in ids])
        S.close()
        got = [hours[ids.index(r)] for r in done]
        print("hardest" if hardest else "fifo", "finish order", got, flush=True)
        assert got == (sorted(hours, reverse=True) if hardest else hours)
    for a, c in zip(envs[0], envs[1]):
        assert np.array_equal(a, c)
