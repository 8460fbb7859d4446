"""HIP time step (libslosh_hip.so, through its C ABI) against the FP32 oracle:
bit-identical HB/UB/VB/HMX after every checkpoint, on the golden cases, on
every smoothing/filter variant of the basin flags, on the three grid types and
on the full polar 100x90 two-day run, for both launch paths (fused LDS tiles
of 8/16/32 rows, and one kernel per routine); and, over the horizon on which
the FP32 and FP64 trajectories stay on one branch, against the reference FP64
build's golden vectors to 0.01 ft."""
import numpy as np
import pytest

import helpers
from conftest import golden_case_files, load_golden, GOLDEN_CASES
from helpers import synth
import sloshgpu

pytestmark = pytest.mark.gpu


def _assert_identical(out, label):
    for step, fields in sorted(out.items()):
        for f, (g, o) in fields.items():
            bad = np.flatnonzero(g.ravel() != o.ravel())
            assert bad.size == 0, "%s step %d %s: %d cells differ, first %s gpu %r oracle %r" % (
                label, step, f, bad.size, np.unravel_index(bad[0], g.shape), g.ravel()[bad[0]],
                o.ravel()[bad[0]])


PATHS = [(sloshgpu.PATH_FUSED, 16), (sloshgpu.PATH_FUSED, 8), (sloshgpu.PATH_FUSED, 32),
         (sloshgpu.PATH_STAGED, 16)]
PATH_IDS = ["fused16", "fused8", "fused32", "staged"]


@pytest.mark.parametrize("path,rows", PATHS, ids=PATH_IDS)
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_golden_cases_bitwise(name, path, rows, tmp_path):
    g = load_golden(name)
    case = golden_case_files(g, str(tmp_path))
    checks = [int(c) for c in g["checks"]]
    print("case", name, "steps", int(g["nsteps"][0]), "path", path, rows, flush=True)
    out, _ = helpers.gpu_vs_oracle(case, int(g["nsteps"][0]), checks, path=path, tile_rows=rows)
    _assert_identical(out, name)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_golden_cases_vs_reference_fp64(name, tmp_path):
    """Surge and envelope within 0.01 ft of the reference FP64 build and wet/dry
    agreement >= 99.9 % at every checkpoint up to step 200."""
    g = load_golden(name)
    case = golden_case_files(g, str(tmp_path))
    import sloshgpu
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, [case["trk"]])
    X.upload()
    z = g["init_ZB"].reshape(B.jmxb, B.imxb)[:-1, :-1]
    for c in [int(c) for c in g["checks"] if int(c) <= 200]:
        X.step(c - X.current_step)
        X.sync()
        for f in ("HB", "HMX"):
            a = X.get(0, f)[:-1, :-1]
            b = g["s%d_%s" % (c, f)].reshape(B.jmxb, B.imxb)[:-1, :-1]
            wa, wb = a + z != 0, b + z != 0
            assert (wa == wb).mean() >= 0.999, (name, c, f)
            assert np.abs(a - b)[wa & wb].max() <= 0.01, (name, c, f)
    X.close()
    B.close()


@pytest.mark.parametrize("path", [sloshgpu.PATH_FUSED, sloshgpu.PATH_STAGED], ids=["fused", "staged"])
@pytest.mark.parametrize("idx", range(len(helpers.VARIANTS)), ids=helpers.variant_id)
def test_variants_bitwise(idx, path, tmp_path):
    case = helpers.variant_case(str(tmp_path), idx)
    print("variant", helpers.variant_id(idx), "path", path, flush=True)
    out, _ = helpers.gpu_vs_oracle(case, 300, [1, 5, 31, 100, 200], path=path)
    _assert_identical(out, helpers.variant_id(idx))


@pytest.mark.parametrize("path", [sloshgpu.PATH_FUSED, sloshgpu.PATH_STAGED], ids=["fused", "staged"])
def test_polar_100x90_two_days_bitwise(path, tmp_path):
    """The reference's own CPU-runnable workload size: polar 100x90 basin,
    60 s step, 48 h (2880 steps) with barriers and drying."""
    b, t = synth.config1()
    case = helpers.make_case(str(tmp_path), b, t)
    print("polar 100x90 2880 steps path", path, flush=True)
    out, events = helpers.gpu_vs_oracle(case, 2880, [1, 100, 720, 1440, 2160], path=path)
    _assert_identical(out, "polar100x90")
    assert events > 1000          # the drying cascade is exercised


@pytest.mark.parametrize("rows", [16, 32])
def test_elliptic_250x300_prefix_bitwise(rows, tmp_path):
    b, t = synth.config2()
    case = helpers.make_case(str(tmp_path), b, t)
    print("elliptic 250x300 300 steps rows", rows, flush=True)
    out, _ = helpers.gpu_vs_oracle(case, 300, [1, 50, 150], tile_rows=rows)
    _assert_identical(out, "elliptic250x300")


def test_batch_member_equals_alone(tmp_path):
    """A storm's fields do not depend on what it is batched with."""
    b, t = synth.tiny(seed=31, ni=30, nj=26, hours=6)
    case = helpers.make_case(str(tmp_path), b, t)
    tracks = synth.meow_tracks(b, n=1024, hours=6, landfall=3)[::150]
    paths = []
    for i, tr in enumerate(tracks):
        p = str(tmp_path / ("m%d.trk" % i))
        open(p, "w").write(synth.trk_text(tr))
        paths.append(p)
    import sloshgpu
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    alone = []
    for p in [case["trk"]] + paths:
        X = sloshgpu.Batch(B, [p])
        X.upload()
        X.step(X.nsteps[0])
        X.sync()
        alone.append({f: X.get(0, f) for f in helpers.FIELDS})
        X.close()
    X = sloshgpu.Batch(B, [case["trk"]] + paths)
    X.upload()
    X.step(max(X.nsteps))
    X.sync()
    for s in range(len(alone)):
        for f in helpers.FIELDS:
            assert np.array_equal(X.get(s, f), alone[s][f]), (s, f)
    X.close()
    B.close()


def test_fused_equals_staged_ensemble(tmp_path):
    """Sixteen MEOW storms on the 250x300 polar basin for 400 steps: the fused
    path (LDS tiles, XCD-ordered workgroups) and the one-kernel-per-routine
    path give identical fields for every storm."""
    b = synth.config3_basin()
    tracks = synth.meow_tracks(b, n=1024)[::64]
    case = helpers.make_case(str(tmp_path), b, tracks[0])
    paths = [case["trk"]]
    for k, tr in enumerate(tracks[1:]):
        p = str(tmp_path / ("e%d.trk" % k))
        open(p, "w").write(synth.trk_text(tr))
        paths.append(p)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    res = []
    for path, rows in [(sloshgpu.PATH_STAGED, 16), (sloshgpu.PATH_FUSED, 16), (sloshgpu.PATH_FUSED, 32)]:
        print("ensemble path", path, rows, flush=True)
        X = sloshgpu.Batch(B, paths)
        X.set_path(path, rows)
        X.upload()
        X.step(400)
        X.sync()
        res.append(([{f: X.get(s, f) for f in helpers.FIELDS} for s in range(len(paths))],
                    [X.stats(s) for s in range(len(paths))]))
        X.close()
    B.close()
    for r in res[1:]:
        for s in range(len(paths)):
            for f in helpers.FIELDS:
                assert np.array_equal(r[0][s][f], res[0][0][s][f]), (s, f)
            assert np.array_equal(r[1][s], res[0][1][s]), ("stats", s)


# ------------------------------------------------- sub-grid 1-D flow sides
# Basins with 1-D flow sides, banks, raised weirs and cuts: SMPT2G's side pass
# and CHANNL / FLW1DM_C / FLW1DM_M between its routines, on the staged path and
# on the fused one (SMPT2G, FLT2G and BDRYHT into a third HB buffer, the side
# passes there, the step kernel from HMXSV on).

SIDE_PATHS = [(sloshgpu.PATH_STAGED, 16), (sloshgpu.PATH_FUSED, 16), (sloshgpu.PATH_FUSED, 32)]


@pytest.mark.parametrize("path", SIDE_PATHS, ids=["staged", "fused16", "fused32"])
@pytest.mark.parametrize("idx", range(len(helpers.EDGE_VARIANTS)), ids=helpers.edge_id)
def test_sides_bitwise(idx, path, tmp_path):
    case = helpers.edge_case(str(tmp_path), idx)
    print("sides", helpers.edge_id(idx), path, flush=True)
    out, _ = helpers.gpu_vs_oracle(case, 480, [1, 5, 31, 100, 200, 400], path=path[0], tile_rows=path[1])
    assert all("FLWSQR" in f for f in out.values())
    _assert_identical(out, helpers.edge_id(idx))


@pytest.mark.parametrize("path", SIDE_PATHS[:2], ids=["staged", "fused16"])
@pytest.mark.parametrize("idx", range(len(helpers.EDGE_VARIANTS)), ids=helpers.edge_id)
def test_fp64_sides_bitwise(idx, path, tmp_path):
    """FP64: bit for bit without channels.  CHANNL's fractional powers use the
    device pow, which can differ from the host libm's in the last bit, so with
    channels the FP64 step is held to 1e-6 ft / 1e-6 relative instead."""
    case = helpers.edge_case(str(tmp_path), idx)
    print("fp64 sides", helpers.edge_id(idx), path, flush=True)
    out, _ = helpers.gpu_vs_oracle(case, 480, [1, 31, 200, 400], path=path[0], tile_rows=path[1], precision=64)
    if not helpers.EDGE_VARIANTS[idx][5]:
        _assert_identical(out, helpers.edge_id(idx))
        return
    for step, fields in sorted(out.items()):
        for f, (g, o) in fields.items():
            d = np.abs(g - o)
            rel = d / np.maximum(np.abs(o), 1.0)
            print("step %d %s max abs %.3g max rel %.3g" % (step, f, d.max(), rel.max()))
            assert rel.max() <= 1e-6, (step, f)


# ---------------------------------------------------------------- FP64 step
# The same kernels built in double reproduce the reference's default FP64
# build bit for bit: against the reference's own golden state dumps directly,
# and against the FP64 oracle on every variant.

@pytest.mark.parametrize("path", [sloshgpu.PATH_FUSED, sloshgpu.PATH_STAGED], ids=["fused", "staged"])
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_fp64_golden_equals_reference_dumps(name, path, tmp_path):
    g = load_golden(name)
    case = golden_case_files(g, str(tmp_path))
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, [case["trk"]], precision=64)
    X.set_path(path, 16)
    X.upload()
    print("fp64", name, "path", path, flush=True)
    for c in [int(c) for c in g["checks"]] + [int(g["nsteps"][0])]:
        X.step(c - X.current_step)
        X.sync()
        pre = "final_" if c == int(g["nsteps"][0]) else "s%d_" % c
        for f in helpers.FIELDS:
            ref = g[pre + f].reshape(B.jmxb, B.imxb)
            got = X.get(0, f)
            bad = np.flatnonzero(got.ravel() != ref.ravel())
            assert bad.size == 0, "%s step %d %s: %d cells differ" % (name, c, f, bad.size)
    X.close()
    B.close()


@pytest.mark.parametrize("idx", range(len(helpers.VARIANTS)), ids=helpers.variant_id)
def test_fp64_variants_bitwise(idx, tmp_path):
    case = helpers.variant_case(str(tmp_path), idx)
    print("fp64 variant", helpers.variant_id(idx), flush=True)
    out, _ = helpers.gpu_vs_oracle(case, 300, [1, 5, 31, 100, 200], precision=64)
    _assert_identical(out, helpers.variant_id(idx))


def test_fp64_polar_100x90_two_days_bitwise(tmp_path):
    b, t = synth.config1()
    case = helpers.make_case(str(tmp_path), b, t)
    print("fp64 polar 100x90 2880 steps", flush=True)
    out, events = helpers.gpu_vs_oracle(case, 2880, [1, 720, 1440, 2160], precision=64)
    _assert_identical(out, "polar100x90-fp64")


def test_uneven_run_lengths_equal_alone(tmp_path):
    """Storms of different run lengths in one batch: each stops at its own
    MHALT, and its fields equal those of the same storm run alone."""
    b, t = synth.tiny(seed=41, ni=30, nj=26, hours=8)
    case = helpers.make_case(str(tmp_path), b, t)
    paths = [case["trk"]]
    for k, hours in enumerate((6, 7, 10)):
        tr = synth.straight_track(b, 10.0 + 3 * k, 300.0 + 20 * k, 50.0 + 10 * k, 20.0, hours=hours,
                                  landfall=hours // 2)
        p = str(tmp_path / ("u%d.trk" % k))
        open(p, "w").write(synth.trk_text(tr))
        paths.append(p)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    for path in (sloshgpu.PATH_FUSED, sloshgpu.PATH_STAGED):
        X = sloshgpu.Batch(B, paths)
        X.set_path(path, 16)
        X.upload()
        assert len(set(X.nsteps)) == len(paths)
        X.step(max(X.nsteps))
        X.sync()
        mix = [{f: X.get(s, f) for f in helpers.FIELDS} for s in range(len(paths))]
        X.close()
        for s, p in enumerate(paths):
            A = sloshgpu.Batch(B, [p])
            A.set_path(path, 16)
            A.upload()
            A.step(A.nsteps[0])
            A.sync()
            for f in helpers.FIELDS:
                assert np.array_equal(A.get(0, f), mix[s][f]), (path, s, f)
            A.close()
    B.close()


@pytest.mark.parametrize("precision", [32, 64])
def test_graph_replay_and_autotune_bitwise(precision, tmp_path):
    """The step loop replayed as hipGraphs (storm-own step flags), a run whose
    first steps went to the autotuner's trials, and a tuned run then replayed as
    hipGraphs (the tuner's staged trial filters on step 79 and leaves HB in the
    second buffer, where the graph's fused steps keep it) all give the fields
    of the plain fused loop."""
    b, t = synth.tiny(seed=51, ni=30, nj=26, hours=6)
    case = helpers.make_case(str(tmp_path), b, t)
    tracks = synth.meow_tracks(b, n=1024, hours=6, landfall=3)[::200]
    paths = [case["trk"]]
    for k, tr in enumerate(tracks):
        p = str(tmp_path / ("g%d.trk" % k))
        open(p, "w").write(synth.trk_text(tr))
        paths.append(p)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    res = {}
    for mode in ("plain", "graph", "tuned", "tuned_graph"):
        X = sloshgpu.Batch(B, paths, precision=precision)
        X.upload()
        if mode == "graph":
            X.set_graph(16)
        if mode.startswith("tuned"):
            rows, path, ms = X.autotune(20)
            print("autotune", rows, path, ms, flush=True)
            if mode == "tuned_graph":
                X.set_path(sloshgpu.PATH_FUSED, rows)
                X.set_graph(16)
        X.step(max(X.nsteps) - X.current_step)
        X.sync()
        res[mode] = [{f: X.get(s, f) for f in helpers.FIELDS} for s in range(len(paths))]
        X.close()
    B.close()
    for mode in ("graph", "tuned", "tuned_graph"):
        for s in range(len(paths)):
            for f in helpers.FIELDS:
                assert np.array_equal(res[mode][s][f], res["plain"][s][f]), (mode, s, f)


def test_split_streams_bitwise(tmp_path):
    """The batch stepped in 2, 3 or 4 parts on as many streams gives the fields
    of the one-stream loop, storm by storm (storm counts that do not divide
    evenly, runs of different lengths, two tile heights, a change of the part
    count between calls)."""
    b, t = synth.tiny(seed=61, ni=30, nj=26, hours=8)
    root = str(tmp_path)
    case = helpers.make_case(root, b, t)
    rng = np.random.default_rng(61)
    trks = [case["trk"]]
    for k in range(6):
        tr = synth.random_track(b, rng, hours_range=(6, 8))
        p = "%s/h%d.trk" % (root, k)
        with open(p, "w") as fh:
            fh.write(synth.trk_text(tr))
        trks.append(p)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    res = []
    for streams, rows in ((1, 16), (2, 16), (2, 8), (3, 8), (4, 16)):
        X = sloshgpu.Batch(B, trks)
        X.upload()
        X.set_path(sloshgpu.PATH_FUSED, rows)
        X.set_streams(streams)
        assert X.streams == streams
        X.step(300)
        ms, cnt = X.step_timed(100)
        X.set_streams(5 - streams if streams > 1 else 1)
        X.step(max(X.nsteps))
        X.sync()
        res.append([[X.get(s, f) for f in helpers.FIELDS] for s in range(len(trks))])
        X.close()
    B.close()
    for r in res[1:]:
        for s in range(len(trks)):
            for k, f in enumerate(helpers.FIELDS):
                assert np.array_equal(res[0][s][k], r[s][k]), (s, f)


@pytest.mark.parametrize("precision", [32, 64])
def test_storm_groups_bitwise(precision, tmp_path):
    """The fused kernels' workgroup order in storm groups of 1, 3, 4 or 16
    (groups that do not divide the batch, a group larger than it) gives the
    same fields storm by storm; the autotuner's choice is one of 1, 4, 8, 16."""
    b, t = synth.tiny(seed=71, ni=40, nj=34, hours=8)
    root = str(tmp_path)
    case = helpers.make_case(root, b, t)
    rng = np.random.default_rng(71)
    trks = [case["trk"]]
    for k in range(10):
        p = "%s/q%d.trk" % (root, k)
        with open(p, "w") as fh:
            fh.write(synth.trk_text(synth.random_track(b, rng, hours_range=(8, 9))))
        trks.append(p)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    res = []
    for sg in (1, 3, 4, 16):
        X = sloshgpu.Batch(B, trks, precision=precision)
        X.upload()
        X.set_path(sloshgpu.PATH_FUSED, 16)
        X.set_storm_group(sg)
        assert X.storm_group == sg
        X.step(max(X.nsteps))
        X.sync()
        res.append([[X.get(s, f) for f in helpers.FIELDS] for s in range(len(trks))])
        X.close()
    X = sloshgpu.Batch(B, trks, precision=precision)
    X.upload()
    X.autotune(10)
    assert X.storm_group in (1, 4, 8, 16)
    X.close()
    B.close()
    for r in res[1:]:
        for s in range(len(trks)):
            for k, f in enumerate(helpers.FIELDS):
                assert np.array_equal(res[0][s][k], r[s][k]), (s, f)


# ----------------------------------------------------- closed island basins
# FLTER2, MOMNTM to row JMXB without J side boundaries and the periodic row
# copies (flter2.f, momntm.f:96-112, 464-469), on both launch paths.

@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("path", [sloshgpu.PATH_FUSED, sloshgpu.PATH_STAGED], ids=["fused", "staged"])
@pytest.mark.parametrize("idx", range(len(helpers.ISLAND_VARIANTS)), ids=helpers.island_id)
def test_island_bitwise(idx, path, precision, tmp_path):
    case = helpers.island_case(str(tmp_path), idx)
    print("island", helpers.island_id(idx), "path", path, "fp%d" % precision, flush=True)
    out, _ = helpers.gpu_vs_oracle(case, 300, [1, 19, 20, 40, 100, 200], path=path, precision=precision)
    _assert_identical(out, helpers.island_id(idx))


# '1$' (squares dried at the start) and '2$' (MSY: DEPMSY, INTLHT's MSY
# branch) basins: the initial water made on the device and the steps after it
# equal the oracle from the host's initialisation, bit for bit.
@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("idx", range(len(helpers.DOLLAR_VARIANTS)), ids=helpers.dollar_id)
def test_dollar_bitwise(idx, precision, tmp_path):
    case = helpers.dollar_case(str(tmp_path), idx)
    print("variant", helpers.dollar_id(idx), "fp%d" % precision, flush=True)
    out, _ = helpers.gpu_vs_oracle(case, 300, [1, 20, 100, 300], path=sloshgpu.PATH_FUSED, precision=precision)
    _assert_identical(out, helpers.dollar_id(idx))


def test_island_graph_and_batch_equal_alone(tmp_path):
    """Island storms batched (runs of different lengths, hipGraph replay with
    each storm's own filter steps) equal the same storms run alone."""
    b, t = synth.tiny(seed=204, ni=30, nj=26, hours=9)
    b.extra["dollar"] = "$ "
    root = str(tmp_path)
    case = helpers.make_case(root, b, t)
    rng = np.random.default_rng(204)
    trks = [case["trk"]]
    for k in range(5):
        p = "%s/i%d.trk" % (root, k)
        with open(p, "w") as fh:
            fh.write(synth.trk_text(synth.random_track(b, rng, hours_range=(8, 10))))
        trks.append(p)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, trks)
    X.upload()
    X.set_path(sloshgpu.PATH_FUSED, 16)
    X.set_graph(8)
    X.step(max(X.nsteps))
    X.sync()
    mix = [{f: X.get(s, f) for f in helpers.FIELDS} for s in range(len(trks))]
    X.close()
    for s, p in enumerate(trks):
        A = sloshgpu.Batch(B, [p])
        A.set_path(sloshgpu.PATH_STAGED, 16)
        A.upload()
        A.step(A.nsteps[0])
        A.sync()
        for f in helpers.FIELDS:
            assert np.array_equal(A.get(0, f), mix[s][f]), (s, f)
        A.close()
    B.close()


# ------------------------------------------- Ver2019-03 basins (Manning's n)
@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("idx", range(len(helpers.MANNING_VARIANTS)), ids=helpers.manning_id)
def test_manning_bitwise(idx, precision, tmp_path):
    case = helpers.manning_case(str(tmp_path), idx)
    print("manning", helpers.manning_id(idx), "fp%d" % precision, flush=True)
    out, _ = helpers.gpu_vs_oracle(case, 300, [1, 20, 100, 200], precision=precision)
    _assert_identical(out, helpers.manning_id(idx))


@pytest.mark.parametrize("idx", [1, 3, 5], ids=lambda i: helpers.edge_id(i))
def test_sides_graph_and_batch_equal_alone(idx, tmp_path):
    """Side basins batched on the fused path with runs of different lengths,
    replayed as hipGraphs (each storm's own SMPT2G and FILTER steps), and split
    over streams: every storm's fields and 1-D transports equal the same storm
    run alone on the staged path."""
    b, t = helpers.edge_basin(idx, hours=9)
    root = str(tmp_path)
    case = helpers.make_case(root, b, t)
    rng = np.random.default_rng(300 + idx)
    trks = [case["trk"]]
    for k in range(5):
        p = "%s/e%d.trk" % (root, k)
        with open(p, "w") as fh:
            fh.write(synth.trk_text(synth.random_track(b, rng, hours_range=(8, 10))))
        trks.append(p)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    runs = []
    for graph, streams in ((8, 1), (0, 2)):
        X = sloshgpu.Batch(B, trks)
        X.upload()
        X.set_path(sloshgpu.PATH_FUSED, 16)
        if streams > 1:
            X.set_streams(streams)
        X.set_graph(graph)
        X.step(max(X.nsteps))
        X.sync()
        runs.append(([{f: X.get(s, f) for f in helpers.FIELDS} for s in range(len(trks))],
                     [X.flwsqr(s) for s in range(len(trks))]))
        X.close()
    for s, p in enumerate(trks):
        A = sloshgpu.Batch(B, [p])
        A.set_path(sloshgpu.PATH_STAGED, 16)
        A.upload()
        A.step(A.nsteps[0])
        A.sync()
        for fields, flw in runs:
            for f in helpers.FIELDS:
                assert np.array_equal(A.get(0, f), fields[s][f]), (s, f)
            assert np.array_equal(A.flwsqr(0), flw[s]), (s, "FLWSQR")
        A.close()
    B.close()
