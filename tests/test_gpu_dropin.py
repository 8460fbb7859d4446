"""The GPU drop-in end to end: the reference's own C driver and CLI linked over
libslosh_ftn.so (the Fortran-ABI entry points over the HIP time step) writes
the same env and rex files, byte for byte, as the same driver over the FP32
oracle (oracle/orc_ftn.c + liborc32).  Against the reference FP64 build the
FP32 envelope of the 48 h polar run agrees on wet/dry for >= 99 % of cells
(measured 99.5 %); the two precisions take different branches at wetting
fronts (see test_oracle_golden), so no cell-wise tolerance is asserted."""
import os

import numpy as np
import pytest

import helpers
import sloshgpu
from helpers import synth
from test_dropin_cli import CASES

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (os.path.exists(helpers.GPU_CLI) and os.path.exists(helpers.ORC_CLI["32"])),
                                 reason="oracle/_ref drop-in binaries not built")]


def _run_pair(b, t, root):
    bsn = synth.write_case(root, b, t)["basin"]
    print("drop-in", bsn, b.imxb, b.jmxb, flush=True)
    env_g, rex_g = helpers.run_cli(helpers.GPU_CLI, root, bsn, "gpu", env_extra={"SLG_PRECISION": "32"})
    env_o, rex_o = helpers.run_cli(helpers.ORC_CLI["32"], root, bsn, "orc32")
    return bsn, env_g, rex_g, env_o, rex_o


@pytest.mark.parametrize("cid,b,t", CASES, ids=[c[0] for c in CASES])
def test_gpu_cli_matches_oracle32_cli(cid, b, t, tmp_path):
    _, env_g, rex_g, env_o, rex_o = _run_pair(b, t, str(tmp_path))
    assert helpers.same_cli_outputs(env_g, rex_g, env_o, rex_o) == []


def test_gpu_cli_polar_100x90_vs_reference(tmp_path):
    """Polar 100x90, 48 h: identical bytes to the FP32 oracle through the same
    driver, and the envelope's agreement with the reference FP64 CLI.

    FP32 is not the precision that meets the 99.9 % wet/dry bar: it agrees on
    99.53 % of this run's squares, while one ulp of input moves the reference
    build itself by 1-3 squares (99.97-99.99 %, test_sensitivity).  The FP64
    step, which is the reference's default build, equals it bit for bit
    (test_gpu_cli_fp64_polar_100x90_matches_reference).  The bound here only
    guards the FP32 path against regressions."""
    b, t = synth.config1()
    root = str(tmp_path)
    bsn, env_g, rex_g, env_o, rex_o = _run_pair(b, t, root)
    assert helpers.same_cli_outputs(env_g, rex_g, env_o, rex_o) == []
    env_r, _ = helpers.run_cli(helpers.REF_CLI, root, bsn, "ref")
    g, r = helpers.read_env(env_g)[3], helpers.read_env(env_r)[3]
    wet_g, wet_r = g != 999, r != 999
    both = wet_g & wet_r
    d = np.abs(g.astype(int) - r.astype(int))[both]
    print("fp32 vs reference fp64: wet/dry agreement %.4f, within 0.1 ft %.4f, max %.1f ft"
          % ((wet_g == wet_r).mean(), (d <= 1).mean(), d.max() / 10.0))
    assert (wet_g == wet_r).mean() >= 0.99


def _run_gpu64(root, bsn, tag="gpu64", extra=()):
    old = os.environ.get("SLG_PRECISION")
    os.environ["SLG_PRECISION"] = "64"
    try:
        return helpers.run_cli(helpers.GPU_CLI, root, bsn, tag, extra=extra)
    finally:
        if old is None:
            del os.environ["SLG_PRECISION"]
        else:
            os.environ["SLG_PRECISION"] = old


@pytest.mark.parametrize("cid,b,t", CASES, ids=[c[0] for c in CASES])
def test_gpu_cli_fp64_matches_reference_cli(cid, b, t, tmp_path):
    """With the FP64 step the drop-in writes the reference CLI's own env and rex
    bytes."""
    root = str(tmp_path)
    bsn = synth.write_case(root, b, t)["basin"]
    env_g, rex_g = _run_gpu64(root, bsn)
    env_r, rex_r = helpers.run_cli(helpers.REF_CLI, root, bsn, "ref")
    assert helpers.same_cli_outputs(env_g, rex_g, env_r, rex_r) == []


def test_gpu_cli_fp64_polar_100x90_matches_reference_cli(tmp_path):
    b, t = synth.config1()
    root = str(tmp_path)
    bsn = synth.write_case(root, b, t)["basin"]
    env_g, rex_g = _run_gpu64(root, bsn)
    env_r, rex_r = helpers.run_cli(helpers.REF_CLI, root, bsn, "ref")
    assert helpers.same_cli_outputs(env_g, rex_g, env_r, rex_r) == []


def test_gpu_cli_fp64_f_stat_matches_reference_cli(tmp_path):
    """P-Surge windowed maxima (-f_stat 1): the C driver's MaxStatUpdateGrid
    (slosh2.c:1264-1316) reads COMMON HB after every step, so the drop-in keeps
    it current in this mode.  The env file of every window (ExtendEnvName,
    slosh2.c:1998-2011) and the rex file of the window maxima equal the
    reference CLI's."""
    import glob
    import subprocess
    b, t = synth.tiny(seed=5, ni=30, nj=26, hours=6)
    outs = {}
    for tag, exe in (("gpu64", helpers.GPU_CLI), ("ref", helpers.REF_CLI)):
        root = str(tmp_path / tag)
        bsn = synth.write_case(root, b, t)["basin"]
        env = dict(os.environ, SLG_PRECISION="64")
        r = subprocess.run([exe, "-basin", bsn, "-rootDir", root, "-trk", os.path.join(root, "storm.trk"),
                            "-rex", os.path.join(root, "out.rex"), "-env", os.path.join(root, "out.env"),
                            "-verbose", "0", "-f_stat", "1"], cwd=root, capture_output=True, text=True,
                           timeout=600, env=env, preexec_fn=helpers._big_stack)
        assert r.returncode == 0, (tag, r.stdout[-500:], r.stderr[-500:])
        outs[tag] = (root, sorted(os.path.basename(f) for f in glob.glob(os.path.join(root, "*.env"))))
    (rg, fg), (rr, fr) = outs["gpu64"], outs["ref"]
    assert fg == fr and len(fr) >= 3, (fg, fr)
    for f in fr:
        a, c = helpers.read_env(os.path.join(rg, f)), helpers.read_env(os.path.join(rr, f))
        assert a[:3] == c[:3] and a[4] == c[4], f
        assert np.array_equal(a[3], c[3]), f
    ra, rb = open(os.path.join(rg, "out.rex"), "rb").read(), open(os.path.join(rr, "out.rex"), "rb").read()
    ia, ib = ra.rfind(b"envl"), rb.rfind(b"envl")
    assert len(ra) == len(rb) and ia == ib
    if ia >= 0:   # the track records Data_WriteTrk copies are unset for this track layout
        ra, rb = ra[:ia + 4] + ra[ia + 4 + 2400:], rb[:ib + 4] + rb[ib + 4 + 2400:]
    assert ra == rb


def test_gpu_cli_fp64_envsave2_matches_reference_cli(tmp_path):
    """-f_stat with -envSave2 60 (slosh2.c:2477-2520, 2639-2668): the envDir
    and envDir2 files of the drop-in equal the reference CLI's."""
    import subprocess
    b, t = synth.tiny(seed=6, ni=30, nj=26, hours=6)
    outs = {}
    for tag, exe in (("gpu64", helpers.GPU_CLI), ("ref", helpers.REF_CLI)):
        root = str(tmp_path / tag)
        bsn = synth.write_case(root, b, t)["basin"]
        for d in ("e1", "e2"):
            os.makedirs(os.path.join(root, d, b.name.lower()))
        env = dict(os.environ, SLG_PRECISION="64")
        r = subprocess.run([exe, "-basin", bsn, "-rootDir", root, "-trk", os.path.join(root, "storm.trk"),
                            "-env", "yes", "-envDir", os.path.join(root, "e1"), "-envDir2", os.path.join(root, "e2"),
                            "-envSave2", "60", "-verbose", "0", "-f_stat", "1"], cwd=root, capture_output=True,
                           text=True, timeout=600, env=env, preexec_fn=helpers._big_stack)
        assert r.returncode == 0, (tag, r.stdout[-500:], r.stderr[-500:])
        outs[tag] = root
    for d in ("e1", "e2"):
        rg, rr = (os.path.join(outs[k], d, b.name.lower()) for k in ("gpu64", "ref"))
        fg, fr = sorted(os.listdir(rg)), sorted(os.listdir(rr))
        assert fg == fr and len(fr) >= 5, (d, fg, fr)
        for f in fr:
            a, c = helpers.read_env(os.path.join(rg, f)), helpers.read_env(os.path.join(rr, f))
            assert a[:3] == c[:3] and a[4] == c[4], (d, f)
            assert np.array_equal(a[3], c[3]), (d, f)


def test_gpu_list_run_matches_reference(tmp_path):
    """A list run (-lst, DoStormListV0, cstart.c:92-245; two storms on two
    basins in one process, so the library is initialised and cleaned up twice):
    with the FP64 step every run's env and rex equal the reference CLI's; with
    the FP32 step they equal the same driver over the FP32 oracle."""
    root = str(tmp_path)
    cases = helpers.list_run_case(root)
    print("list run fp64", flush=True)
    gpu64 = helpers.run_cli_list(helpers.GPU_CLI, root, cases, "gpu64", env_extra={"SLG_PRECISION": "64"})
    ref = helpers.run_cli_list(helpers.REF_CLI, root, cases, "ref")
    for (ea, ra), (eb, rb) in zip(gpu64, ref):
        assert helpers.same_cli_outputs(ea, ra, eb, rb) == []
    print("list run fp32", flush=True)
    gpu32 = helpers.run_cli_list(helpers.GPU_CLI, root, cases, "gpu32", env_extra={"SLG_PRECISION": "32"})
    orc32 = helpers.run_cli_list(helpers.ORC_CLI["32"], root, cases, "orc32")
    for (ea, ra), (eb, rb) in zip(gpu32, orc32):
        assert helpers.same_cli_outputs(ea, ra, eb, rb) == []


def test_gpu_f_stat_without_hb_mirror_differs(tmp_path):
    """The HB mirror is what the -f_stat windows read: with SLG_HB_MIRROR=0 the
    drop-in skips it, and the window maxima then come from a stale COMMON HB
    (so the default keeps it on for every driver)."""
    import subprocess
    b, t = synth.tiny(seed=5, ni=30, nj=26, hours=6)
    envs = {}
    for tag, mirror in (("on", "1"), ("off", "0")):
        root = str(tmp_path / tag)
        bsn = synth.write_case(root, b, t)["basin"]
        env = dict(os.environ, SLG_PRECISION="64", SLG_HB_MIRROR=mirror)
        r = subprocess.run([helpers.GPU_CLI, "-basin", bsn, "-rootDir", root, "-trk", os.path.join(root, "storm.trk"),
                            "-env", os.path.join(root, "out.env"), "-verbose", "0", "-f_stat", "1"], cwd=root,
                           capture_output=True, text=True, timeout=600, env=env, preexec_fn=helpers._big_stack)
        assert r.returncode == 0, (tag, r.stdout[-500:], r.stderr[-500:])
        envs[tag] = sorted(os.listdir(root))
        envs[tag + "_vals"] = [helpers.read_env(os.path.join(root, f))[3] for f in envs[tag] if f.endswith(".env")
                               and f != "out.env"]
    assert envs["on"] == envs["off"]
    assert any(not np.array_equal(a, c) for a, c in zip(envs["on_vals"], envs["off_vals"]))


@pytest.mark.parametrize("idx", range(len(helpers.TIDE_CASES)), ids=[c[0] for c in helpers.TIDE_CASES])
def test_gpu_cli_tide_coupling(idx, tmp_path):
    """-f_tide through the drop-in: the driver's changes to COMMON HB and
    TIDESH between steps reach the device (slg_batch_put,
    slg_batch_set_boundary_tide), STATICHT and ISSUBGRID answer from the
    product.  FP64: env and rex byte-identical to the reference CLI; FP32:
    byte-identical to the same driver over the FP32 oracle."""
    root = str(tmp_path)
    case, flag = helpers.tide_case(root, idx)
    print("tide", helpers.TIDE_CASES[idx], flush=True)
    bsn = case["basin"]
    env_r, rex_r = helpers.run_cli(helpers.REF_CLI, root, bsn, "ref", extra=["-f_tide", flag])
    env_g, rex_g = helpers.run_cli(helpers.GPU_CLI, root, bsn, "gpu64", extra=["-f_tide", flag],
                                   env_extra={"SLG_PRECISION": "64"})
    assert helpers.same_cli_outputs(env_g, rex_g, env_r, rex_r) == []
    env_o, rex_o = helpers.run_cli(helpers.ORC_CLI["32"], root, bsn, "orc32", extra=["-f_tide", flag])
    env_h, rex_h = helpers.run_cli(helpers.GPU_CLI, root, bsn, "gpu32", extra=["-f_tide", flag],
                                   env_extra={"SLG_PRECISION": "32"})
    assert helpers.same_cli_outputs(env_h, rex_h, env_o, rex_o) == []


def test_gpu_cli_full_depth_matches_reference_cli(tmp_path):
    """-fullDepth through the drop-in (depths deeper than 600 ft kept): FP64 env
    and rex byte-identical to the reference CLI's."""
    root = str(tmp_path)
    case = helpers.full_depth_case(root)
    bsn = case["basin"]
    env_r, rex_r = helpers.run_cli(helpers.REF_CLI, root, bsn, "ref", extra=["-fullDepth", "1"])
    env_g, rex_g = helpers.run_cli(helpers.GPU_CLI, root, bsn, "gpu64", extra=["-fullDepth", "1"],
                                   env_extra={"SLG_PRECISION": "64"})
    assert helpers.same_cli_outputs(env_g, rex_g, env_r, rex_r) == []


@pytest.mark.parametrize("idx", range(len(helpers.TRK1992_CASES)), ids=helpers.trk1992_id)
def test_gpu_cli_and_batch_trk1992(idx, tmp_path):
    """Tracks in the original 1992 format (the reference's default): the drop-in
    CLI's env and rex files equal the reference CLI's byte for byte (FP64), and
    a batched run's env file (its comment the two title lines, its datums from
    the datum line, ReadTrkFile slosh2.c:1940-1990) equals the same env."""
    root = str(tmp_path)
    case = helpers.trk1992_case(root, idx)
    print("trk1992", helpers.TRK1992_CASES[idx], flush=True)
    bsn = case["basin"]
    env_r, rex_r = helpers.run_cli(helpers.REF_CLI, root, bsn, "ref")
    env_g, rex_g = helpers.run_cli(helpers.GPU_CLI, root, bsn, "gpu64", env_extra={"SLG_PRECISION": "64"})
    assert helpers.same_cli_outputs(env_g, rex_g, env_r, rex_r) == []
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, [case["trk"]], precision=64)
    X.upload()
    X.step(X.nsteps[0])
    env, _ = X.envelopes()
    X.close()
    B.close()
    ht1, ht2 = sloshgpu.track_datums(case["trk"])
    mine = os.path.join(root, "batch.env")
    sloshgpu.write_env(mine, env[0], ht1, ht2, comment=sloshgpu.track_env_comment(case["trk"]))
    assert open(mine, "rb").read() == open(env_r, "rb").read()
