"""HIP time step (libslosh_hip.so, through its C ABI) against the FP32 oracle:
bit-identical HB/UB/VB/HMX after every checkpoint, on the golden cases, on
every smoothing/filter variant of the basin flags, on the three grid types and
on the full polar 100x90 two-day run; and, over the horizon on which the FP32
and FP64 trajectories stay on one branch, against the reference FP64 build's
golden vectors to 0.01 ft."""
import numpy as np
import pytest

import helpers
from conftest import golden_case_files, load_golden, GOLDEN_CASES
from helpers import synth

pytestmark = pytest.mark.gpu


def _assert_identical(out, label):
    for step, fields in sorted(out.items()):
        for f, (g, o) in fields.items():
            bad = np.flatnonzero(g.ravel() != o.ravel())
            assert bad.size == 0, "%s step %d %s: %d cells differ, first %s gpu %r oracle %r" % (
                label, step, f, bad.size, np.unravel_index(bad[0], g.shape), g.ravel()[bad[0]],
                o.ravel()[bad[0]])


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_golden_cases_bitwise(name, tmp_path):
    g = load_golden(name)
    case = golden_case_files(g, str(tmp_path))
    checks = [int(c) for c in g["checks"]]
    print("case", name, "steps", int(g["nsteps"][0]), flush=True)
    out, _ = helpers.gpu_vs_oracle(case, int(g["nsteps"][0]), checks)
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


@pytest.mark.parametrize("idx", range(len(helpers.VARIANTS)), ids=helpers.variant_id)
def test_variants_bitwise(idx, tmp_path):
    case = helpers.variant_case(str(tmp_path), idx)
    print("variant", helpers.variant_id(idx), flush=True)
    out, _ = helpers.gpu_vs_oracle(case, 300, [1, 5, 31, 100, 200])
    _assert_identical(out, helpers.variant_id(idx))


def test_polar_100x90_two_days_bitwise(tmp_path):
    """The reference's own CPU-runnable workload size: polar 100x90 basin,
    60 s step, 48 h (2880 steps) with barriers and drying."""
    b, t = synth.config1()
    case = helpers.make_case(str(tmp_path), b, t)
    print("polar 100x90 2880 steps", flush=True)
    out, events = helpers.gpu_vs_oracle(case, 2880, [1, 100, 720, 1440, 2160])
    _assert_identical(out, "polar100x90")
    assert events > 1000          # the drying cascade is exercised


def test_elliptic_250x300_prefix_bitwise(tmp_path):
    b, t = synth.config2()
    case = helpers.make_case(str(tmp_path), b, t)
    print("elliptic 250x300 300 steps", flush=True)
    out, _ = helpers.gpu_vs_oracle(case, 300, [1, 50, 150])
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
