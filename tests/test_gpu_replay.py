"""The CONTTY drying-cascade replay's two fallbacks (contty.f:146-213), run on
ordinary cases by lowering the replay's limits from the host
(slg_batch_set_replay_limits):

* the full sweep: once a storm's list of drying squares passes its capacity,
  at the start of a step or while a cascade adds squares, the walk visits every
  square; with capacities of 1 to 64 squares it runs on most drying steps of
  the polar 100x90 two-day run, and the fields stay bit-identical to the
  oracle's serial sweep;
* the visit budget: a run whose walk exhausts it is stopped and flagged, and
  every result call refuses its fields with SLG_ELIMIT.
"""
import numpy as np
import pytest

import helpers
from helpers import synth
import sloshgpu

pytestmark = pytest.mark.gpu


def _assert_identical(out, label):
    for step, fields in sorted(out.items()):
        for f, (g, o) in fields.items():
            bad = np.flatnonzero(g.ravel() != o.ravel())
            assert bad.size == 0, "%s step %d %s: %d cells differ" % (label, step, f, bad.size)


@pytest.mark.parametrize("evcap", [1, 16, 64])
def test_full_sweep_fallback_bitwise(evcap, tmp_path):
    b, t = synth.config1()
    case = helpers.make_case(str(tmp_path), b, t)
    print("full-sweep fallback, evcap", evcap, flush=True)
    out, events = helpers.gpu_vs_oracle(case, 2880, [100, 720, 1440, 2160], tile_rows=32,
                                        setup=lambda X: X.set_replay_limits(evcap, 0))
    _assert_identical(out, "evcap%d" % evcap)
    assert events > 1000


def test_full_sweep_fallback_fp64_bitwise(tmp_path):
    b, t = synth.config1()
    case = helpers.make_case(str(tmp_path), b, t)
    print("full-sweep fallback fp64, evcap 8", flush=True)
    out, _ = helpers.gpu_vs_oracle(case, 2880, [720, 1440, 2160], precision=64,
                                   setup=lambda X: X.set_replay_limits(8, 0))
    _assert_identical(out, "evcap8-fp64")


def test_exhausted_visit_budget_is_refused(tmp_path):
    b, t = synth.config1()
    case = helpers.make_case(str(tmp_path), b, t)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, [case["trk"]])
    X.upload()
    X.set_replay_limits(4096, 1)
    print("visit budget 1, 2880 steps", flush=True)
    X.step(2880)
    X.sync()
    assert X.stats(0)[4] > 0          # cascades were replayed (and cut short)
    for call in (lambda: X.get(0, "HB"), lambda: X.get(0, "HMX"), lambda: X.envelopes(),
                 lambda: X.rex_points([(5, 5)], field="HMX")):
        with pytest.raises(sloshgpu.SloshError, match="ran out of visits"):
            call()
    X.close()
    # the default budget: the same run completes and its fields are handed out
    X = sloshgpu.Batch(B, [case["trk"]])
    X.upload()
    X.step(2880)
    X.get(0, "HB")
    X.envelopes()
    X.close()
    B.close()


@pytest.mark.parametrize("path", [sloshgpu.PATH_FUSED, sloshgpu.PATH_STAGED], ids=["fused", "staged"])
def test_nested_cascade_leaves_hsub(path, tmp_path):
    """Storm 254 of the MEOW ensemble (250x300, FP64) at step 3880: square
    (9, 97) dries with a zero rescale factor, the back-step dries (10, 96) a
    second time, so row 97 is swept again from its start and the revisit of
    (9, 97), its corner transports now zero, stores nothing (contty.f:128-133):
    the square keeps HSUB.  Checked against the oracle just after it."""
    root = str(tmp_path)
    b = synth.config3_basin()
    tracks = synth.meow_tracks(b, n=1024)
    case = synth.write_case(root, b, tracks[0], trk_name="s0000.trk")
    p = root + "/s254.trk"
    with open(p, "w") as fh:
        fh.write(synth.trk_text(tracks[254]))
    c = helpers.track_case(case, p)
    print("storm 254 to step 3900, path", path, flush=True)
    out, events = helpers.gpu_vs_oracle(c, 3900, [3870, 3881, 3900], path=path, tile_rows=32, precision=64)
    _assert_identical(out, "s254")
