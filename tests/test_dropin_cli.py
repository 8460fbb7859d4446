"""The reference's own C driver and CLI (cstart.c, slosh2.c, rex.c ...) linked
over the FP64 oracle through the Fortran-ABI glue (oracle/orc_ftn.c) must
write the same env and rex files as the reference FP64 build itself.  This
pins, end to end through the reference's writers, the glue semantics the GPU
drop-in (libslosh_ftn.so) shares: what INITAL returns (date, DELT, ZB, initial
water), HHB = min(HB, HMX) after HMXSV on rex steps, the storm values TMSTPM
returns, and the envelope CLNUP hands over.  Skipped where the reference
build (oracle/_ref) is absent."""
import os

import pytest

import helpers
from helpers import synth

pytestmark = pytest.mark.skipif(not (os.path.exists(helpers.REF_CLI) and os.path.exists(helpers.ORC_CLI["64"])),
                                reason="oracle/_ref not built")


def cli_cases():
    """(id, basin, track) -- grid types, smoothing variants, a deepening storm."""
    b, t = synth.tiny(seed=5)
    yield "polar", b, t
    for idx in (9, 18, 19, 20, 21):   # xfilter, ellip, ellip+filter, hyper, hyper+filter
        g, e, m = helpers.VARIANTS[idx]
        b, t = synth.tiny(seed=100 + idx, gtype=g, ni=28, nj=24, hours=6)
        t.dp = [85.0] * len(t.dp)
        b.extra["ebsn1"] = e
        b.extra["ismooth"] = m
        yield helpers.variant_id(idx), b, t
    b, t = synth.tiny(seed=7, gtype="$", ni=30, nj=24, hours=8)
    import numpy as np
    t.dp = list(np.linspace(40.0, 80.0, len(t.dp)))
    t.rmw = list(np.linspace(12.0, 20.0, len(t.dp)))
    yield "ellip-deepening", b, t
    for idx in (0, 1):                   # sub-grid 1-D flow sides, banks, weirs, cuts, channels
        b, t = helpers.edge_basin(idx)
        yield "sides-" + helpers.edge_id(idx), b, t
    b, t = synth.tiny(seed=205, ni=30, nj=26, hours=6)   # closed island: FLTER2, periodic rows
    b.extra["dollar"] = "$ "
    yield "island-polar", b, t


CASES = list(cli_cases())


@pytest.mark.parametrize("cid,b,t", CASES, ids=[c[0] for c in CASES])
def test_reference_cli_over_oracle64_matches_reference(cid, b, t, tmp_path):
    root = str(tmp_path)
    bsn = synth.write_case(root, b, t)["basin"]
    env_r, rex_r = helpers.run_cli(helpers.REF_CLI, root, bsn, "ref")
    env_o, rex_o = helpers.run_cli(helpers.ORC_CLI["64"], root, bsn, "orc")
    assert helpers.same_cli_outputs(env_r, rex_r, env_o, rex_o) == []


def test_polar_100x90_cli(tmp_path):
    """The reference's CPU-runnable workload (polar 100x90, 48 h) through the CLI."""
    b, t = synth.config1()
    root = str(tmp_path)
    bsn = synth.write_case(root, b, t)["basin"]
    env_r, rex_r = helpers.run_cli(helpers.REF_CLI, root, bsn, "ref")
    env_o, rex_o = helpers.run_cli(helpers.ORC_CLI["64"], root, bsn, "orc")
    assert helpers.same_cli_outputs(env_r, rex_r, env_o, rex_o) == []


def test_list_run_matches_reference(tmp_path):
    """A list run (-lst, DoStormListV0, cstart.c:92-245): two storms on two
    basins in one process, so INITAL .. CLNUP run twice through the same glue;
    every run's env and rex equal the reference CLI's."""
    root = str(tmp_path)
    cases = helpers.list_run_case(root)
    ref = helpers.run_cli_list(helpers.REF_CLI, root, cases, "ref")
    orc = helpers.run_cli_list(helpers.ORC_CLI["64"], root, cases, "orc64")
    for (ea, ra), (eb, rb) in zip(ref, orc):
        assert helpers.same_cli_outputs(ea, ra, eb, rb) == []


@pytest.mark.parametrize("idx", range(len(helpers.TIDE_CASES)), ids=[c[0] for c in helpers.TIDE_CASES])
def test_tide_coupling_over_oracle64_matches_reference(idx, tmp_path):
    """-f_tide: the reference's driver adds the tide between steps to COMMON HB
    and TIDESH (CalcTideGrid2, slosh2.c:813-1027) and calls STATICHT and
    ISSUBGRID; the glue hands those changes to the model before the next step
    and BDRYHT adds TIDESH (bdryht.f:85).  The env and rex files equal the
    reference CLI's byte for byte."""
    root = str(tmp_path)
    case, flag = helpers.tide_case(root, idx)
    env_r, rex_r = helpers.run_cli(helpers.REF_CLI, root, case["basin"], "ref", extra=["-f_tide", flag])
    env_o, rex_o = helpers.run_cli(helpers.ORC_CLI["64"], root, case["basin"], "orc", extra=["-f_tide", flag])
    assert helpers.same_cli_outputs(env_r, rex_r, env_o, rex_o) == []


def test_tide_changes_the_run(tmp_path):
    """The tide cases are sensitive: with the tide the reference's env differs
    from its surge-only run on many squares (a glue that dropped the driver's
    changes would write the surge-only files), and the deep case has squares
    that V2.1 sets to STATICHT + tide."""
    import numpy as np
    root = str(tmp_path)
    case, flag = helpers.tide_case(root, 0)
    env_t, _ = helpers.run_cli(helpers.REF_CLI, root, case["basin"], "tide", extra=["-f_tide", flag])
    env_s, _ = helpers.run_cli(helpers.REF_CLI, root, case["basin"], "surge")
    a, c = helpers.read_env(env_t)[3], helpers.read_env(env_s)[3]
    assert int((a != c).sum()) > 50
    # the deep case forces STATICHT + tide on its deep squares
    root2 = str(tmp_path / "deep")
    case2, flag2 = helpers.tide_case(root2, 2)
    zb = np.array(case2["b"].zb)
    assert (zb >= 290).sum() > 20


def test_full_depth_over_oracle64_matches_reference(tmp_path):
    """-fullDepth: the depths are taken as read instead of capped at 600 ft
    (crdrd2.f:441-444).  The env and rex files equal the reference CLI's, and
    differ from the capped run's (the basin has squares deeper than 600 ft)."""
    import numpy as np
    root = str(tmp_path)
    case = helpers.full_depth_case(root)
    assert (np.array(case["b"].zb) > 600).sum() > 20
    bsn = case["basin"]
    env_r, rex_r = helpers.run_cli(helpers.REF_CLI, root, bsn, "ref", extra=["-fullDepth", "1"])
    env_o, rex_o = helpers.run_cli(helpers.ORC_CLI["64"], root, bsn, "orc", extra=["-fullDepth", "1"])
    assert helpers.same_cli_outputs(env_r, rex_r, env_o, rex_o) == []
    env_c, rex_c = helpers.run_cli(helpers.REF_CLI, root, bsn, "capped")
    assert open(rex_c, "rb").read() != open(rex_r, "rb").read()
