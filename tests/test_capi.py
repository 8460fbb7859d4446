"""The C-ABI library loads without a GPU and exports every entry point its
public header declares; host-only entry points behave as documented."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import helpers
import sloshgpu

HEADERS = glob.glob(os.path.join(helpers.REPO, "include", "*.h"))


def declared():
    names = set()
    for h in HEADERS:
        txt = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:int|void|double|const char \*)\s+\**(sl\w+)\s*\(", txt, re.M))
    return sorted(names)


def test_header_declares_entry_points():
    assert len(declared()) >= 16


@pytest.mark.parametrize("sym", declared())
def test_library_exports(sym):
    L = sloshgpu.lib()
    assert hasattr(L, sym), "%s declared but not exported" % sym


def test_basin_open_reports_errors(tmp_path):
    p = tmp_path / "bad.dta"
    p.write_text("not a basin\n")
    with pytest.raises(sloshgpu.SloshError):
        sloshgpu.Basin(str(p), "x")


def test_unsupported_track_format_rejected(tmp_path):
    b, t = helpers.synth.tiny(seed=3)
    case = helpers.make_case(str(tmp_path), b, t)
    bad = tmp_path / "old.trk"
    bad.write_text("this is not a 2013-format track\n")
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    with pytest.raises(sloshgpu.SloshError):
        sloshgpu.Batch(B, [str(bad)])
    B.close()


def test_host_state_without_gpu(tmp_path):
    """Opening a basin and a batch, and exporting its host-side state, needs no
    GPU; the step itself does and says so."""
    b, t = helpers.synth.tiny(seed=4)
    case = helpers.make_case(str(tmp_path), b, t)
    rows, tabs, nsteps = helpers.product_export(case, str(tmp_path))
    assert nsteps == rows.shape[0] - 1 and len(tabs) >= 1


def test_sides_basin_takes_every_path(tmp_path):
    """A basin with sub-grid 1-D flow sides opens with its sides counted and
    accepts both launch paths and hipGraphs (no GPU needed)."""
    import helpers
    case = helpers.edge_case(str(tmp_path), 0)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    assert B.nsides >= 15
    X = sloshgpu.Batch(B, [case["trk"]])
    X.set_path(sloshgpu.PATH_FUSED, 16)
    X.set_graph(16)
    X.set_path(sloshgpu.PATH_STAGED, 16)
    X.close()
    B.close()


def test_batch_size_limit(tmp_path):
    """A batch holds 1 .. SLG_MAX_STORMS (65 535) runs: storms index a launch
    grid dimension.  More is refused at open with a message, before any track
    is read."""
    b, t = helpers.synth.tiny(seed=4)
    case = helpers.make_case(str(tmp_path), b, t)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    with pytest.raises(sloshgpu.SloshError, match="65535"):
        sloshgpu.Batch(B, [case["trk"]] * 65536)
    B.close()


def test_latitude_outside_coriolis_table_refused(tmp_path):
    """MOMNTM takes its Coriolis factor from CORL(INT(YLT)+1..+2)
    (momntm.f:294-295), defined for latitudes -1 < YLT < 89 only; beyond that
    the reference reads outside the table.  A basin whose momentum points
    reach such latitudes is refused at upload, before any device work, and the
    oracle refuses the step (-2) instead of reading out of bounds."""
    root = str(tmp_path)
    # a hyperbolic grid 350 cells long with small axes: cells grow to tens of
    # miles and the grid runs south of the equator
    b = helpers.synth.make_basin("hyb", "+", 350, 200, seed=35, land_frac=0.25, n_barriers=50, dt=30.0)
    rng = np.random.default_rng(3)
    case = helpers.synth.write_case(root, b, helpers.synth.random_track(b, rng, hours_range=(8, 9)))
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, [case["trk"]])
    with pytest.raises(sloshgpu.SloshError, match="outside the Coriolis table"):
        X.upload()
    rows, tabs, _ = helpers.product_export(case, root)
    with pytest.raises(RuntimeError, match="oracle step 1 returned -2"):
        helpers.run_oracle("32", root + "/init.bin", rows, tabs, 1)
    X.close()
    B.close()
