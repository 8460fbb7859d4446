"""The product's host side (basin reader, initial water, storm engine; all in
double, libslosh_hip.so, no GPU needed) against what the reference computed
from the same input files: the COMMON state the time step reads, every storm
row STMVAL leaves, and every radial table set FORCE1 produces."""
import numpy as np
import pytest

import helpers
from conftest import golden_case_files
from golden.make_golden import INIT_KEYS, TABS


def _export(g, tmp_path):
    case = golden_case_files(g, str(tmp_path))
    rows, tabs, nsteps = helpers.product_export(case, str(tmp_path))
    init = helpers.read_rec(str(tmp_path / "init.bin"))
    return init, rows, tabs, nsteps


def test_host_init_matches_reference(golden, tmp_path):
    name, g = golden
    init, _, _, _ = _export(g, tmp_path)
    for k in INIT_KEYS:
        assert np.array_equal(init[k], g["init_" + k]), "%s: %s differs" % (name, k)
    for k in TABS:
        mine = init[k].reshape(600, 245)[:, 0]
        assert np.array_equal(mine, g["init_%s_iik1" % k]), "%s: %s differs" % (name, k)


def test_storm_schedule_matches_reference(golden, tmp_path):
    name, g = golden
    _, rows, tabs, nsteps = _export(g, tmp_path)
    ref = g["rows"]
    n = min(len(ref), len(rows))
    assert n == int(g["nsteps"][0]) + 1       # initial row + one per step
    assert np.array_equal(rows[:n], ref[:n]), "%s: storm rows differ" % name
    ref_steps = [int(s) for s in g["tab_steps"]]
    assert [s for s in sorted(tabs) if s < n] == ref_steps
    for s in ref_steps:
        mine = np.concatenate([tabs[s][t] for t in helpers.TABLES])
        assert np.array_equal(mine, g["tab_%d" % s]), "%s: table set at step %d differs" % (name, s)


def test_rejects_bad_inputs(tmp_path):
    import sloshgpu
    with pytest.raises(sloshgpu.SloshError):
        sloshgpu.Basin(str(tmp_path / "missing.dta"), "x")
