"""Pins the oracle and the product's host side on every basin-flag variant
and grid type by running the reference FP64 build itself (oracle/_ref, built
from the reference sources by oracle/Makefile).  Skipped where that build is
absent."""
import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.skipif(not helpers.have_reference(), reason="oracle/_ref not built")

CHECKS = [1, 5, 31, 100, 200, 300]


@pytest.mark.parametrize("idx", range(len(helpers.VARIANTS)), ids=helpers.variant_id)
def test_variant_oracle64_and_host_match_reference(idx, tmp_path):
    root = str(tmp_path)
    case = helpers.variant_case(root, idx)
    ref = helpers.run_reference(case, root + "/ref", 300, ",".join(map(str, CHECKS)))
    rows, tabs, _ = helpers.product_export(case, root)
    ri = helpers.read_rec(ref + "/init.bin")
    mi = helpers.read_rec(root + "/init.bin")
    assert len(mi) >= 50
    for k in mi:
        assert np.array_equal(ri[k], mi[k]), "init record %s differs" % k
    rrows = np.fromfile(ref + "/storm.bin").reshape(-1, 90)
    assert np.array_equal(rows[:len(rrows)], rrows)
    snaps = helpers.run_oracle("64", root + "/init.bin", rows, tabs, 300, set(CHECKS))
    for c in CHECKS:
        st = helpers.read_rec(ref + "/s%06d.bin" % c)
        for f in helpers.FIELDS:
            assert np.array_equal(snaps[c][f].ravel(), st[f]), "step %d %s differs" % (c, f)


@pytest.mark.parametrize("idx", range(len(helpers.ISLAND_VARIANTS)), ids=helpers.island_id)
def test_island_oracle64_and_host_match_reference(idx, tmp_path):
    """Closed island basins: the product's initialisation and the FP64 oracle
    (FLTER2, MOMNTM's island rows and periodic copy) equal the reference build
    at every checkpoint, filter steps included."""
    root = str(tmp_path)
    case = helpers.island_case(root, idx)
    ref = helpers.run_reference(case, root + "/ref", 300, ",".join(map(str, CHECKS)))
    rows, tabs, _ = helpers.product_export(case, root)
    ri = helpers.read_rec(ref + "/init.bin")
    mi = helpers.read_rec(root + "/init.bin")
    for k in mi:
        assert np.array_equal(ri[k], mi[k]), "init record %s differs" % k
    snaps = helpers.run_oracle("64", root + "/init.bin", rows, tabs, 300, set(CHECKS))
    for c in CHECKS:
        st = helpers.read_rec(ref + "/s%06d.bin" % c)
        for f in helpers.FIELDS:
            assert np.array_equal(snaps[c][f].ravel(), st[f]), "step %d %s differs" % (c, f)


@pytest.mark.parametrize("idx", range(len(helpers.MANNING_VARIANTS)), ids=helpers.manning_id)
def test_manning_oracle64_and_host_match_reference(idx, tmp_path):
    """Ver2019-03 basins: the Manning's n map read by the product's host side
    sets the same initial records (bottom stress classes included) as the
    reference build, and the FP64 oracle then equals it at every checkpoint."""
    root = str(tmp_path)
    case = helpers.manning_case(root, idx)
    ref = helpers.run_reference(case, root + "/ref", 300, ",".join(map(str, CHECKS)))
    rows, tabs, _ = helpers.product_export(case, root)
    ri = helpers.read_rec(ref + "/init.bin")
    mi = helpers.read_rec(root + "/init.bin")
    for k in mi:
        assert np.array_equal(ri[k], mi[k]), "init record %s differs" % k
    snaps = helpers.run_oracle("64", root + "/init.bin", rows, tabs, 300, set(CHECKS))
    for c in CHECKS:
        st = helpers.read_rec(ref + "/s%06d.bin" % c)
        for f in helpers.FIELDS:
            assert np.array_equal(snaps[c][f].ravel(), st[f]), "step %d %s differs" % (c, f)


def test_manning_map_changes_the_run(tmp_path):
    """The same basin read as Ver2014-08 (no map: every class from the fixed
    default) and as Ver2019-03 with a map gives different initial records: the
    Manning pinning above covers a map that matters."""
    root = str(tmp_path)
    a = helpers.manning_case(root + "/a", 0)
    g, seed = helpers.MANNING_VARIANTS[0]
    b, t = helpers.synth.tiny(seed=seed, gtype=g, ni=30, nj=26, hours=6)
    t.dp = [85.0] * len(t.dp)
    c = helpers.make_case(root + "/b", b, t)
    helpers.product_export(a, root + "/a")
    helpers.product_export(c, root + "/b")
    ra = helpers.read_rec(root + "/a/init.bin")
    rb = helpers.read_rec(root + "/b/init.bin")
    assert any(not np.array_equal(ra[k], rb[k]) for k in ra)


def test_island_planted_mistake_is_caught(tmp_path):
    """FLTER2 run as the plain FILTER (no periodic neighbour, no row copy) on an
    island basin must change the run: the island tests see the periodic rules."""
    root = str(tmp_path)
    case = helpers.island_case(root, 0)
    rows, tabs, _ = helpers.product_export(case, root)
    a = helpers.run_oracle("64", root + "/init.bin", rows, tabs, 200, {200})
    recs = {k: np.array(v) for k, v in helpers.read_rec(root + "/init.bin").items()}
    fl = recs["FLAGS"].copy()
    fl[0:2] = np.frombuffer(b"  ", dtype=np.uint8)   # DOLLAR '$ ' -> '  '
    recs["FLAGS"] = fl
    helpers.write_rec(root + "/init2.bin", recs)
    b = helpers.run_oracle("64", root + "/init2.bin", rows, tabs, 200, {200})
    assert any(not np.array_equal(a[200][f], b[200][f]) for f in helpers.FIELDS)


@pytest.mark.parametrize("idx", range(len(helpers.EDGE_VARIANTS)), ids=helpers.edge_id)
def test_sides_oracle64_and_host_match_reference(idx, tmp_path):
    """Basins with sub-grid 1-D flow sides and channels: the product's reading
    of the sides and channels (CRDRD2), DEPSFC's bank, weir and barrier
    heights, SETDST's channel end areas and CNSTNT's channel constants equal
    the reference's COMMON blocks; the FP64 oracle's CHANNL, FLW1DM_C /
    FLW1DM_M and SMPT2G side pass reproduce HB/UB/VB/HMX and every side's
    transport FLWSQR bit for bit."""
    root = str(tmp_path)
    case = helpers.edge_case(root, idx)
    checks = [1, 5, 31, 100, 200, 400]
    ref = helpers.run_reference(case, root + "/ref", 400, ",".join(map(str, checks)))
    rows, tabs, _ = helpers.product_export(case, root)
    ri = helpers.read_rec(ref + "/init.bin")
    mi = helpers.read_rec(root + "/init.bin")
    for k in mi:
        assert np.array_equal(ri[k], mi[k]), "init record %s differs" % k
    nl = int(mi["INTS"][18])
    assert nl >= 15 and len(mi["CUTL"]) >= 1 and int(mi["INTS"][17]) > int(mi["INTS"][16])
    assert int(mi["INTS"][19]) == helpers.EDGE_VARIANTS[idx][5]     # channels
    rrows = np.fromfile(ref + "/storm.bin").reshape(-1, 90)
    assert np.array_equal(rows[:len(rrows)], rrows)
    o = helpers.Oracle("64", root + "/init.bin")
    active = 0
    for k in range(400):
        if k in tabs:
            o.set_tables(tabs[k])
        rc, _ = o.advance(rows[k])
        assert rc == 0
        if k + 1 in checks:
            st = helpers.read_rec(ref + "/s%06d.bin" % (k + 1))
            for f in helpers.FIELDS:
                assert np.array_equal(o.get(f).ravel(), st[f]), "step %d %s differs" % (k + 1, f)
            assert np.array_equal(o.flwsqr(), st["FLWSQR"]), "step %d FLWSQR differs" % (k + 1)
            active = max(active, int((st["FLWSQR"] != 0).sum()))
    o.close()
    if helpers.EDGE_VARIANTS[idx][4] != "+":
        assert active >= nl // 2      # the sides carry flow
    else:
        assert active == 0            # NOF1D '+': FLW1DM off


def test_side_planted_mistake_is_caught(tmp_path):
    """The sides' wind-stress rotation (COS1D, SIN1D) turned by about 0.6 degrees
    in the oracle's input must change the run: the comparison above sees the
    side momentum terms."""
    root = str(tmp_path)
    case = helpers.edge_case(root, 0)
    rows, tabs, _ = helpers.product_export(case, root)
    a = helpers.run_oracle("64", root + "/init.bin", rows, tabs, 200, {200})
    recs = helpers.read_rec(root + "/init.bin")
    recs = {k: np.array(v) for k, v in recs.items()}
    ang = np.arctan2(recs["SIN1D"], recs["COS1D"]) + 0.01
    recs["COS1D"], recs["SIN1D"] = np.cos(ang), np.sin(ang)
    helpers.write_rec(root + "/init2.bin", recs)
    b = helpers.run_oracle("64", root + "/init2.bin", rows, tabs, 200, {200})
    assert any(not np.array_equal(a[200][f], b[200][f]) for f in helpers.FIELDS)


def test_channel_planted_mistake_is_caught(tmp_path):
    """The channels' widths changed by 10 % in the oracle's input must change the
    run: the channel flow reaches the surge field."""
    root = str(tmp_path)
    case = helpers.edge_case(root, 1)
    rows, tabs, _ = helpers.product_export(case, root)
    a = helpers.run_oracle("64", root + "/init.bin", rows, tabs, 200, {200})
    recs = {k: np.array(v) for k, v in helpers.read_rec(root + "/init.bin").items()}
    recs["CHNLD"][5:10] *= 1.1
    helpers.write_rec(root + "/init2.bin", recs)
    b = helpers.run_oracle("64", root + "/init2.bin", rows, tabs, 200, {200})
    assert any(not np.array_equal(a[200][f], b[200][f]) for f in helpers.FIELDS)


@pytest.mark.parametrize("idx", range(len(helpers.DOLLAR_VARIANTS)), ids=helpers.dollar_id)
def test_dollar_oracle64_and_host_match_reference(idx, tmp_path):
    """'1$' basins (squares dried at the start) and '2$' (MSY: DEPMSY and
    INTLHT's MSY branch): the product's initialisation equals the reference
    build's, and the FP64 oracle from it equals the reference at every
    checkpoint."""
    root = str(tmp_path)
    case = helpers.dollar_case(root, idx)
    ref = helpers.run_reference(case, root + "/ref", 300, ",".join(map(str, CHECKS)))
    rows, tabs, _ = helpers.product_export(case, root)
    ri = helpers.read_rec(ref + "/init.bin")
    mi = helpers.read_rec(root + "/init.bin")
    for k in mi:
        assert np.array_equal(ri[k], mi[k]), "init record %s differs" % k
    rrows = np.fromfile(ref + "/storm.bin").reshape(-1, 90)
    assert np.array_equal(rows[:len(rrows)], rrows)
    snaps = helpers.run_oracle("64", root + "/init.bin", rows, tabs, 300, set(CHECKS))
    for c in CHECKS:
        st = helpers.read_rec(ref + "/s%06d.bin" % c)
        for f in helpers.FIELDS:
            assert np.array_equal(snaps[c][f].ravel(), st[f]), "step %d %s differs" % (c, f)


def test_dollar_variants_change_the_start(tmp_path):
    """The variants are seen: the dried squares start dry, and MSY's rows
    1..50 are lowered and filled from both datums."""
    for idx in range(len(helpers.DOLLAR_VARIANTS)):
        d, g = helpers.DOLLAR_VARIANTS[idx]
        root = str(tmp_path / ("v%d" % idx))
        case = helpers.dollar_case(root, idx)
        helpers.product_export(case, root)
        mi = helpers.read_rec(root + "/init.bin")
        ni, nj = mi["DUMB3"][0], mi["DUMB3"][1]
        hb = mi["HB"].reshape(nj, ni)
        zb = mi["ZB"].reshape(nj, ni)
        if d == "1$":
            b, _ = helpers.dollar_basin(idx)
            dried = [(i, j) for i, j in b.extra["dry"] if i <= ni - 1 and j <= nj - 1]
            assert dried and all(hb[j - 1, i - 1] == -zb[j - 1, i - 1] for i, j in dried)
        else:
            assert np.any(hb[:nj - 1, :ni - 1] == 1.5 - 1.0) and np.any(hb[:nj - 1, :ni - 1] == 0.7)


@pytest.mark.parametrize("idx", range(len(helpers.TRK1992_CASES)), ids=helpers.trk1992_id)
def test_trk1992_host_matches_reference(idx, tmp_path):
    """Tracks in the original 1992 format (INPUTN2's default branch,
    inputn2.f:610-627; the landfall time INITAL parses, inital.f:206-215): the
    product's host initialisation and storm rows equal the reference build's,
    and the FP64 oracle from them equals it at every checkpoint."""
    root = str(tmp_path)
    case = helpers.trk1992_case(root, idx)
    ref = helpers.run_reference(case, root + "/ref", 300, ",".join(map(str, CHECKS)))
    rows, tabs, _ = helpers.product_export(case, root)
    ri = helpers.read_rec(ref + "/init.bin")
    mi = helpers.read_rec(root + "/init.bin")
    for k in mi:
        assert np.array_equal(ri[k], mi[k]), "init record %s differs" % k
    rrows = np.fromfile(ref + "/storm.bin").reshape(-1, 90)
    assert np.array_equal(rows[:len(rrows)], rrows)
    snaps = helpers.run_oracle("64", root + "/init.bin", rows, tabs, 300, set(CHECKS))
    for c in CHECKS:
        st = helpers.read_rec(ref + "/s%06d.bin" % c)
        for f in helpers.FIELDS:
            assert np.array_equal(snaps[c][f].ravel(), st[f]), "step %d %s differs" % (c, f)
