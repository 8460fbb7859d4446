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
