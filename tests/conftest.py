import glob
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import helpers  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


GOLDEN_CASES = sorted(os.path.basename(p)[5:-4] for p in glob.glob(os.path.join(helpers.GOLDEN, "case_*.npz")))


def load_golden(name):
    z = np.load(os.path.join(helpers.GOLDEN, "case_%s.npz" % name))
    return {k: z[k] for k in z.files}


def golden_case_files(g, root):
    """Write the golden inputs (dta, trk, bnt line) into root; return a case dict."""
    os.makedirs(os.path.join(root, "dta"), exist_ok=True)
    dta = os.path.join(root, "dta", bytes(g["in_dta_name"]).decode())
    open(dta, "wb").write(bytes(g["in_dta"]))
    trk = os.path.join(root, "storm.trk")
    open(trk, "wb").write(bytes(g["in_trk"]))
    bnt = bytes(g["in_bnt"]).decode()
    open(os.path.join(root, "bnt.txt"), "w").write(bnt + "\n")
    return {"dta": dta, "trk": trk, "bntline": bnt, "bnt_path": os.path.join(root, "bnt.txt")}


def golden_init(g, path):
    """Rebuild an oracle init file from the golden record (tables: the IIK=1
    slip column the basin uses; other columns are never read)."""
    recs = {}
    for k in g:
        if k.startswith("init_") and not k.endswith("_iik1"):
            recs[k[5:]] = g[k]
    for t in ("AR2", "AI2", "BR2", "BI2", "CR2", "CI2"):
        full = np.zeros((600, 245))
        full[:, 0] = g["init_%s_iik1" % t]
        recs[t] = full.ravel()
    helpers.write_rec(path, recs)
    return path


def golden_tables(g):
    out = {}
    for s in g["tab_steps"]:
        v = g["tab_%d" % s].reshape(6, 800)
        out[int(s)] = {n: v[i] for i, n in enumerate(helpers.TABLES)}
    return out


@pytest.fixture(params=GOLDEN_CASES)
def golden(request):
    return request.param, load_golden(request.param)
