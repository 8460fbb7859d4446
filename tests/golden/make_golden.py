#!/usr/bin/env python3
"""Writes the golden vectors under tests/golden/ by running the reference SLOSH
build (oracle/_ref/refdrv64, compiled from /root/reference by oracle/Makefile) on
small seeded synthetic cases.

Each case_<name>.npz holds the inputs exactly as the reference read them (dta,
trk and ?basins.dta text) and what the reference computed from them in its
default FP64 build: the initialised COMMON state the step reads (bottom-stress
tables reduced to the slip-coefficient columns these basins use), every storm
row STMVAL left after each step, each radial table set FORCE1 produced, and
HB/UB/VB/HMX at checkpoints.

    python tests/golden/make_golden.py          (needs oracle/_ref/refdrv64)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import helpers  # noqa: E402
from helpers import synth  # noqa: E402

INIT_KEYS = ["DUMB3", "ZB", "ZBM", "ITREE", "IMANN", "KSKP", "YLT", "HB", "UB", "VB", "HMX",
             "ELPCT", "ELPDT", "ELPCL", "ELPDL", "ELPCL2", "ELPDL2", "SINL2", "SINT", "COST",
             "SINL", "COSL", "IS", "MS", "IE", "ME", "X1B", "CORL", "GAMAF", "HTER", "DATUM",
             "NBCPT", "ISH", "JSH", "INTS", "FLAGS", "STA", "S", "C", "P", "DELP", "SW", "CW",
             "W1218", "WMAX", "X12", "STRMSB", "WTEMP", "EGTH", "POLAR", "ELLIP", "ORIGN"]
TABS = ["AR2", "AI2", "BR2", "BI2", "CR2", "CI2"]


def cases():
    b1, t1 = synth.tiny(seed=5)
    t1.dp = [90.0] * len(t1.dp)
    yield "polar_small", b1, t1, 360, [1, 2, 10, 60, 120, 240]
    b2, t2 = synth.tiny(seed=7, gtype="$", ni=30, nj=24, hours=8)
    n = len(t2.dp)                       # deepening, widening storm: FORCE1 reruns
    t2.dp = list(np.linspace(40.0, 80.0, n))
    t2.rmw = list(np.linspace(12.0, 20.0, n))
    yield "ellip_small", b2, t2, 480, [1, 50, 200]
    b3, t3 = synth.tiny(seed=9, ni=28, nj=22, hours=6)
    b3.extra["ebsn1"] = "#"
    b3.extra["ismooth"] = "+"
    yield "polar_xfilter", b3, t3, 360, [1, 100]


def main():
    if not helpers.have_reference():
        sys.exit("oracle/_ref/refdrv64 is not built (make -C oracle)")
    for name, b, t, nsteps, checks in cases():
        root = helpers.tmpdir("gold_")
        case = helpers.make_case(root, b, t)
        out = helpers.run_reference(case, root + "/dump", nsteps, ",".join(str(c) for c in checks))
        init = helpers.read_rec(out + "/init.bin")
        rec = {"in_dta": np.frombuffer(open(case["dta"], "rb").read(), np.uint8),
               "in_trk": np.frombuffer(open(case["trk"], "rb").read(), np.uint8),
               "in_bnt": np.frombuffer(case["bntline"].encode(), np.uint8),
               "in_dta_name": np.frombuffer(os.path.basename(case["dta"]).encode(), np.uint8)}
        for k in INIT_KEYS:
            rec["init_" + k] = init[k]
        for k in TABS:   # IIK = 1 column (water/land default slip), all 600 depths
            rec["init_" + k + "_iik1"] = init[k].reshape(600, 245)[:, 0].copy()
        rec["rows"] = np.fromfile(out + "/storm.bin").reshape(-1, 90)
        tabs = helpers.read_tables(out + "/tables.bin")
        rec["tab_steps"] = np.array(sorted(tabs), dtype=np.int32)
        for s in sorted(tabs):
            rec["tab_%d" % s] = np.concatenate([tabs[s][n] for n in helpers.TABLES])
        for c in checks:
            st = helpers.read_rec(out + "/s%06d.bin" % c)
            for f in helpers.FIELDS:
                rec["s%d_%s" % (c, f)] = st[f]
        fin = helpers.read_rec(out + "/final.bin")
        for f in helpers.FIELDS:
            rec["final_%s" % f] = fin[f]
        rec["nsteps"] = np.array([nsteps])
        rec["checks"] = np.array(checks)
        np.savez_compressed(os.path.join(HERE, "case_%s.npz" % name), **rec)
        print(name, "steps", nsteps, "tables", len(tabs),
              "size %.0f KB" % (os.path.getsize(os.path.join(HERE, "case_%s.npz" % name)) / 1e3))


if __name__ == "__main__":
    main()
