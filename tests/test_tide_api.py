"""Host side of the batched gridded tide (no GPU): the -f_tide values the
batch takes (ParseTide, usrparse.c:14-52), the errors slg_batch_set_tide
gives before it touches the device, and the datum-adjustment file the
synthetic tide database writes (ReadAdjDatum's layout, tideutil.c:647-734)."""
import os

import pytest

import helpers
import sloshgpu
from helpers import synth


def test_parse_tide_follows_parsetide():
    assert sloshgpu.parse_tide("V2") == (2, 0)
    assert sloshgpu.parse_tide("V2.1") == (21, 0)
    assert sloshgpu.parse_tide("V2.1.-5") == (21, -5)
    assert sloshgpu.parse_tide("V2.2.-3") == (22, -3)
    assert sloshgpu.parse_tide("V2.4.2") == (24, 2)
    for bad in ("V1", "V3", "T1", "0", "VDEF"):
        with pytest.raises(sloshgpu.SloshError):
            sloshgpu.parse_tide(bad)


def _batch(root, abbrev="tny"):
    b, t = synth.tiny(seed=41, ni=30, nj=26)
    case = helpers.make_case(root, b, t)
    d = synth.write_tide_db(root, b, case["basin"], seed=50, adj=True)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, [case["trk"]], precision=64)
    return b, case, d, B, X


def test_set_tide_refuses_bad_input_before_the_device(tmp_path):
    root = str(tmp_path)
    b, case, d, B, X = _batch(root)
    ft03, bhc = os.path.join(d, "ft03.dta"), os.path.join(d, case["basin"] + ".bhc")
    try:
        with pytest.raises(sloshgpu.SloshError, match="tide mode 3"):
            X.set_tide(ft03, bhc, None, 3)
        with pytest.raises(sloshgpu.SloshError, match="cannot open the harmonic-constant file"):
            X.set_tide(ft03, bhc + ".missing", None, 2)
        # a harmonic grid of another size
        other = os.path.join(root, "other.bhc")
        open(other, "wb").write(synth.bhc_bytes(synth.tiny(seed=41, ni=20, nj=18)[0], case["basin"], 50))
        with pytest.raises(sloshgpu.SloshError, match="harmonic constants for 20 x 18 squares, the basin has 30 x 26"):
            X.set_tide(ft03, other, None, 2)
    finally:
        X.close()
        B.close()


def test_adj_file_layout(tmp_path):
    b, _ = synth.tiny(seed=41, ni=30, nj=26)
    lines = synth.adj_text(b, 50).splitlines()
    assert lines[0].startswith("Longitude, Latitude, I, J, Adj")
    assert len(lines) == 1 + (b.imxb - 1) * (b.jmxb - 1)
    f = [c.strip() for c in lines[2].split(",")]
    assert (int(f[2]), int(f[3])) == (2, 1)      # I fastest, the harmonic file's order
    assert abs(float(f[0]) - int(-b.alno * 1e6) / 1e6) < 1e-7
