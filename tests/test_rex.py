"""rex v2 files of batched runs, on the CPU.

* The oracle's frame encoder (oracle/orc_rex.c) equals the reference's own
  writer (rex.c RexSaveStep / RexSaveEnv with pack.c Stuff2_xxx, run through
  oracle/_ref/rexref64) frame for frame, on fields with dry cells, clipped
  values, negative heights and runs longer than 256.
* The product's writer (slg_rex_*), fed those frames, writes files equal byte
  for byte to the reference's: header and comment padding, track rounding,
  dates from the model clock, frame links, the track-file record of the
  envelope frame (2013 and 1992 layouts) and the header patch.
* The datum override of the env and rex trailers equals the reference's
  InitWater_TideModeOverride.
* The committed fixture (tests/golden/rex_small.npz, made by the reference's
  writer) pins the same without the reference build.
"""
import os

import numpy as np
import pytest

import helpers
import rexutil
import sloshgpu

need_ref = pytest.mark.skipif(not rexutil.have_rexref(), reason="reference rex writer (oracle/_ref) not built")

# clocks: 2025-09-01 00:10:00, a leap day, the last second of a year with a
# fraction, the first minute of a year, a day in 2003
CLOCKS = [1756685400.0, 1709208000.0, 1704067199.75, 1735689660.0, 1050000000.5]


def _frames(rng, zb, kinds):
    jm, im = zb.shape
    out = []
    for k, kind in enumerate(kinds):
        hb, _ = rexutil.frame_fields(rng, im, jm, kind, zb)
        t6 = [25.123456 + k, 80.987654 - k, 12.345, 271.115, 55.555, 17.125]
        out.append((t6, CLOCKS[k % len(CLOCKS)], hb))
    return out


@need_ref
@pytest.mark.parametrize("im,jm,seed", [(41, 31, 1), (70, 23, 2), (9, 300, 3)])
def test_oracle_frames_match_reference(tmp_path, im, jm, seed):
    rng = np.random.default_rng(seed)
    _, zb = rexutil.frame_fields(rng, im, jm)
    frames = _frames(rng, zb, ["mixed", "dry", "flat", "mixed"])
    hb_env = rexutil.frame_fields(rng, im, jm, "mixed", zb)[0]
    trk = str(tmp_path / "t.trk")
    rexutil.trk_2013(trk, 48, rng)
    spec = str(tmp_path / "spec.bin")
    rexutil.write_spec(spec, zb, frames, trk, hb_env)
    ref = rexutil.reference_rex(spec, str(tmp_path / "ref.rex"))
    _, rf, (_, env_bits) = rexutil.split_rex(ref)
    assert len(rf) == len(frames)
    for (t6, c, hb), (_, bits) in zip(frames, rf):
        assert rexutil.oracle_frame(hb, zb) == bits
    assert rexutil.oracle_frame(hb_env, zb) == env_bits


@need_ref
@pytest.mark.parametrize("case", ["2013", "2013_short", "2013_long", "legacy", "legacy_okee", "comment_short",
                                  "comment_quotes", "abbrev4"])
def test_product_writer_matches_reference(tmp_path, case):
    rng = np.random.default_rng(sum(case.encode()))
    im, jm = 37, 29
    hb0, zb = rexutil.frame_fields(rng, im, jm)
    frames = _frames(rng, zb, ["mixed"] * 5)
    hb_env = np.maximum(hb0, frames[0][2])
    trk = str(tmp_path / "t.trk")
    if case.startswith("legacy"):
        rexutil.trk_legacy(trk, rng, okee=case.endswith("okee"))
    else:
        rows = {"2013_short": 20, "2013_long": 130}.get(case)
        rexutil.trk_2013(trk, 48, rng, ocean="0.3", lake="-1.2", nrows=rows)
    comment = {"comment_short": "abc", "comment_quotes": 'storm "X" of 2025, a long comment of some length'}.get(
        case, "FileVersion=2013")
    abbrev = "hsf1" if case == "abbrev4" else "tny"
    spec = str(tmp_path / "spec.bin")
    rexutil.write_spec(spec, zb, frames, trk, hb_env, comment, abbrev)
    ref = rexutil.reference_rex(spec, str(tmp_path / "ref.rex"))
    bits = [(t6, c, rexutil.oracle_frame(hb, zb)) for t6, c, hb in frames]
    mine = rexutil.product_rex(str(tmp_path / "mine.rex"), zb, bits, trk, rexutil.oracle_frame(hb_env, zb), comment,
                               abbrev)
    assert len(mine) == len(ref)
    assert mine == ref


@need_ref
def test_datum_override_matches_reference(tmp_path):
    vals = ["0.0", "0.3", "-1.2", "1.7", "12.34", "49.9", "50.2", "149.7", "150.2", "199.9", "250.5", "-49.99",
            "99.9", "999.9", "-0.0", "3.3"]
    pairs = [(np.float32(float(a)), np.float32(float(b))) for a in vals for b in ("0.0", "-0.7")]
    ref = rexutil.reference_datums(pairs, str(tmp_path))
    rng = np.random.default_rng(0)
    for k, a in enumerate(vals):
        for m, b in enumerate(("0.0", "-0.7")):
            trk = str(tmp_path / "d.trk")
            rexutil.trk_2013(trk, 6, rng, ocean=a, lake=b)
            h1, h2 = sloshgpu.track_datums(trk)
            want = ref[2 * k + m]
            assert np.float32(h1).tobytes() == want[0].tobytes(), (a, b, h1, want)
            assert np.float32(h2).tobytes() == want[1].tobytes(), (a, b, h2, want)


def _py_frame(hb, zb, cap=256, trunc=True):
    """A slow restatement with two knobs, to show the fixture tells apart a
    run cap of 255 and rounding half up from the reference's encoding."""
    jm, im = hb.shape
    codes = []
    for i in range(im - 1):
        for j in range(jm - 1):
            if hb[j, i] + zb[j, i] == 0.0:
                v = 999
            else:
                x = hb[j, i] * 10 + .5
                v = int(x) if trunc else int(np.floor(x))
            v = -320 if v < -320 else (700 if (v > 700 and v != 999) else v)
            codes.append((701 if v == 999 else v) + 320)
    bits, run, prev = [], 0, None
    for c in codes:
        if c == prev and run < cap:
            run += 1
            continue
        if run:
            bits.append(format(run - 1, "08b") + format(prev, "010b"))
        prev, run = c, 1
    bits.append(format(run - 1, "08b") + format(prev, "010b"))
    s = "".join(bits)
    s += "0" * (-len(s) % 8)
    return bytes(int(s[k:k + 8], 2) for k in range(0, len(s), 8))


def test_rex_golden(tmp_path):
    """The reference's rex file of a committed case (frames, envelope and
    track record); the oracle and the product's writer reproduce it."""
    z = np.load(os.path.join(helpers.GOLDEN, "rex_small.npz"))
    zb, hbs, tracks, clocks, hb_env = z["zb"], z["hb"], z["track"], z["clock"], z["hb_env"]
    ref = bytes(z["ref_rex"])
    trk = str(tmp_path / "t.trk")
    with open(trk, "wb") as fh:
        fh.write(bytes(z["trk"]))
    _, rf, (_, env_bits) = rexutil.split_rex(ref)
    cap_seen = trunc_seen = False
    for k, (_, bits) in enumerate(rf):
        assert rexutil.oracle_frame(hbs[k], zb) == bits
        assert _py_frame(hbs[k], zb) == bits
        cap_seen |= _py_frame(hbs[k], zb, cap=255) != bits
        trunc_seen |= _py_frame(hbs[k], zb, trunc=False) != bits
    # a run cap of 255 or rounding half up would not pass
    assert cap_seen and trunc_seen
    assert rexutil.oracle_frame(hb_env, zb) == env_bits
    bits = [(tracks[k], float(clocks[k]), rexutil.oracle_frame(hbs[k], zb)) for k in range(len(hbs))]
    mine = rexutil.product_rex(str(tmp_path / "mine.rex"), zb, bits, trk, rexutil.oracle_frame(hb_env, zb),
                               str(z["comment"]), str(z["abbrev"]))
    assert mine == ref
