"""Helpers of the rex tests: seeded frame fields, the reference's rex writer
(oracle/_ref/rexref64) driven from a spec file, the oracle's frame encoder,
the product's writer, and the layout of a rex v2 file (rex.c:264-650)."""
from __future__ import annotations

import ctypes
import os
import struct
import subprocess

import numpy as np

import helpers
import sloshgpu

REXREF = os.path.join(helpers.REPO, "oracle", "_ref", "rexref64")


REXOUT = os.path.join(helpers.REPO, "oracle", "_ref", "rexout")


def have_rexref():
    return os.path.exists(REXREF)


def have_rexout():
    return os.path.exists(REXOUT)


def rexout_abbrev(abbrev):
    """The rex header's abbreviation as rexout normalises it (rexout.c:534-543):
    the pnt file's first line must equal it (rexout.c:544-547)."""
    a = (abbrev + "\0" * 4)[:4]
    a = a[1:] + " " if a[0] == " " else a
    return (a[:3] if a[3] == " " else a).split("\0")[0]


def reference_rexout(rexfile, abbrev, xy, root):
    """The reference's rexout -style 0 over rexfile at squares xy (its pnt
    file: the basin abbreviation, then one "I J" line per square); the text."""
    pnt, out = os.path.join(root, "pts.pnt"), os.path.join(root, "pts.ref.txt")
    with open(pnt, "w") as fh:
        fh.write(abbrev + "\n" + "".join("%d %d\n" % (x, y) for x, y in xy))
    subprocess.run([REXOUT, "-rex", rexfile, "-pnt", pnt, "-style", "0", "-out", out], check=True, timeout=60,
                   stdout=subprocess.DEVNULL)
    with open(out) as fh:
        return fh.read()


def rexout_values(text):
    """Values of rexout style-0 text: float [frames, points], tenths as int."""
    rows = [ln.split(", ")[2:] for ln in text.splitlines()[4:]]
    return np.array([[round(float(v) * 10) for v in r] for r in rows], dtype=np.int64)


def compare_series(mine, ref, storm):
    """rexout's style-0 text against the product's: equal line by line, except
    the Storm line.  rexout builds the storm name with an overlapping strcpy
    (rexout.c:296-298, undefined behaviour) that garbles it with this C
    library; the product prints the name the header holds."""
    a, b = mine.splitlines(), ref.splitlines()
    assert len(a) == len(b)
    assert a[1] == "Storm=" + storm and b[1].startswith("Storm=") and len(b[1]) == len(a[1])
    for k, (x, y) in enumerate(zip(a, b)):
        if k != 1:
            assert x == y, (k, x, y)


def oracle_frame(hb, zb):
    """Bit stream of one frame by the oracle (orc_rex_frame); hb, zb [jmxb, imxb]."""
    L = ctypes.CDLL(helpers.ORACLE["64"])
    L.orc_rex_frame.restype = ctypes.c_long
    L.orc_rex_frame.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_int] * 2 + [ctypes.c_void_p]
    jm, im = hb.shape
    h = np.ascontiguousarray(hb, dtype=np.float64)
    z = np.ascontiguousarray(zb, dtype=np.float64)
    out = np.zeros((18 * (im - 1) * (jm - 1) + 7) // 8 + 8, dtype=np.uint8)
    n = L.orc_rex_frame(h.ctypes.data_as(ctypes.c_void_p), z.ctypes.data_as(ctypes.c_void_p), im, jm,
                        out.ctypes.data_as(ctypes.c_void_p))
    return bytes(out[:n])


def frame_fields(rng, im, jm, kind="mixed", zb=None):
    """zb (given or new) and an HB field on it with dry cells (HB = -ZB),
    values beyond the clip range [-32, 70] ft, negative heights and runs
    longer than 256 cells."""
    if zb is None:
        zb = np.round(rng.normal(-10.0, 25.0, size=(jm, im)))
    hb = rng.normal(1.0, 3.0, size=(jm, im))
    if kind == "mixed":
        dry = rng.random((jm, im)) < 0.3
        hb[dry] = -zb[dry]
        hb[:, : im // 3] = -zb[:, : im // 3]            # a dry block: long runs of 999
        hb[rng.random((jm, im)) < 0.02] = 85.0          # clipped to 70.0
        hb[rng.random((jm, im)) < 0.02] = -41.0         # clipped to -32.0
        flat = (slice(jm // 2, jm), slice(im // 2, im))
        hb[flat] = 2.34                                 # long runs of one value
        hb[rng.random((jm, im)) < 0.05] = -0.26         # rounds toward zero
    elif kind == "dry":
        hb = -zb.copy()
    elif kind == "flat":
        hb[:] = 1.25
    return hb, zb


def trk_2013(path, hours, rng, ocean="0.3", lake="-1.2", nrows=None):
    """Track file in the fixed 2013 layout Data_WriteTrk reads (rex.c:81-142):
    16 header lines with LandfallHour / BeginHour / EndHour on lines 7-9 and
    the datums on lines 11-12, then space-separated hourly rows."""
    L = ["FileVersion=2013", "# rex test track", "DateTime=2025-09-01T00:00:00", "DateHour=0",
         "InterpolationType=Linear", "PressureType=0", "LandfallHour=%d" % (hours // 2), "BeginHour=0",
         "EndHour=%d" % hours, "# datums", "OceanDatum=%s" % ocean, "LakeDatum=%s" % lake, "#", "#", "#",
         "Hour,Lat,Lon,DeltaP,RMW"]
    n = hours + 1 if nrows is None else nrows
    for h in range(n):
        L.append("%d %.4f %.4f %.2f %.2f" % (h, 25.0 + 0.1 * h + rng.random() * 0.01, -80.0 - 0.05 * h,
                                              40.0 + rng.random() * 20, 15.0 + rng.random() * 10))
    L.append("999 0 0 0 0")
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")


def trk_legacy(path, rng, okee=False):
    """Track file in the 1992 layout of the other branch of Data_WriteTrk
    (rex.c:143-190): two header lines, 100 rows whose first 18 columns are
    blanked, the hours line, then the datums line (an 'x' in column 11 marks
    Okeechobee datums)."""
    L = ["LEGACY TRACK", "HEADER"]
    for h in range(100):
        L.append("%-18s%5d%8.2f%8.2f%8.2f%8.2f%8.2f%8.2f%5d" % ("STORM", h, 25.0 + 0.1 * h, 80.0 + 0.05 * h,
                                                            10.0 + rng.random(), 300.0 + rng.random(),
                                                            50.0 + rng.random(), 20.0, 1))
    L.append("  0 48 24")
    L.append("datums follow")
    L.append("  1.2x 0.3  -0.8" if okee else "  0.7 -0.4")
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")


def write_spec(path, zb, frames, trk, hb_env, comment="FileVersion=2013", abbrev="tny"):
    """frames: list of (track6, clock, hb)."""
    jm, im = zb.shape
    with open(path, "wb") as fh:
        fh.write(struct.pack("<3i", im, jm, len(frames)))
        fh.write(comment.encode().ljust(201, b"\0")[:201])
        fh.write(abbrev.encode().ljust(5, b"\0")[:5])
        fh.write(np.ascontiguousarray(zb, dtype="<f8").tobytes())
        for t6, clock, hb in frames:
            fh.write(np.asarray(list(t6) + [clock], dtype="<f8").tobytes())
            fh.write(np.ascontiguousarray(hb, dtype="<f8").tobytes())
        fh.write(trk.encode().ljust(257, b"\0")[:257])
        fh.write(np.ascontiguousarray(hb_env, dtype="<f8").tobytes())


def reference_rex(spec, out):
    subprocess.run([REXREF, "rex", spec, out], check=True, timeout=120)
    with open(out, "rb") as fh:
        return fh.read()


def reference_datums(pairs, root):
    a = np.asarray(pairs, dtype="<f4").reshape(-1, 2)
    fi, fo = os.path.join(root, "d.in"), os.path.join(root, "d.out")
    a.tofile(fi)
    subprocess.run([REXREF, "datum", fi, fo], check=True, timeout=60)
    return np.fromfile(fo, dtype="<f4").reshape(-1, 2)


def product_rex(path, zb, frames_bits, trk, env_bits, comment="FileVersion=2013", abbrev="tny"):
    """frames_bits: list of (track6, clock, bits)."""
    jm, im = zb.shape
    R = sloshgpu.Rex(path, comment, abbrev, im, jm)
    for t6, clock, bits in frames_bits:
        R.frame(t6, clock, bits)
    R.env(trk, env_bits)
    R.close()
    with open(path, "rb") as fh:
        return fh.read()


def split_rex(data):
    """(header, [(frame record, bits)], (env record, bits)) of a rex v2 file:
    each frame record is 35 bytes (6 floats, 5 date bytes, the year, the
    next-frame offset); the envelope record is 2 415 bytes."""
    c0 = data[0]
    # str_len - 12 bytes of identifier and comment, fixed fields, 12 + 1 + 4
    hlen = 1 + c0 + 4 + 4 + 8 + 10 + 12 + 1 + 4
    off, frames = hlen, []
    while data[off:off + 4] != b"envl":
        nxt = struct.unpack("<i", data[off + 31:off + 35])[0]
        frames.append((data[off:off + 35], data[off + 35:nxt]))
        off = nxt
    return data[:hlen], frames, (data[off:off + 2415], data[off + 2415:])


def trk_canonical(t, path):
    """A synth track written in the fixed 2013 layout of trk_2013 (so that
    Data_WriteTrk reads defined values), with the track's own rows."""
    L = ["FileVersion=2013", "# rex test track", "DateTime=2025-09-01T00:00:00", "DateHour=%d" % t.hours[0],
         "InterpolationType=%s" % t.interp, "PressureType=0", "LandfallHour=%d" % t.landfall,
         "BeginHour=%d" % t.begin, "EndHour=%d" % t.end, "# datums", "OceanDatum=0.0", "LakeDatum=0.0", "#", "#",
         "#", "Hour,Lat,Lon,DeltaP,RMW"]
    for h, la, lo, d, r in zip(t.hours, t.lat, t.lon, t.dp, t.rmw):
        L.append("%d %.4f %.4f %.2f %.2f" % (h, la, lo, d, r))
    L.append("999 0 0 0 0")
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")


def rex_abbrev(basin):
    """bsnAbrev as setFileNames makes it (setup.c:118-146): a blank before a
    three-letter (polar) basin name."""
    return " " + basin if len(basin) == 3 else basin
