"""Shared test helpers: record-file I/O, the oracle (CPU restatement, ctypes),
seeded cases, and the reference build when it is present."""
from __future__ import annotations

import ctypes
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "slosh-test_amd"))
import synth  # noqa: E402
import sloshgpu  # noqa: E402

ORACLE = {"32": os.path.join(REPO, "oracle", "liborc32.so"),
          "64": os.path.join(REPO, "oracle", "liborc64.so")}
REFDRV = os.path.join(REPO, "oracle", "_ref", "refdrv64")
GOLDEN = os.path.join(REPO, "tests", "golden")
TABLES = ["S", "C", "P", "DELP", "SW", "CW"]
FIELDS = ["HB", "UB", "VB", "HMX"]
_DT = {"d": np.float64, "i": np.int32, "h": np.int16, "c": np.uint8}


def read_rec(path):
    """Tagged records: name[16] type[1] pad[3] n(int64) data."""
    out = {}
    data = open(path, "rb").read()
    p = 0
    while p < len(data):
        name = data[p:p + 16].split(b"\0")[0].decode()
        t = chr(data[p + 16])
        n = struct.unpack("<q", data[p + 20:p + 28])[0]
        p += 28
        a = np.frombuffer(data, dtype=_DT[t], count=n, offset=p)
        p += n * np.dtype(_DT[t]).itemsize
        out[name] = a
    return out


def write_rec(path, recs):
    with open(path, "wb") as fh:
        for name, a in recs.items():
            a = np.ascontiguousarray(a)
            t = {np.dtype(np.float64): "d", np.dtype(np.int32): "i", np.dtype(np.int16): "h",
                 np.dtype(np.uint8): "c"}[a.dtype]
            fh.write(name.encode().ljust(16, b"\0") + t.encode() + b"\0\0\0")
            fh.write(struct.pack("<q", a.size))
            fh.write(a.tobytes())


def read_tables(path):
    """Table sets by the step after which they took effect."""
    tabs = {}
    raw = open(path, "rb").read()
    p = 0
    cur = None
    while p < len(raw):
        name = raw[p:p + 16].split(b"\0")[0].decode()
        t = chr(raw[p + 16])
        n = struct.unpack("<q", raw[p + 20:p + 28])[0]
        p += 28
        es = 8 if t == "d" else 4
        a = np.frombuffer(raw, dtype=np.float64 if t == "d" else np.int32, count=n, offset=p)
        p += n * es
        if name == "STEP":
            cur = int(a[0])
            tabs[cur] = {}
        else:
            tabs[cur][name] = a
    return tabs


class Oracle:
    """The CPU restatement of one step (oracle/orc.c), REAL = float ('32') or
    double ('64'), opened from an init record file."""

    def __init__(self, prec, init_path):
        L = ctypes.CDLL(ORACLE[prec])
        L.orc_open.restype = ctypes.c_void_p
        L.orc_open.argtypes = [ctypes.c_char_p]
        for f in ("orc_advance", "orc_get", "orc_put", "orc_set_tables", "orc_call", "orc_close",
                  "orc_info"):
            getattr(L, f).argtypes = None
        self.L = L
        self.c = ctypes.c_void_p(L.orc_open(init_path.encode()))
        if not self.c:
            raise RuntimeError("oracle could not open %s" % init_path)
        a, b = ctypes.c_int(), ctypes.c_int()
        L.orc_info(self.c, ctypes.byref(a), ctypes.byref(b))
        self.imxb, self.jmxb = a.value, b.value
        self.diag = (ctypes.c_int * 8)()

    def set_tables(self, tab: dict):
        tb = np.ascontiguousarray(np.concatenate([tab[n] for n in TABLES]), dtype=np.float64)
        self.L.orc_set_tables(self.c, tb.ctypes.data_as(ctypes.c_void_p))

    def advance(self, row):
        r = np.ascontiguousarray(row, dtype=np.float64)
        rc = self.L.orc_advance(self.c, r.ctypes.data_as(ctypes.c_void_p), self.diag)
        return rc, list(self.diag)[:5]

    def call(self, routine, row):
        r = np.ascontiguousarray(row, dtype=np.float64)
        return self.L.orc_call(self.c, routine.encode(), r.ctypes.data_as(ctypes.c_void_p), self.diag)

    def get(self, name):
        out = np.zeros(self.imxb * self.jmxb)
        self.L.orc_get(self.c, name.encode(), out.ctypes.data_as(ctypes.c_void_p))
        return out.reshape(self.jmxb, self.imxb)

    def flwsqr(self):
        out = np.zeros(41000)
        n = self.L.orc_get(self.c, b"FLWSQR", out.ctypes.data_as(ctypes.c_void_p))
        return out[:n]

    def put(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.L.orc_put(self.c, name.encode(), a.ctypes.data_as(ctypes.c_void_p))

    def close(self):
        if self.c:
            self.L.orc_close(self.c)
            self.c = ctypes.c_void_p()


def run_oracle(prec, init_path, rows, tabs, nsteps, checkpoints=(), callback=None):
    """Advance the oracle nsteps from init; rows[k] is the storm row after step k."""
    o = Oracle(prec, init_path)
    snaps = {}
    events = 0
    for k in range(nsteps):
        if k in tabs:
            o.set_tables(tabs[k])
        rc, d = o.advance(rows[k])
        if rc != 0:
            raise RuntimeError("oracle step %d returned %d" % (k + 1, rc))
        events += d[0]
        if (k + 1) in checkpoints:
            snaps[k + 1] = {f: o.get(f) for f in FIELDS}
        if callback:
            callback(k + 1, o)
    snaps["final"] = {f: o.get(f) for f in FIELDS}
    snaps["events"] = events
    o.close()
    return snaps


def make_case(root, basin, track):
    os.makedirs(root, exist_ok=True)
    p = synth.write_case(root, basin, track)
    with open(os.path.join(root, "bnt.txt"), "w") as fh:
        fh.write(p["bntline"] + "\n")
    p["bnt_path"] = os.path.join(root, "bnt.txt")
    return p


def product_export(case, root):
    """Host initialisation of the product (no GPU needed) exported to records."""
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, [case["trk"]])
    X.export(0, root + "/init.bin", root + "/storm.bin", root + "/tables.bin")
    nsteps = X.nsteps[0]
    X.close()
    B.close()
    rows = np.fromfile(root + "/storm.bin").reshape(-1, 90)
    return rows, read_tables(root + "/tables.bin"), nsteps


def have_reference():
    return os.path.exists(REFDRV)


def run_reference(case, outdir, nsteps, dump):
    os.makedirs(outdir, exist_ok=True)
    r = subprocess.run([REFDRV, case["bnt_path"], case["dta"], case["trk"], outdir, str(nsteps), dump],
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError("reference run failed: %s %s" % (r.stdout[-500:], r.stderr[-500:]))
    return outdir


def tmpdir(prefix="slg_"):
    return tempfile.mkdtemp(prefix=prefix)


def gpu_vs_oracle(case, nsteps, checks, storms=None, path=0, tile_rows=16, precision=32, setup=None):
    """Run one storm of a case on the GPU and the FP32 oracle side by side.

    Returns {step: {field: (gpu, oracle)}} at each checkpoint (plus the last
    step) and the oracle's drying-cascade event count.  ``storms`` lists extra
    track files batched with the case's track (the case's own storm is 0);
    ``path``/``tile_rows`` select the launch path (sloshgpu.PATH_*); ``precision``
    the arithmetic (32 against the FP32 oracle, 64 against the FP64 one);
    ``setup(batch)`` runs after the upload."""
    root = tmpdir("gvo_")
    rows, tabs, total = product_export(case, root)
    n = min(nsteps, total)
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    X = sloshgpu.Batch(B, [case["trk"]] + list(storms or []), precision=precision)
    X.set_path(path, tile_rows)
    X.upload()
    if setup:
        setup(X)
    o = Oracle(str(precision), root + "/init.bin")
    out, events = {}, 0
    want = set(c for c in checks if c <= n) | {n}
    for k in range(n):
        if k in tabs:
            o.set_tables(tabs[k])
        rc, d = o.advance(rows[k])
        assert rc == 0, "oracle step %d returned %d" % (k + 1, rc)
        events += d[0]
        if (k + 1) in want:
            X.step(k + 1 - X.current_step)
            X.sync()
            out[k + 1] = {f: (X.get(0, f), o.get(f)) for f in FIELDS}
            if B.nsides:
                out[k + 1]["FLWSQR"] = (X.flwsqr(0), o.flwsqr())
    o.close()
    X.close()
    B.close()
    return out, events


# Basin flag variants: (grid type, ebsn1, ismooth).  ebsn1 selects the smoothing
# cadence INCSM and whether SMPT2G runs (inital.f:468-497); ismooth selects the
# XFILTER weights (xfilter.f); grid ' ' polar, '$' elliptic, '+' hyperbolic.
VARIANTS = [
    (" ", " ", " "), (" ", "+", " "), (" ", "=", " "), (" ", "!", " "), (" ", "@", " "),
    (" ", "%", " "), (" ", "^", " "), (" ", "*", " "), (" ", "&", " "), (" ", "#", "-"),
    (" ", " ", "@"), (" ", "!", "$"), (" ", "@", "*"), (" ", "^", "%"), (" ", " ", "&"),
    (" ", " ", "="), (" ", " ", "^"), (" ", " ", "#"), ("$", " ", " "), ("$", "#", "+"),
    ("+", " ", " "), ("+", "&", "-"),
]


def variant_id(idx):
    g, e, m = VARIANTS[idx]
    return "%s-e%s-s%s" % ({" ": "polar", "$": "ellip", "+": "hyper"}[g],
                            "%02x" % ord(e), "%02x" % ord(m))


def variant_case(root, idx, ni=28, nj=24, hours=6):
    g, e, m = VARIANTS[idx]
    b, t = synth.tiny(seed=100 + idx, gtype=g, ni=ni, nj=nj, hours=hours)
    t.dp = [85.0] * len(t.dp)
    b.extra["ebsn1"] = e
    b.extra["ismooth"] = m
    return make_case(root, b, t)


# Closed island basins (DOLLAR '$ ', crdrd1.f:76: periodic in J): FLTER2
# instead of FILTER/XFILTER (tmstpc2.f:101-111), MOMNTM to row JMXB with no J
# side boundaries and row 1 copied from row JMXB (momntm.f:96-112, 464-469):
# (grid type, ebsn1, ismooth).
ISLAND_VARIANTS = [(" ", " ", " "), (" ", "!", "@"), ("$", " ", " "), ("+", "+", " ")]


def island_id(idx):
    g, e, m = ISLAND_VARIANTS[idx]
    return "island-%s-e%02x-s%02x" % ({" ": "polar", "$": "ellip", "+": "hyper"}[g], ord(e), ord(m))


def island_case(root, idx, ni=30, nj=26, hours=6):
    g, e, m = ISLAND_VARIANTS[idx]
    b, t = synth.tiny(seed=200 + idx, gtype=g, ni=ni, nj=nj, hours=hours)
    t.dp = [85.0] * len(t.dp)
    b.extra["ebsn1"] = e
    b.extra["ismooth"] = m
    b.extra["dollar"] = "$ "
    return make_case(root, b, t)


# The DOLLAR variants of CRDRD1's line 1 that change the initialisation:
# '1$' lists squares dried at the start (crdrd2.f:447-461, intlht.f:81-92);
# '2$' is the MSY basin: DEPMSY lowers ZB of rows 1..50 and the lake datum
# (initlz.f:52, depmsy.f), and INTLHT fills it by LDTMG's lake columns, ITREE
# '2' ocean and the sea datum elsewhere (intlht.f:96-121).  (flag, grid type).
DOLLAR_VARIANTS = [("1$", " "), ("1$", "$"), ("2$", " "), ("2$", "+")]


def dollar_id(idx):
    d, g = DOLLAR_VARIANTS[idx]
    return "dollar%s-%s" % (d[0], {" ": "polar", "$": "ellip", "+": "hyper"}[g])


def dollar_basin(idx, hours=6):
    d, g = DOLLAR_VARIANTS[idx]
    ni, nj = (40, 56) if d == "2$" else (30, 26)
    b, t = synth.tiny(seed=400 + idx, gtype=g, ni=ni, nj=nj, hours=hours)
    t.dp = [85.0] * len(t.dp)
    t.ocean, t.lake = 0.7, 1.5
    b.extra["dollar"] = d
    if d == "1$":
        rng = np.random.default_rng(400 + idx)
        # wet squares below the datum, plus one outside the grid (INTLHT skips it)
        wet = [(i + 1, j + 1) for j in range(nj - 1) for i in range(ni - 1) if b.zb[j, i] > 0]
        pick = rng.choice(len(wet), size=min(23, len(wet)), replace=False)
        b.extra["dry"] = [wet[k] for k in sorted(pick)] + [(ni + 3, 2)]
    return b, t


def dollar_case(root, idx, hours=6):
    b, t = dollar_basin(idx, hours)
    return make_case(root, b, t)


# Ver2019-03 basins with a Manning's n map (crdrd2.f:379-385): the bottom
# stress class IIK of each point follows its n (inital.f), (grid type, seed).
MANNING_VARIANTS = [(" ", 301), ("$", 302), ("+", 303)]


def manning_id(idx):
    g, _ = MANNING_VARIANTS[idx]
    return "manning-%s" % {" ": "polar", "$": "ellip", "+": "hyper"}[g]


def manning_case(root, idx, ni=30, nj=26, hours=6):
    g, seed = MANNING_VARIANTS[idx]
    b, t = synth.tiny(seed=seed, gtype=g, ni=ni, nj=nj, hours=hours)
    t.dp = [85.0] * len(t.dp)
    synth.add_manning(b, seed)
    return make_case(root, b, t)


# Tide coupling through the reference's C driver (-f_tide; slosh2.c:813-1027):
# (id, -f_tide value, basin kind).  'deep' has squares at least 290 ft deep,
# which V2.1 / V2.2 set to STATICHT + tide every step; 'sides' has 1-D sides
# with cuts narrower than a cell, which V2.2 / V2.4 take out of the tide
# (ISSUBGRID); VDEF takes the version from tide_flavor.txt.
TIDE_CASES = [("v2-polar", "V2", "polar"), ("v2-hyper", "V2", "hyper"), ("v21-deep", "V2.1.-5", "deep"),
              ("v22-sides", "V2.2.-3", "sides"), ("v24-sides", "V2.4", "sides"), ("vdef-deep", "VDEF", "deep")]


def tide_basin(kind, hours=6):
    if kind == "polar":
        return synth.tiny(seed=41, ni=30, nj=26, hours=hours)
    if kind == "hyper":
        return synth.tiny(seed=42, gtype="+", ni=30, nj=26, hours=hours)
    if kind == "deep":
        b = synth.make_basin("tdp", " ", 30, 26, seed=43, land_frac=0.2, n_barriers=20, dt=30.0,
                             size_km=(1.5, 4.0), depth_m=150.0)
        return b, synth.straight_track(b, 12.0, 330.0, 60.0, 15.0, hours=hours, landfall=hours // 2)
    b, t = edge_basin(1, hours=hours)
    return b, t


# Tracks in INPUTN2's original (1992) format: (grid, seed, IBGNT, ITEND, JHR,
# ocean datum, lake datum)
TRK1992_CASES = [(" ", 61, 40, 52, 46, 1.2, 0.8), ("+", 62, 30, 44, 38, 0.0, 0.0), ("$", 63, 44, 54, 48, 99.9, 2.0)]


def trk1992_id(idx):
    g, seed, beg, end, land, oc, lk = TRK1992_CASES[idx]
    return "trk1992-%s-s%d" % ({" ": "polar", "$": "ellip", "+": "hyper"}[g], seed)


def trk1992_case(root, idx):
    """A case whose trk file is in the 1992 format (the reference's default)."""
    g, seed, beg, end, land, oc, lk = TRK1992_CASES[idx]
    b, _ = synth.tiny(seed=seed, gtype=g, ni=30, nj=26)
    t = synth.straight_track(b, 11.0, 320.0, 65.0, 18.0, hours=99, landfall=land)
    t.ocean, t.lake = oc, lk
    t.date = (2024, 9, 12, 6, 30)
    case = make_case(root, b, t)
    with open(case["trk"], "w") as fh:
        fh.write(synth.trk1992_text(t, beg, end, land))
    case["b"] = b
    return case


def full_depth_case(root, hours=6):
    """A basin with squares deeper than 600 ft (cells of 3-8 km at dt = 20 s keep
    them within the gravity-wave limit), for -fullDepth (crdrd2.f:441-444)."""
    b = synth.make_basin("tfd", " ", 30, 26, seed=44, land_frac=0.2, n_barriers=20, dt=20.0,
                         size_km=(3.0, 8.0), depth_m=300.0)
    t = synth.straight_track(b, 12.0, 330.0, 60.0, 20.0, hours=hours, landfall=hours // 2)
    case = make_case(root, b, t)
    case["b"] = b
    return case


def tide_case(root, idx, hours=6, adj=False):
    """A case with its tide inputs (adj: with a datum-adjustment file); returns
    (case, -f_tide value)."""
    cid, flag, kind = TIDE_CASES[idx]
    b, t = tide_basin(kind, hours)
    case = make_case(root, b, t)
    d = synth.write_tide_db(root, b, case["basin"], seed=50 + idx, flavor="V2.2.-3" if flag == "VDEF" else None,
                            adj=adj)
    case["b"] = b
    case["t"] = t
    case["tidedir"] = d
    return case, flag


# Basins with sub-grid 1-D flow sides (1-D sides with and without banks, raised
# weirs, cuts; 'T' canal sheltering and rotated stress directions) and
# channels: (grid type, seed, ebsn1, ismooth, nof1d, channels).  NOF1D '+'
# keeps SMPT2G's side pass but turns FLW1DM off (tmstpc1.f:92, tmstpc2.f:69).
EDGE_VARIANTS = [(" ", 5, " ", " ", " ", 0), ("$", 6, " ", " ", " ", 2), ("+", 7, " ", " ", " ", 0),
                 (" ", 8, "!", "@", " ", 3), (" ", 9, "+", " ", " ", 0), (" ", 10, " ", " ", "+", 1)]


def edge_id(idx):
    g, seed, e, m, n, c = EDGE_VARIANTS[idx]
    return "%s-s%d-e%02x-m%02x%s%s" % ({" ": "polar", "$": "ellip", "+": "hyper"}[g], seed, ord(e), ord(m),
                                       "-nof1d" if n == "+" else "", "-ch%d" % c if c else "")


def edge_basin(idx, hours=8):
    g, seed, e, m, n, c = EDGE_VARIANTS[idx]
    b, t = synth.tiny(seed=seed, gtype=g, ni=40, nj=34, hours=hours)
    b.extra["ebsn1"] = e
    b.extra["ismooth"] = m
    b.extra["nof1d"] = n
    synth.add_edges(b, seed=seed + 2)
    if c:
        synth.add_channels(b, seed=seed + 3, n=c)
    return b, t


def edge_case(root, idx, hours=8):
    b, t = edge_basin(idx, hours)
    return make_case(root, b, t)


# ------------------------------------------------------------------ drop-in CLI
REF_CLI = os.path.join(REPO, "oracle", "_ref", "slosh64")          # reference FP64 build
ORC_CLI = {"64": os.path.join(REPO, "oracle", "_ref", "slosh_orc64"),  # its C driver over the oracle
           "32": os.path.join(REPO, "oracle", "_ref", "slosh_orc32")}
GPU_CLI = os.path.join(REPO, "oracle", "_ref", "slosh_gpu")         # its C driver over libslosh_ftn


def _big_stack():
    import resource
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def run_cli(exe, case_root, basin, tag, timeout=900, extra=(), env_extra=None):
    """Run a reference-style CLI (slosh -basin -rootDir -trk -rex -env) on a case
    written by synth.write_case, with extra options and environment variables;
    returns (env_path, rex_path)."""
    env = os.path.join(case_root, tag + ".env")
    rex = os.path.join(case_root, tag + ".rex")
    r = subprocess.run([exe, "-basin", basin, "-rootDir", case_root, "-trk",
                        os.path.join(case_root, "storm.trk"), "-rex", rex, "-env", env, "-verbose", "0"]
                       + list(extra),
                       cwd=case_root, capture_output=True, text=True, timeout=timeout,
                       preexec_fn=_big_stack, env=dict(os.environ, **(env_extra or {})))
    if r.returncode != 0 or not os.path.exists(env):
        raise RuntimeError("%s failed (%d): %s %s" % (exe, r.returncode, r.stdout[-800:], r.stderr[-800:]))
    return env, rex


def read_env(path):
    """env file (EnvSave, slosh2.c:1611-1671): imxb, jmxb, 160-char comment,
    imxb*jmxb int16 tenths of a foot (999 dry), then 2 floats."""
    b = open(path, "rb").read()
    imxb, jmxb = struct.unpack("<ii", b[:8])
    comment = b[8:168].split(b"\0")[0]
    vals = np.frombuffer(b, dtype="<i2", count=imxb * jmxb, offset=168).reshape(jmxb, imxb)
    tail = b[168 + 2 * imxb * jmxb:]
    return imxb, jmxb, comment, vals, tail


def same_cli_outputs(env_a, rex_a, env_b, rex_b):
    """Byte comparison of two runs' env and rex files, except bytes the
    reference writes from uninitialised memory: the env comment after its
    terminating NUL (slosh2.c:1643 writes 160 bytes of a shorter string) and the
    100 track records RexSaveEnv copies after 'envl' (rex.c:57-120 parses a
    2013-format trk with a space-separated pattern that never matches, leaving
    the floats it writes unset).  Returns a list of differences (empty = same)."""
    out = []
    a, b = read_env(env_a), read_env(env_b)
    if a[:3] != b[:3]:
        out.append("env header")
    if not np.array_equal(a[3], b[3]):
        d = np.flatnonzero(a[3].ravel() != b[3].ravel())
        out.append("env: %d cells differ, first %d: %d vs %d" % (d.size, d[0], a[3].ravel()[d[0]],
                                                                  b[3].ravel()[d[0]]))
    if a[4] != b[4]:
        out.append("env datums")
    ra, rb = open(rex_a, "rb").read(), open(rex_b, "rb").read()
    ia, ib = ra.rfind(b"envl"), rb.rfind(b"envl")
    if len(ra) != len(rb) or ia != ib:
        out.append("rex size/layout %d/%d %d/%d" % (len(ra), len(rb), ia, ib))
    else:
        if ia >= 0:
            ra = ra[:ia + 4] + ra[ia + 4 + 2400:]
            rb = rb[:ib + 4] + rb[ib + 4 + 2400:]
        if ra != rb:
            d = next(k for k in range(len(ra)) if ra[k] != rb[k])
            out.append("rex differs from byte %d" % d)
    return out


def list_run_case(root):
    """Two storms on two basins (polar 'tst', elliptic 'etnz') written for a
    list run: a list file of the reference's version-0 layout (cstart.c:92-245:
    three header lines, then basin,trk,rex,env per run).  Consecutive runs must
    name different basins (setFileNames, setup.c, refuses a repeat)."""
    b1, t1 = synth.tiny(seed=71, ni=28, nj=24, hours=6)
    b2, t2 = synth.tiny(seed=72, gtype="$", ni=26, nj=22, hours=8)
    b2.name = "tnz"
    c1 = synth.write_case(root, b1, t1, trk_name="run1.trk")
    c2 = synth.write_case(root, b2, t2, trk_name="run2.trk")
    return [c1, c2]


def run_cli_list(exe, root, cases, tag, timeout=900, env_extra=None):
    """Run a reference-style CLI over a list file (-lst, lstType 0) of the
    cases; returns [(env, rex)] per run."""
    lst = os.path.join(root, tag + ".lst")
    outs = []
    with open(lst, "w") as fh:
        fh.write("list run\nbasin,trk,rex,env\n-----\n")
        for k, c in enumerate(cases):
            rex = os.path.join(root, "%s_%d.rex" % (tag, k))
            envf = os.path.join(root, "%s_%d.env" % (tag, k))
            fh.write("%s,%s,%s,%s\n" % (c["basin"], c["trk"], rex, envf))
            outs.append((envf, rex))
    env = dict(os.environ)
    env.update(env_extra or {})
    r = subprocess.run([exe, "-lst", lst, "-rootDir", root, "-rex", "yes", "-env", "yes", "-verbose", "0"],
                       cwd=root, capture_output=True, text=True, timeout=timeout, preexec_fn=_big_stack, env=env)
    if r.returncode != 0 or not all(os.path.exists(e) for e, _ in outs):
        raise RuntimeError("%s list run failed (%d): %s %s" % (exe, r.returncode, r.stdout[-800:], r.stderr[-800:]))
    return outs


def oracle_checkpoints(case, prec, checkpoints, root):
    """The oracle's HB/UB/VB/HMX of one storm (case: dta, bntline, trk) at each
    checkpoint step up to its run's end, and at the end of the last step run
    ("final"; the whole run when no checkpoint falls inside it); the host
    initialisation is the product's (exported)."""
    os.makedirs(root, exist_ok=True)
    rows, tabs, n = product_export(case, root)
    cps = sorted(c for c in checkpoints if c <= n)
    return run_oracle(prec, root + "/init.bin", rows, tabs, cps[-1] if cps else n, cps)


def oracle_many(jobs, workers=8):
    """oracle_checkpoints for several (case, prec, checkpoints, root) jobs in
    threads (the oracle's C step releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=workers) as ex:
        futs = [ex.submit(oracle_checkpoints, *j) for j in jobs]
        return [f.result() for f in futs]


def track_case(case, trk_path):
    """The case of one of a batch's storms: the same basin, another track."""
    c = dict(case)
    c["trk"] = trk_path
    return c
