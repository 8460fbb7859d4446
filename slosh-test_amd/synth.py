"""Seeded synthetic SLOSH inputs: basin (dta + bnt line) and track (trk 2013).

No real basin or track ships with the reference (they are downloaded assets), so
every workload here is generated.  The writers follow the reference readers
column for column:

* dta (Ver2014-08 layout): header read by ``CRDRD1`` (sorc/slosh.fd/crdrd1.f:102-213),
  body by ``CRDRD2`` (crdrd2.f:65-470): barriers ``(7(I4,I4,F3.0))``, flow/weir/
  channel/cut counts, history stations, the ITREE map ``(60A1)`` per J, print
  sheets and depths in 15-column strips ``(15F5.0)``.
* bnt line for ``memSetLLx``: ``A3,A1,A18,A9,1X,A10,2X,A9,1X,A10,1X,F5.1,1X,F5.1,
  1X,F6.4,4F5.1,20X,F6.5`` (savellx.c:392-431); latitudes as DDD.MMSSs (a2lat,
  savellx.c:332-366).
* trk, FileVersion=2013 (inputn2.f:110-605).

Depth comes from a seeded smooth Gaussian random field plus a shelf trend, and is
capped by the explicit scheme's gravity-wave limit on each cell: the largest
total depth allowed in a cell of size h at step dt is (C*h/dt)^2/g.  Without the
cap a random 60 m depth under a 0.5 km cell at 60 s breaks the CFL limit and the
reference itself blows up.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np
from scipy.ndimage import gaussian_filter

FT_PER_M = 3.280839895
G_FT = 32.2
MI_PER_KT = 1.150779


@dataclass
class Basin:
    """A generated basin: grid geometry plus per-point fields (1-based in files)."""
    name: str            # 3-letter code (the dta/bnt key)
    gtype: str           # ' ' polar, '$' elliptic, '+' hyperbolic (dta line-1 col 12)
    imxb: int            # momentum points in I (cells = imxb-1)
    jmxb: int
    irb: int
    ndb: int
    dela: float          # radians
    rmouth: float        # miles
    abqab: float
    azmth: float
    alto: float          # tangent point latitude (deg N)
    alno: float          # tangent point longitude (deg W, positive)
    dltin: float         # seconds
    zb: np.ndarray       # (jmxb, imxb) depth ft (positive water), int
    itree: np.ndarray    # (jmxb, imxb) single chars
    barriers: list = field(default_factory=list)   # (i, j, h) 1-based, h ft
    hist: list = field(default_factory=list)       # 10 (i, j) history points
    extra: dict = field(default_factory=dict)      # dsntn/asmpt/aaxis/baxis/gsize
    # sub-grid 1-D flow (crdrd2.f:104-357): flow sides (i, j, ktree, side, iang),
    # banks (i, j, side, bki, bke, wchn), weirs (i, j, ktree, side, hweir),
    # cuts (i, j, ktree, side, bank1, bank2, cutl, cutli, cutle, iang, hweir)
    edges: dict = field(default_factory=dict)
    # channels (crdrd2.f:286-299): (length mi, width ft, depth ft,
    # entrance squares ((i, j), (i, j)), exit squares ((i, j), (i, j)))
    channels: list = field(default_factory=list)

    @property
    def abbrev(self) -> str:
        """Basin argument of the reference CLI: ' tst' polar, 'etst' e, 'htst' h."""
        return {" ": "", "$": "e", "+": "h"}[self.gtype] + self.name

    @property
    def cells(self) -> int:
        return (self.imxb - 1) * (self.jmxb - 1)


# --------------------------------------------------------------------------------------
# grid geometry (the metric SETDST builds, sorc/slosh.fd/setdst.f:47-87)

def _cell_size_ft(b_imxb, b_jmxb, irb, ndb, dela, rmouth, abqab):
    """Cell size (ft) at height points: DELA*|dz/dw| of z = R(e^w + A e^-w)."""
    i = np.arange(1, b_imxb + 1) - 1 - irb + 0.5
    j = np.arange(1, b_jmxb + 1) - 1 - ndb + 0.5
    s = i[None, :] * dela
    t = j[:, None] * dela
    w = s + 1j * t
    dz = rmouth * np.abs(np.exp(w) - abqab * np.exp(-w))
    return dz * dela * 5280.0


def _grf(rng, shape, corr):
    f = gaussian_filter(rng.standard_normal(shape), sigma=corr / 2.0, mode="reflect")
    return (f - f.mean()) / (f.std() + 1e-12)


def make_basin(name="tst", gtype=" ", ni=100, nj=90, seed=1, land_frac=0.20,
               n_barriers=50, dt=60.0, size_km=(0.5, 3.0), alto=29.5, alno=89.5,
               depth_m=60.0, land_m=5.0, corr=10.0, cfl=0.45, hyp_axes=(60.0, 40.0)) -> Basin:
    """Generate a basin of ni x nj cells (height points), as in the workloads:
    polar with radial cell size growing geometrically from size_km[0] to size_km[1],
    elliptic ('$') or hyperbolic ('+') with the same depth/land generator.
    hyp_axes: the hyperbolic grid's AAXIS and BAXIS (miles); with GSIZE fixed
    the cell size near the focal line stays the same and larger axes make the
    cells grow more slowly with I."""
    rng = np.random.default_rng(seed)
    imxb, jmxb = ni + 1, nj + 1
    extra = {}
    if gtype == " ":
        dela = math.log(size_km[1] / size_km[0]) / ni
        irb = ni // 2
        ndb = nj // 2
        r_inner_mi = size_km[0] / 1.609344 / dela          # radius of ring I=1
        rmouth = r_inner_mi * math.exp(irb * dela)          # radius of ring IRB+1
        abqab = 0.0
    elif gtype == "$":
        # elliptic I: DEGREE = (90-ASMPT)/NDE, BAXIS = DSTNT/cos(ASMPT), AAXIS = 0,
        # so z = 2 RMOUTH sinh(w) (crdrd1.f:175-187, setdst.f:47-87)
        ndb = 0
        asmpt = 5.0
        dela = math.radians((90.0 - asmpt) / nj)
        irb = -int(round(0.25 / dela))                      # keep off the focal line
        w = (np.arange(ni) - irb + 0.5)[None, :] * dela + 1j * ((np.arange(nj) + 0.5) * dela)[:, None]
        dzmin = np.abs(np.cosh(w)).min() * 2.0 * dela       # cell size per mile of RMOUTH
        rmouth = size_km[0] / 1.609344 / dzmin
        dstnt = 2.0 * rmouth * math.cos(math.radians(asmpt))
        abqab = -1.0
        wm = ((ni // 2) - irb + 0.5) * dela + 1j * ((nj // 2) + 0.5) * dela
        zm = 2.0 * rmouth * np.sinh(wm)                      # grid midpoint, miles
        extra = dict(dstnt=dstnt, asmpt=asmpt, xmouth=float(zm.real), ymouth=float(zm.imag))
    elif gtype == "+":
        # hyperbolic (elliptic II): DELA = 2 GSIZE/(AAXIS+BAXIS) (crdrd1.f:188-195)
        ndb = nj // 2
        aaxis, baxis = hyp_axes
        gsize = 0.6
        dela = 2.0 * gsize / (aaxis + baxis)
        irb = -int(round(0.1 / dela))
        rmouth = 0.5 * (aaxis + baxis)
        abqab = (aaxis - baxis) / (aaxis + baxis)
        extra = dict(aaxis=aaxis, baxis=baxis, gsize=gsize)
    else:
        raise ValueError(gtype)

    # ---- depth / land field on height points (jmxb-1 x imxb-1) ----
    hsize = _cell_size_ft(imxb - 1, jmxb - 1, irb, ndb, dela, rmouth, abqab)
    f = _grf(rng, (nj, ni), corr)
    trend = np.linspace(-1.6, 1.6, ni)[None, :]             # land toward small I
    z = 0.55 * f + trend
    thr = np.quantile(z, land_frac)
    water = z > thr
    zw = (z - thr) / (z.max() - thr + 1e-9)                  # 0..1 in water
    zl = (thr - z) / (thr - z.min() + 1e-9)                  # 0..1 on land
    # land starts 1 ft above datum: cells exactly at datum would sit at the
    # still-water level and flood as zero-thickness sheets
    depth_ft = np.where(water, 1.0 + zw * (depth_m * FT_PER_M - 1.0),
                        -(1.0 + zl * (land_m * FT_PER_M - 1.0)))
    hmax = (cfl * hsize / dt) ** 2 / G_FT                    # gravity-wave limit, total depth
    depth_ft = np.where(water, np.minimum(depth_ft, np.maximum(hmax - 20.0, 1.0)), depth_ft)
    zb = np.zeros((jmxb, imxb))
    zb[:nj, :ni] = np.round(depth_ft)
    zb[nj, :] = zb[nj - 1, :]
    zb[:, ni] = zb[:, ni - 1]
    zb = zb.astype(int)

    # ---- ITREE map on momentum points ----
    itree = np.full((jmxb, imxb), "3", dtype="<U1")
    tree_f = _grf(rng, (jmxb, imxb), corr)
    zpad = np.pad(zb[:nj, :ni], 1, mode="edge")              # squares around momentum pts
    sq = np.stack([zpad[:-1, :-1], zpad[:-1, 1:], zpad[1:, :-1], zpad[1:, 1:]])
    mind, maxd = sq.min(0), sq.max(0)
    itree[(mind > 10)] = "2"                                 # open water: ocean winds
    itree[(maxd <= 0) & (tree_f > 0.9)] = "1"                # wooded land
    # boundaries: open water -> static height ('6'), land -> closed ('4'); very
    # shallow water next to land gets the shallow-water condition ('9') that
    # MNTMBD evaluates (mntmbd.f:69-101).
    def _bnd(jj, ii):
        if maxd[jj, ii] <= 0:
            return "4"
        if mind[jj, ii] <= 0 and maxd[jj, ii] < 8:
            return "9"
        return "6"
    for jj in range(jmxb):
        itree[jj, imxb - 1] = _bnd(jj, imxb - 1)
        itree[jj, 0] = "4"
    for jj in (0, jmxb - 1):
        for ii in range(1, imxb - 1):
            itree[jj, ii] = _bnd(jj, ii)

    # ---- barriers: coastal momentum points, crest 2..4 m ----
    coast = [(i, j) for j in range(2, jmxb - 1) for i in range(2, imxb - 1)
             if mind[j - 1, i - 1] <= 0 < maxd[j - 1, i - 1] and itree[j - 1, i - 1] not in "46"]
    rng.shuffle(coast)
    barriers = []
    for (i, j) in coast[:n_barriers]:
        h = int(round(rng.uniform(2.0, 4.0) * FT_PER_M))
        barriers.append((i, j, h))

    hist = []
    wet = np.argwhere(zb[:nj, :ni] > 0)
    pick = rng.choice(len(wet), size=10, replace=False)
    for k in pick:
        hist.append((int(wet[k][1]) + 1, int(wet[k][0]) + 1))

    return Basin(name=name, gtype=gtype, imxb=imxb, jmxb=jmxb, irb=irb, ndb=ndb,
                 dela=dela, rmouth=rmouth, abqab=abqab, azmth=180.0, alto=alto,
                 alno=alno, dltin=dt, zb=zb, itree=itree, barriers=barriers,
                 hist=hist, extra=extra)


def add_channels(b: Basin, seed=9, n=2):
    """Channels from pairs of coastal water squares (entrance) to pairs of low
    land squares (exit), 2-5 miles long, 60-200 ft wide, 6-15 ft deep."""
    rng = np.random.default_rng(seed)
    ni, nj = b.imxb - 1, b.jmxb - 1
    wet = [(i, j) for j in range(3, nj - 2) for i in range(3, ni - 3)
           if 1 <= b.zb[j - 1, i - 1] <= 20 and 1 <= b.zb[j - 1, i] <= 20]
    low = [(i, j) for j in range(3, nj - 2) for i in range(3, ni - 3)
           if -8 <= b.zb[j - 1, i - 1] <= -1 and -8 <= b.zb[j - 1, i] <= -1]
    rng.shuffle(wet)
    rng.shuffle(low)
    out = []
    for k in range(min(n, len(wet), len(low), 5)):
        (i0, j0), (il, jl) = wet[k], low[k]
        out.append((float(rng.choice([2.0, 3.5, 5.0])), float(rng.choice([60.0, 120.0, 200.0])),
                    float(rng.choice([6.0, 10.0, 15.0])), ((i0, j0), (i0 + 1, j0)), ((il, jl), (il + 1, jl))))
    b.channels = out
    return b


def add_edges(b: Basin, seed=7, n_flow=12, n_weirs=4, n_cuts=4, banks=True, crest_ft=(8, 14)):
    """Sub-grid 1-D flow on coastal water squares: each side gets barriers at
    both of its end points (so 2-D flow across it is blocked below the crest
    and the side carries 1-D flow), some sides with banks, raised weirs and
    cuts, some with 'T' (canal) sheltering and rotated stress directions.
    Sides stay one square inside the basin's outer rows and columns."""
    rng = np.random.default_rng(seed)
    ni, nj = b.imxb - 1, b.jmxb - 1
    cand = [(i, j) for j in range(3, nj - 2) for i in range(3, ni - 2)
            if 1 <= b.zb[j - 1, i - 1] <= 25]
    rng.shuffle(cand)
    used, out = set(), {"flow": [], "banks": [], "weirs": [], "cuts": []}
    ends = {1: ((0, 1), (0, 0)), 2: ((1, 1), (1, 0)), 3: ((0, 0), (1, 0)), 4: ((0, 1), (1, 1))}
    di = {1: -1, 2: 1, 3: 0, 4: 0}
    dj = {1: 0, 2: 0, 3: -1, 4: 1}
    want = [("flow", n_flow), ("weirs", n_weirs), ("cuts", n_cuts)]
    it = iter(cand)
    for kind, n in want:
        made = 0
        while made < n:
            try:
                i, j = next(it)
            except StopIteration:
                break
            side = int(rng.integers(1, 5))
            ii, jj = i + di[side], j + dj[side]
            if (i, j) in used or (ii, jj) in used or not (1 <= b.zb[jj - 1, ii - 1] <= 25):
                continue
            used.update({(i, j), (ii, jj)})
            crest = int(rng.integers(crest_ft[0], crest_ft[1] + 1))
            for (a, c) in ends[side]:
                b.barriers.append((i + a, j + c, crest))
            ktree = "T" if rng.random() < 0.25 else " "
            iang = int(rng.choice([0, 0, 30, 45, 120]))
            if kind == "flow":
                out["flow"].append((i, j, ktree, side, iang))
                if banks and rng.random() < 0.5:
                    out["banks"].append((i, j, side, int(rng.integers(-3, 4)), int(rng.integers(-3, 4)),
                                         float(rng.choice([0.25, 0.5, 0.75, 1.0]))))
            elif kind == "weirs":
                out["weirs"].append((i, j, ktree, side, float(rng.choice([0.0, -2.5, 1.5, 4.0]))))
            else:
                out["cuts"].append((i, j, ktree, side, int(rng.choice([0, 0, 2, -1])),
                                    int(rng.choice([0, 0, 3])), float(rng.choice([0.5, 1.0, 1.5])),
                                    float(rng.choice([0.0, 0.5, 1.0, 2.0])),
                                    float(rng.choice([0.0, 0.6, 1.0, 1.5])), iang,
                                    float(rng.choice([0.0, 0.0, -3.0]))))
            made += 1
    b.edges = out
    return b


# --------------------------------------------------------------------------------------
# writers

def _f10(*vals):
    return "".join("%10.6f" % v for v in vals)


def dta_text(b: Basin) -> str:
    """The basin as a Ver2014-08 dta file (crdrd1.f:102-213, crdrd2.f:65-470)."""
    L = []
    L.append("Ver2019-03" if "manning" in b.extra else "Ver2014-08")
    L.append("Proj=Clarke")
    nm = ("SYN " + b.name.upper())[:10].ljust(10)
    # col 12 grid type, 13-14 island flag, 15 smoothing, 16 order, 17 hemi, 18-21 opts
    # col 15 EBSN1 (smoothing cadence), col 21 X-filter code (crdrd1.f:72-142)
    # (A10,1X,A1,A2,8A1): STA, EBSN, DOLLAR, EBSN1, EBSN2, ISOUTH, NOFLD, NOF1D, ITERRAIN, ISMOOTH
    L.append(nm + " " + b.gtype + b.extra.get("dollar", "  ") + b.extra.get("ebsn1", " ") + " " * 3 +
             b.extra.get("nof1d", " ") + " " + b.extra.get("ismooth", " "))
    L.append(_f10(b.azmth))
    if b.gtype in "$+":
        L.append(_f10(b.alto, b.alno))                       # TLAT TLONG (unused)
    L.append(_f10(b.alto, b.alno))
    if b.gtype == " ":
        L.append(_f10(math.degrees(b.dela)))
    if b.gtype == "$":
        L.append(_f10(b.extra["dstnt"], b.extra["asmpt"]))
    if b.gtype == "+":
        L.append(_f10(b.extra["aaxis"], b.extra["baxis"], b.extra["gsize"]))
    ire = b.imxb - 1 - b.irb
    nde = b.jmxb - 1 - b.ndb
    L.append("%5d%5d%5d%5d%5d" % (b.irb, ire, b.ndb, nde, 0))
    L.append(_f10(*(tuple(b.dltin) if isinstance(b.dltin, (tuple, list)) else (b.dltin,) * 3)))
    if b.gtype != "+":                                       # XMOUTH YMOUTH
        L.append(_f10(b.extra.get("xmouth", b.rmouth), b.extra.get("ymouth", 0.0)))
    # barriers (crdrd2.f:78-100): count with the letter code for thousands
    n = len(b.barriers)
    L.append("%3d  " % n if n < 1000 else "%3d%s " % (n % 1000, "XYZWVUT"[n // 1000 - 1]))
    for k in range(0, n, 7):
        L.append("".join("%4d%4d%3d" % bb for bb in b.barriers[k:k + 7]))
    E = b.edges
    flow, banks, weirs, cuts = (E.get(k, []) for k in ("flow", "banks", "weirs", "cuts"))
    L.append("%3d %s%3d" % (len(flow), "X" if banks else " ", 0))   # (I3,2A1,I3)
    for k in range(0, len(flow), 6):                                 # (6(2I4,A1,I1,I3))
        L.append("".join("%4d%4d%s%1d%3d" % e for e in flow[k:k + 6]))
    if flow and banks:
        L.append("%3d  " % len(banks))                              # (I3,2A1)
        for k in range(0, len(banks), 4):                            # (4(2I4,A1,I1,2F3.0,F4.2))
            L.append("".join("%4d%4d %1d%3d%3d%4.2f" % e for e in banks[k:k + 4]))
    L.append("%3d  " % len(weirs))
    for k in range(0, len(weirs), 5):                                # (5(2I3,A1,I1,F5.1))
        L.append("".join("%3d%3d%s%1d%5.1f" % e for e in weirs[k:k + 5]))
    L.append("".join(("STA%02d" % (k + 1)).ljust(10) for k in range(6)))
    L.append("".join(("STA%02d" % (k + 7)).ljust(10) for k in range(4)))
    L.append("".join("%3d" % h[0] for h in b.hist))
    L.append("".join("%3d" % h[1] for h in b.hist))
    ch = b.channels
    L.append("%3d" % len(ch))                                # channels
    if ch:
        L.append("".join("%5.1f" % c[0] for c in ch))         # (5F5.1)
        L.append("".join("%6.0f" % c[1] for c in ch))         # (5F6.0)
        L.append("".join("%5.0f" % c[2] for c in ch))         # (5F5.0)
        for part, axis in ((3, 0), (3, 1), (4, 0), (4, 1)):   # IPT0 JPT0 IPTL JPTL (10I3)
            L.append("".join("%3d%3d" % (c[part][0][axis], c[part][1][axis]) for c in ch))
    L.append("%3d  " % len(cuts))                            # cuts, one per record
    for e in cuts:                                           # (2I4,A1,I1,2F4.0,3F5.2,I3,F5.1)
        L.append("%4d%4d%s%1d%4d%4d%5.2f%5.2f%5.2f%3d%5.1f" % e)
    L.append("  1")                                          # MGPT: tree map follows
    for j in range(b.jmxb):
        row = "".join(b.itree[j, :])
        for k in range(0, len(row), 60):
            L.append(row[k:k + 60])
    if "manning" in b.extra:                                 # Ver2019-03: Manning's n x 1000 (20I3)
        L.append("  1")
        for j in range(b.jmxb):
            row = [int(v) for v in b.extra["manning"][j, :]]
            for k in range(0, len(row), 20):
                L.append("".join("%3d" % v for v in row[k:k + 20]))
    L.append("  1 X")                                        # print sheets
    L.append("%3d" % 1)
    L.append("%3d" % min(b.jmxb - 1, 999))
    for k in range(0, b.imxb, 15):
        for j in range(b.jmxb):
            vals = b.zb[j, k:min(k + 15, b.imxb)]
            L.append("".join("%5d" % v for v in vals))
    if b.extra.get("dollar") == "1$":                        # dry cells (I4), then (9(I4,1X,I4,1X))
        dry = b.extra.get("dry", [])
        L.append("%4d" % len(dry))
        for k in range(0, len(dry), 9):
            L.append("".join("%4d %4d " % p for p in dry[k:k + 9]))
    return "\n".join(L) + "\n"


def _dms(x):
    """Decimal degrees -> the DDD.MMSSs text a2lat parses (savellx.c:332-366)."""
    s = 1 if x >= 0 else -1
    x = abs(x)
    d = int(x)
    m = int((x - d) * 60)
    sec10 = int(round(((x - d) * 60 - m) * 600))
    if sec10 >= 600:
        sec10 -= 600
        m += 1
    if m >= 60:
        m -= 60
        d += 1
    return "%.5f" % (s * (d + m / 100.0 + sec10 / 100000.0))


def bnt_line(b: Basin) -> str:
    """One ?basins.dta line for memSetLLx (savellx.c:396-431, 704-765).

    G is the tangent point; P is placed so the projection matches the dta grid:
    polar P is the pole, XMOUTH miles toward local north (AZMTH=180 puts +x south)."""
    lat0 = b.alto
    if b.gtype == " ":
        dist = b.rmouth
        typ = " "
        xig = b.irb + 1.0
        yjg = b.ndb + 1.0
        gsize = 2.0 * dist * math.sinh(b.dela / 2.0)
        elip = 0.0
    elif b.gtype == "$":
        dist = b.extra["dstnt"]
        typ = "e"
        xig = b.irb + 1.0
        yjg = b.ndb + 1.0 + (90.0 - b.extra["asmpt"]) / math.degrees(b.dela)
        gsize = b.extra["asmpt"]
        elip = 0.0
    else:
        a, bb = b.extra["aaxis"], b.extra["baxis"]
        elip = bb / a
        dist = b.rmouth * (1.0 + b.abqab)
        typ = "h"
        xig = b.irb + 1.0
        yjg = b.ndb + 1.0
        gsize = b.dela * b.rmouth
    plat = lat0 + dist / 69.05
    plon = b.alno
    s = b.name.upper()[:3].ljust(3) + typ + ("SYNTHETIC " + b.name.upper()).ljust(18)[:18]
    s += _dms(lat0).rjust(9)[:9] + " " + _dms(b.alno).rjust(10)[:10] + "  "
    s += _dms(plat).rjust(9)[:9] + " " + _dms(plon).rjust(10)[:10] + " "
    s += "%5.1f %5.1f %6.4f" % (xig, yjg, gsize)
    s += "%5.1f%5.1f%5.1f%5.1f" % (1.0, float(b.imxb), 1.0, float(b.jmxb))
    s += " " * 20 + "%6.5f" % elip
    return s


@dataclass
class Track:
    hours: list
    lat: list
    lon: list          # degrees east (negative west), as the trk file holds it
    dp: list
    rmw: list
    begin: int
    end: int
    landfall: int
    interp: str = "Linear"
    ocean: float = 0.0  # OceanDatum, LakeDatum (ft)
    lake: float = 0.0
    date: tuple = (2025, 9, 1, 0, 0)   # DateTime: year, month, day, hour, minute


def straight_track(b: Basin, speed_kt=12.0, heading=330.0, dp=60.0, rmw=25.0,
                   hours=48, landfall=24, offset_mi=0.0) -> Track:
    """Straight track crossing the tangent point at `landfall` (+ lateral offset)."""
    v = speed_kt * MI_PER_KT
    th = math.radians(heading)
    lat0 = b.alto + offset_mi * math.cos(th + math.pi / 2) / 69.05
    lon0 = -b.alno + offset_mi * math.sin(th + math.pi / 2) / (69.05 * math.cos(math.radians(b.alto)))
    H, la, lo = [], [], []
    for h in range(0, hours + 3):
        d = v * (h - landfall)
        la.append(lat0 + d * math.cos(th) / 69.05)
        lo.append(lon0 + d * math.sin(th) / (69.05 * math.cos(math.radians(b.alto))))
        H.append(h)
    n = len(H)
    return Track(H, la, lo, [dp] * n, [rmw] * n, 0, hours, landfall)


def spline_track(b: Basin, seed=2, hours=72, dp=(40.0, 80.0), rmw=(15.0, 30.0),
                 speeds=(8.0, 20.0)) -> Track:
    """Curved track: 4 seeded waypoints, hourly values by the reference's cubic
    spline (trk InterpolationType=Spline, inputn2.f:513-566)."""
    rng = np.random.default_rng(seed)
    step_h = 26
    H = [0, step_h, 2 * step_h, 3 * step_h]
    hd = math.radians(330.0 + rng.uniform(-25, 25))
    la, lo = [], []
    spd = np.linspace(speeds[0], speeds[1], 3) * MI_PER_KT
    # landfall near hour 36: waypoint 1 before the coast, waypoint 2 after
    pts = [(0.0, 0.0)]
    for k in range(3):
        turn = rng.uniform(-0.35, 0.35)
        hd += turn
        d = spd[k] * step_h
        x, y = pts[-1]
        pts.append((x + d * math.sin(hd), y + d * math.cos(hd)))
    xs = np.array([p[0] for p in pts])
    ys = np.array([p[1] for p in pts])
    # shift so the track passes the tangent point near hour 36
    f = 36.0 / step_h
    k = int(f)
    cx = xs[k] + (f - k) * (xs[k + 1] - xs[k])
    cy = ys[k] + (f - k) * (ys[k + 1] - ys[k])
    xs -= cx
    ys -= cy
    for x, y in zip(xs, ys):
        la.append(b.alto + y / 69.05)
        lo.append(-b.alno + x / (69.05 * math.cos(math.radians(b.alto))))
    dps = list(np.linspace(dp[0], dp[1], 4))
    rms = list(np.linspace(rmw[0], rmw[1], 4))
    return Track(H, la, lo, dps, rms, 0, hours, 36, interp="Spline")


def trk_text(t: Track) -> str:
    L = ["FileVersion=2013",
         "# synthetic track (seeded)",
         "DateTime=%04d-%02d-%02dT%02d:%02d:00" % tuple(t.date),
         "DateHour=%d" % t.hours[0],
         "LandfallHour=%d" % t.landfall,
         "BeginHour=%d" % t.begin,
         "EndHour=%d" % t.end,
         "OceanDatum=%.1f" % t.ocean,
         "LakeDatum=%.1f" % t.lake,
         "InterpolationType=%s" % t.interp,
         "PressureType=0",
         "Hour,Lat,Lon,DeltaP,RMW"]
    for h, la, lo, d, r in zip(t.hours, t.lat, t.lon, t.dp, t.rmw):
        L.append("%d,%.4f,%.4f,%.2f,%.2f" % (h, la, lo, d, r))
    L.append("999,0,0,0,0")
    return "\n".join(L) + "\n"


def trk1992_text(t: Track, beg: int, end: int, land: int, title="SYNTHETIC STORM (SEEDED)") -> str:
    """The track in INPUTN2's original (1992) format (inputn2.f:610-627), laid out
    as the reference's stm2trk writes it (sorc/stm2trk.cd/stm2trk.c:447-502):
    two title lines, 100 hourly rows (hour I = 1..100: latitude, longitude west
    positive, forward speed, direction, deficit, RMW), IBGNT ITEND JHR, the
    landfall time (t.date is taken as the time of hour JHR) and the datums.
    t must hold hours 1..100."""
    at = {int(h): k for k, h in enumerate(t.hours)}
    L = ["HUR   %-60s" % title, " DELTA-P = %.0fMB ; RMW= %.0f ST. MI." % (t.dp[0], t.rmw[0])]
    for i in range(1, 101):
        k = at[i]
        row = "%17s%3d%8.4f%8.3f%8.2f%8.2f%8.2f%8.2f%5d" % ("NAP-----  " if i == land else " ", i, t.lat[k],
                                                         -t.lon[k], 0.0, 0.0, t.dp[k], t.rmw[k], i)
        L.append(row + (" ---NAP" if i == land else ""))
    L.append("%3d%3d%3d%16sIBGNT ITEND JHR" % (beg, end, land, " "))
    y, mo, d, h, mi = t.date
    L.append("HR%02d%02d %02d %3s %04d%7sNEAREST APPROACH, OR LANDFALL, TIME"
             % (h, mi, d, ["JAN", "FEB", "MAR", "APR", "MAY", "JUN", "JUL", "AUG", "SEP", "OCT", "NOV",
                           "DEC"][mo - 1], y, " "))
    L.append("%5.1f%5.1f%15sSEA AND LAKE DATUM" % (t.ocean, t.lake, " "))
    return "\n".join(L) + "\n"


def track_dict(t: Track) -> dict:
    """The track as slg_track fields (sloshgpu.track_struct): the values a trk
    file of trk_text(t) holds, rounded as that file writes them."""
    y, mo, d, h, mi = t.date
    return dict(year=y, month=mo, day=d, hour_of_day=h, minute=mi, date_hour=t.hours[0],
                landfall_hour=t.landfall, begin_hour=t.begin, end_hour=t.end,
                ocean_datum=float("%.1f" % t.ocean), lake_datum=float("%.1f" % t.lake),
                spline=1 if t.interp == "Spline" else 0,
                hour=[float(h) for h in t.hours], lat=[float("%.4f" % x) for x in t.lat],
                lon=[float("%.4f" % x) for x in t.lon], delp=[float("%.2f" % x) for x in t.dp],
                rmw=[float("%.2f" % x) for x in t.rmw])


def write_case(root: str, b: Basin, t: Track, trk_name="storm.trk"):
    """Write rootDir layout the reference CLI expects (usrparse.c:421-436):
    <root>/dta/<bsn>dta, <root>/bnt/[e|h]basins.dta, and the trk file."""
    os.makedirs(os.path.join(root, "dta"), exist_ok=True)
    os.makedirs(os.path.join(root, "bnt"), exist_ok=True)
    # elliptic / hyperbolic basins go by 'e' / 'h' + name (setup.c:41-64)
    pre = {" ": "", "$": "e", "+": "h"}[b.gtype]
    dta = os.path.join(root, "dta", pre + b.name.lower() + "dta")
    with open(dta, "w") as fh:
        fh.write(dta_text(b))
    bnt = os.path.join(root, "bnt", pre + "basins.dta")
    with open(bnt, "w") as fh:
        fh.write(bnt_line(b) + "\n")
    trk = os.path.join(root, trk_name)
    with open(trk, "w") as fh:
        fh.write(trk_text(t))
    return dict(dta=dta, bnt=bnt, trk=trk, bntline=bnt_line(b), basin=pre + b.name.lower())


# --------------------------------------------------------------------------------------
# the named workloads

def config1(seed=1):
    """Polar 100 x 90, 0.5 -> 3 km, ~20 % land, 50 barriers, 48 h at 60 s."""
    b = make_basin("tst", " ", 100, 90, seed=seed, land_frac=0.20, n_barriers=50, dt=60.0)
    return b, straight_track(b, 12.0, 330.0, 60.0, 25.0, hours=48, landfall=24)


def config2(seed=11):
    """Elliptic 250 x 300, ~35 % land, 400 barriers, spline track, 72 h at 30 s."""
    b = make_basin("ell", "$", 250, 300, seed=seed, land_frac=0.35, n_barriers=400, dt=30.0)
    return b, spline_track(b, seed=seed + 1)


def config3_basin(seed=21):
    """The ensemble basin: config-1 (polar) generator at 250 x 300 cells, 30 s step."""
    return make_basin("mow", " ", 250, 300, seed=seed, land_frac=0.20, n_barriers=400,
                      dt=30.0)


def meow_tracks(b: Basin, n=1024, hours=48, landfall=24):
    """MEOW-style ensemble: 8 headings x {5,10,15,25} kt x {30,50,70,90} mb x
    {15,30} mi x 4 parallel offsets (the first n of the 1024 combinations)."""
    out = []
    for off in (-24.0, -8.0, 8.0, 24.0):
        for rmw in (15.0, 30.0):
            for dp in (30.0, 50.0, 70.0, 90.0):
                for spd in (5.0, 10.0, 15.0, 25.0):
                    for hd in range(0, 360, 45):
                        out.append(straight_track(b, spd, float(hd), dp, rmw, hours=hours,
                                                  landfall=landfall, offset_mi=off))
    return out[:n]


def tiny(seed=5, gtype=" ", ni=24, nj=20, dt=60.0, hours=6):
    """Small basin for per-step parity tests (oracle finishes in milliseconds)."""
    b = make_basin("tny", gtype, ni, nj, seed=seed, land_frac=0.25,
                   n_barriers=max(4, ni * nj // 60), dt=dt, size_km=(1.0, 3.0), corr=4.0)
    return b, straight_track(b, 12.0, 330.0, 60.0, 15.0, hours=hours, landfall=hours // 2)


def add_manning(b: Basin, seed=13):
    """Make b a Ver2019-03 basin with a seeded Manning's n map (n x 1000, I3):
    land mostly 100-300 (slip classes 95-245), some below 100 (class 95)."""
    rng = np.random.default_rng(seed)
    m = rng.integers(100, 300, size=(b.jmxb, b.imxb))
    m[rng.random((b.jmxb, b.imxb)) < 0.2] = rng.integers(20, 99)
    b.extra["manning"] = m
    return b


# uneven ensemble: six basins of uneven size (name, grid type, ni, nj, barrier edges)
C4_BASINS = [("pla", " ", 80, 60, 100), ("plb", " ", 150, 120, 250), ("ela", "$", 200, 250, 400),
             ("elb", "$", 300, 300, 600), ("hya", "+", 350, 200, 500), ("elc", "$", 400, 350, 800)]


def config4_basins(seed=31, dt=30.0):
    """The six basins of the uneven ensemble: polar 80x60 and 150x120, elliptic
    200x250, 300x300 and 400x350, hyperbolic 350x200 (seeded depth fields)."""
    out = []
    for k, (nm, g, ni, nj, nbar) in enumerate(C4_BASINS):
        # 350 cells along I: axes of 260 x 174 mi keep the cells 0.55-1.6 mi and
        # the grid within 24-30 N (60 x 40 mi would grow the cells to 44 mi and
        # take the grid south of the equator, outside the Coriolis table)
        out.append(make_basin(nm, g, ni, nj, seed=seed + k, land_frac=0.25, n_barriers=nbar, dt=dt,
                              hyp_axes=(260.0, 174.0)))
    return out


def random_track(b: Basin, rng, hours_range=(24, 72)):
    """Uniformly drawn storm: heading 0-360, 5-25 kt, 20-100 mb, RMW 10-40 mi,
    run length in hours_range, landfall at mid-run, offset +-20 mi."""
    hours = int(rng.integers(hours_range[0], hours_range[1] + 1))
    return straight_track(b, float(rng.uniform(5, 25)), float(rng.uniform(0, 360)),
                          float(rng.uniform(20, 100)), float(rng.uniform(10, 40)), hours=hours,
                          landfall=hours // 2, offset_mi=float(rng.uniform(-20, 20)))


def forecast_track(b: Basin, rng, hours=72):
    """Forecast-style track: 6-hourly points (5-13 of them over the horizon) with
    deficit, RMW and speed varying point to point, interpolated hourly by the
    reference's linear interpolation (trk InterpolationType=Linear)."""
    npts = int(rng.integers(5, 14))
    step = max(6, int(math.ceil(hours / (npts - 1) / 6.0)) * 6)
    H = [k * step for k in range(npts)]
    while H[-1] < hours + 2:
        H.append(H[-1] + step)
    hd = math.radians(rng.uniform(0, 360))
    la, lo, dps, rms = [], [], [], []
    land = hours // 2
    x = y = 0.0
    pts = []
    for h in H:
        pts.append((x, y))
        spd = rng.uniform(5, 25) * MI_PER_KT
        hd += rng.uniform(-0.2, 0.2)
        x += spd * step * math.sin(hd)
        y += spd * step * math.cos(hd)
        dps.append(float(rng.uniform(20, 100)))
        rms.append(float(rng.uniform(10, 40)))
    # pass the basin's tangent point at the landfall hour
    kl = min(range(len(H)), key=lambda k: abs(H[k] - land))
    cx, cy = pts[kl]
    for (px, py) in pts:
        la.append(b.alto + (py - cy) / 69.05)
        lo.append(-b.alno + (px - cx) / (69.05 * math.cos(math.radians(b.alto))))
    return Track(H, la, lo, dps, rms, 0, hours, H[kl], interp="Linear")


# ---------------------------------------------------------------- tide inputs
# The 37 constituents in the order the reference's harmonic files list them
# ("Tides and Currents" order, tideutil.c:155-163) with their speeds in degrees
# per hour (tideutil.c:301-315); the synthetic amplitudes favour the
# semidiurnal and diurnal ones, as coastal tides do.
TIDE_NAMES = ["M2", "S2", "N2", "K1", "M4", "O1", "M6", "MK3", "S4", "MN4", "NU2", "S6", "MU2", "2N2", "OO1",
              "LAM2", "S1", "M1", "J1", "MM", "SSA", "SA", "MSF", "MF", "RHO", "Q1", "T2", "R2", "2Q1", "P1",
              "2SM2", "M3", "L2", "2MK3", "K2", "M8", "MS4"]
_TIDE_SCALE = {"M2": 1.6, "S2": 0.4, "N2": 0.35, "K1": 0.5, "O1": 0.45, "P1": 0.15, "K2": 0.1, "Q1": 0.08}


def ft03_text(seed=3, year0=2000, nyears=41) -> str:
    """Yearly node factors and equilibrium arguments in the layout LoadYearlyConstit
    reads (tideutil.c:82-283): five lines per year of 8, 8, 8, 8 and 5
    constituents (ft03 order), each line 'YYYY 0 L' then per constituent the node
    factor x 1000 (I4) and V+U x 10 (I4); the first line's year starts the file."""
    rng = np.random.default_rng(seed)
    out = []
    for y in range(year0, year0 + nyears):
        xode = rng.integers(850, 1150, size=37)
        vpu = rng.integers(0, 3600, size=37)
        k = 0
        for line, n in enumerate((8, 8, 8, 8, 5), start=1):
            s = "%4d 0 %1d" % (y, line)
            for _ in range(n):
                s += "%4d%4d" % (xode[k], vpu[k])
                k += 1
            out.append(s)
    return "\n".join(out) + "\n"


def bhc_bytes(b: Basin, abbrev: str, seed=4) -> bytes:
    """A harmonic-constant file as ReadHarmonicHeader/ReadHarmonicBody read it
    (tideutil.c:470-645): 'HarmConstit' + NUL, version 1 (int16), the basin
    abbreviation (5 bytes), the cell counts IMXB-1 and JMXB-1 (int16), 37 names
    of 5 bytes, then per cell (I fastest, tideInd, slosh2.c:492-518) longitude
    and latitude x 1e6 and per constituent amplitude (ft) and phase (deg)
    x 1e6, all little-endian int32.  Amplitudes vary smoothly over the basin."""
    import struct
    rng = np.random.default_rng(seed)
    ni, nj = b.imxb - 1, b.jmxb - 1
    head = b"HarmConstit\0" + struct.pack("<h", 1) + abbrev.encode()[:5].ljust(5, b"\0")
    head += struct.pack("<hh", ni, nj) + b"".join(n.encode().ljust(5, b"\0") for n in TIDE_NAMES)
    base_amp = np.array([_TIDE_SCALE.get(n, 0.02) for n in TIDE_NAMES])
    base_ph = rng.uniform(0, 360, size=37)
    gi = np.linspace(0, 1, ni)[None, :]
    gj = np.linspace(0, 1, nj)[:, None]
    cells = np.zeros((nj, ni, 2 + 74), dtype=np.int64)
    cells[:, :, 0] = int(-b.alno * 1e6)
    cells[:, :, 1] = int(b.alto * 1e6)
    for k in range(37):
        amp = base_amp[k] * (0.8 + 0.4 * gi + 0.1 * gj)
        ph = (base_ph[k] + 20.0 * gi - 10.0 * gj) % 360.0
        cells[:, :, 2 + 2 * k] = np.round(amp * 1e6)
        cells[:, :, 3 + 2 * k] = np.round(ph * 1e6)
    return head + cells.astype("<i4").tobytes()


def adj_text(b: Basin, seed=4) -> str:
    """A datum-adjustment file as ReadAdjDatum reads it (tideutil.c:647-734):
    the header line, then per square of the harmonic grid (I fastest) its
    longitude, latitude, I, J and the NAVD88 - MTL offset (ft)."""
    rng = np.random.default_rng(seed + 1000)
    ni, nj = b.imxb - 1, b.jmxb - 1
    lon, lat = int(-b.alno * 1e6) / 1e6, int(b.alto * 1e6) / 1e6
    adj = np.round(rng.uniform(-0.6, 1.2) + 0.3 * np.linspace(0, 1, ni)[None, :]
                   - 0.2 * np.linspace(0, 1, nj)[:, None], 3)
    rows = ["Longitude, Latitude, I, J, Adj"]
    for j in range(nj):
        for i in range(ni):
            rows.append("%.6f, %.6f, %d, %d, %.3f" % (lon, lat, i + 1, j + 1, adj[j, i]))
    return "\n".join(rows) + "\n"


def write_tide_db(root: str, b: Basin, abbrev: str, seed=4, database=2014, flavor=None, adj=False):
    """The tide inputs -f_tide reads under -rootDir (usrparse.c:432-436,
    slosh2.c:2173-2217): <root>/tidefile.ec<database>/ft03.dta and <abbrev>.bhc,
    with adj <abbrev>.adj (without it the driver takes the adjustment as 0),
    and with flavor the basin's line of tide_flavor.txt for -f_tide VDEF
    (slosh2.c:2098-2160)."""
    d = os.path.join(root, "tidefile.ec%4d" % database)
    os.makedirs(d, exist_ok=True)
    if adj:
        with open(os.path.join(d, abbrev + ".adj"), "w") as fh:
            fh.write(adj_text(b, seed))
    if flavor:
        with open(os.path.join(d, "tide_flavor.txt"), "w") as fh:
            fh.write("# basin:tide version\n%s:%s\n" % (abbrev.strip(), flavor))
    with open(os.path.join(d, "ft03.dta"), "w") as fh:
        fh.write(ft03_text(seed))
    with open(os.path.join(d, abbrev + ".bhc"), "wb") as fh:
        fh.write(bhc_bytes(b, abbrev, seed))
    return d
