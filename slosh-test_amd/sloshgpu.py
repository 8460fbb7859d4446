"""ctypes binding of libslosh_hip.so (include/slosh_hip.h) for tests and the bench.

The library is built in-tree (``make -C slosh-test_amd``); this module only loads
it and never falls back to anything else: a missing or unloadable library is an
error.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libslosh_hip.so")

_lib = None

# kernel slots of one step (SLG_K_* in include/slosh_hip.h) and their symbols
KERNELS = ["contty", "replay", "pflag", "hclass", "smooth", "hmxsv", "filter", "momntm", "stepB"]
PATH_FUSED, PATH_STAGED = 0, 1


class SloshError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SloshError("%s is not built (run make -C slosh-test_amd)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, cp, ip, i = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int
        L.slg_basin_open.argtypes = [cp, cp, ctypes.POINTER(vp), cp, i]
        L.slg_basin_set_ylt.argtypes = [vp, vp]
        L.slg_basin_dims.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]
        L.slg_basin_sides.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]
        L.slg_basin_close.argtypes = [vp]
        L.slg_basin_close.restype = None
        L.slg_batch_open.argtypes = [vp, i, ctypes.POINTER(cp), ctypes.POINTER(vp), cp, i]
        L.slg_batch_open_prec.argtypes = [vp, i, ctypes.POINTER(cp), i, ctypes.POINTER(vp), cp, i]
        L.slg_batch_open_opts.argtypes = [vp, i, ctypes.POINTER(cp), i, i, ctypes.POINTER(vp), cp, i]
        L.slg_batch_open_tracks.argtypes = [vp, i, vp, i, i, ctypes.POINTER(vp), cp, i]
        L.slg_batch_schedule.argtypes = [vp, i, i, vp, cp, i]
        L.slg_meow_seeds.argtypes = [vp, vp, i, vp, vp, vp, cp, i]
        L.slg_meow_tracks.argtypes = [vp, vp, i, vp, vp, cp, i]
        L.slg_batch_precision.argtypes = [vp]
        L.slg_batch_nsteps.argtypes = [vp, i]
        L.slg_batch_upload.argtypes = [vp, cp, i]
        L.slg_batch_step.argtypes = [vp, i, cp, i]
        L.slg_batch_sync.argtypes = [vp, cp, i]
        L.slg_batch_step_timed.argtypes = [vp, i, vp, vp, cp, i]
        L.slg_batch_set_path.argtypes = [vp, i, i]
        L.slg_batch_set_graph.argtypes = [vp, i]
        L.slg_batch_set_streams.argtypes = [vp, i]
        L.slg_batch_streams.argtypes = [vp]
        L.slg_batch_autotune.argtypes = [vp, i, ctypes.POINTER(i), ctypes.POINTER(i), vp, cp, i]
        L.slg_batch_envelopes.argtypes = [vp, vp, vp, cp, i]
        L.slg_track_datums.argtypes = [cp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.slg_track_env_comment.argtypes = [cp, cp, i]
        L.slg_track_datums_tide.argtypes = [cp, i, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.slg_batch_set_tide.argtypes = [vp, cp, cp, cp, i, i, cp, i]
        L.slg_batch_rex_frames.argtypes = [vp, cp, i, i, vp, ctypes.c_long, vp, cp, i]
        L.slg_batch_track.argtypes = [vp, i, i, vp]
        L.slg_basin_field.argtypes = [vp, cp, vp]
        L.slg_batch_set_maxstat.argtypes = [vp, i]
        L.slg_batch_maxstat_reset.argtypes = [vp]
        L.slg_batch_maxstat_env.argtypes = [vp, vp, cp, i]
        L.slg_batch_set_maxstat2.argtypes = [vp, i]
        L.slg_batch_maxstat2_reset.argtypes = [vp]
        L.slg_batch_maxstat2_env.argtypes = [vp, vp, cp, i]
        L.slg_batch_start_date.argtypes = [vp, i, vp]
        L.slg_rex_open.argtypes = [cp, cp, cp, i, i, ctypes.POINTER(vp)]
        L.slg_rex_frame.argtypes = [vp, vp, ctypes.c_double, vp, i]
        L.slg_rex_env.argtypes = [vp, cp, vp, i]
        L.slg_rex_close.argtypes = [vp]
        L.slg_rex_close.restype = None
        L.slg_batch_rex_points.argtypes = [vp, cp, i, i, i, vp, vp, cp, i]
        L.slg_pnt_open.argtypes = [cp, cp, cp, cp, i, vp, ctypes.POINTER(vp)]
        L.slg_pnt_row.argtypes = [vp, ctypes.c_double, vp]
        L.slg_pnt_close.argtypes = [vp]
        L.slg_pnt_close.restype = None
        L.slg_write_env.argtypes = [cp, i, i, cp, vp, ctypes.c_float, ctypes.c_float]
        L.slg_server_open.argtypes = [i, ctypes.c_double, ctypes.POINTER(vp), cp, i]
        L.slg_server_add_basin.argtypes = [vp, cp, cp, i, i, i, ctypes.POINTER(i), cp, i]
        L.slg_server_slots.argtypes = [vp, i]
        L.slg_server_submit.argtypes = [vp, i, cp, ctypes.POINTER(i), cp, i]
        L.slg_server_pump.argtypes = [vp, i, cp, i]
        L.slg_server_result.argtypes = [vp, i, vp, ctypes.POINTER(ctypes.c_double)]
        L.slg_server_stats.argtypes = [vp, vp]
        L.slg_server_run_error.argtypes = [vp, i, cp, i]
        L.slg_server_tiles.argtypes = [vp, vp, i]
        L.slg_server_set_order.argtypes = [vp, i]
        L.slg_batch_set_replay_limits.argtypes = [vp, i, ctypes.c_long]
        L.slg_batch_set_storm_group.argtypes = [vp, i]
        L.slg_batch_storm_group.argtypes = [vp]
        L.slg_server_close.argtypes = [vp]
        L.slg_server_close.restype = None
        L.slg_batch_current_step.argtypes = [vp]
        L.slg_batch_get.argtypes = [vp, i, cp, vp, cp, i]
        L.slg_batch_stats.argtypes = [vp, i, vp, cp, i]
        L.slg_batch_close.argtypes = [vp]
        L.slg_batch_close.restype = None
        L.slg_export_init.argtypes = [vp, i, cp]
        L.slg_export_storm.argtypes = [vp, i, cp, cp]
        _lib = L
    return _lib


def _chk(rc, err):
    if rc != 0:
        raise SloshError("slosh_hip error %d: %s" % (rc, err.value.decode(errors="replace")))


class Basin:
    def __init__(self, dta: str, bnt_line: str | None):
        L = lib()
        self.h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        _chk(L.slg_basin_open(dta.encode(), (bnt_line or "").encode(), ctypes.byref(self.h),
                              err, 512), err)
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L.slg_basin_dims(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        self.imxb, self.jmxb, self.nbcpt = a.value, b.value, c.value
        L.slg_basin_sides(self.h, None, None, ctypes.byref(c))
        self.nsides = c.value

    def field(self, name: str) -> np.ndarray:
        """Static field "ZB", "ZBM" or "YLT" as double [jmxb, imxb]."""
        out = np.zeros((self.jmxb, self.imxb))
        if lib().slg_basin_field(self.h, name.encode(), out.ctypes.data_as(ctypes.c_void_p)):
            raise SloshError("no basin field %s" % name)
        return out

    @property
    def cells(self):
        return (self.imxb - 1) * (self.jmxb - 1)

    def close(self):
        if self.h:
            lib().slg_basin_close(self.h)
            self.h = ctypes.c_void_p()


SCHED_HOST = 1
NPAR = 26


class TrackC(ctypes.Structure):
    """struct slg_track (include/slosh_hip.h): a trk file's contents in memory."""
    _dp = ctypes.POINTER(ctypes.c_double)
    _fields_ = [("year", ctypes.c_int), ("month", ctypes.c_int), ("day", ctypes.c_int),
                ("hour_of_day", ctypes.c_int), ("minute", ctypes.c_int), ("date_hour", ctypes.c_int),
                ("landfall_hour", ctypes.c_int), ("begin_hour", ctypes.c_int), ("end_hour", ctypes.c_int),
                ("ocean_datum", ctypes.c_double), ("lake_datum", ctypes.c_double),
                ("wind_factor", ctypes.c_double), ("rmw_factor", ctypes.c_double), ("spline", ctypes.c_int),
                ("npts", ctypes.c_int), ("hour", _dp), ("lat", _dp), ("lon", _dp), ("delp", _dp), ("rmw", _dp)]


def track_struct(t: dict):
    """A TrackC from a dict with the trk header fields (year, month, day,
    hour_of_day, minute, date_hour, landfall_hour, begin_hour, end_hour,
    ocean_datum, lake_datum, optional wind_factor / rmw_factor / spline) and row
    arrays hour, lat, lon (deg E), delp (mb), rmw (mi).  Returns (struct, the
    arrays it points into, which the caller keeps alive)."""
    rows = [np.ascontiguousarray(t[k], dtype=np.float64) for k in ("hour", "lat", "lon", "delp", "rmw")]
    dp = ctypes.POINTER(ctypes.c_double)
    c = TrackC(int(t["year"]), int(t["month"]), int(t["day"]), int(t["hour_of_day"]), int(t["minute"]),
               int(t["date_hour"]), int(t["landfall_hour"]), int(t["begin_hour"]), int(t["end_hour"]),
               float(t["ocean_datum"]), float(t["lake_datum"]), float(t.get("wind_factor", 1.0)),
               float(t.get("rmw_factor", 1.0)), int(t.get("spline", 0)), len(rows[0]),
               *[r.ctypes.data_as(dp) for r in rows])
    return c, rows


class MeowC(ctypes.Structure):
    """struct slg_meow (include/slosh_hip.h)."""
    _fields_ = [("xyflag", ctypes.c_int), ("lat", ctypes.c_double), ("lon", ctypes.c_double),
                ("dir", ctypes.c_double), ("apart", ctypes.c_double), ("r34", ctypes.c_double),
                ("num", ctypes.c_int), ("speed", ctypes.c_double), ("delp", ctypes.c_double),
                ("rmw", ctypes.c_double), ("landfall_hour", ctypes.c_int), ("begin_hour", ctypes.c_int),
                ("end_hour", ctypes.c_int), ("ocean_datum", ctypes.c_double), ("lake_datum", ctypes.c_double)]


MEOW_HOURS = 100


def meow(xyflag=0, lat=0.0, lon=0.0, dir=0.0, apart=10.0, r34=0.0, num=-1, speed=10.0, delp=50.0, rmw=20.0,
         landfall_hour=50, begin_hour=20, end_hour=70, ocean_datum=0.0, lake_datum=0.0):
    return MeowC(xyflag, lat, lon, dir, apart, r34, num, speed, delp, rmw, landfall_hour, begin_hour, end_hour,
                 ocean_datum, lake_datum)


def meow_seeds(basin, m: MeowC, cap=4096):
    """MEOW seed points (slg_meow_seeds): (index, lat, lon) arrays."""
    idx = np.zeros(cap, dtype=np.int32)
    la, lo = np.zeros(cap), np.zeros(cap)
    err = ctypes.create_string_buffer(512)
    n = lib().slg_meow_seeds(basin.h, ctypes.byref(m), cap, idx.ctypes.data_as(ctypes.c_void_p),
                             la.ctypes.data_as(ctypes.c_void_p), lo.ctypes.data_as(ctypes.c_void_p), err, 512)
    _chk(min(n, 0), err)
    return idx[:n], la[:n], lo[:n]


def meow_tracks(basin, m: MeowC, cap=4096):
    """MEOW tracks (slg_meow_tracks) as dicts Batch(tracks=...) takes."""
    tr = (TrackC * cap)()
    rows = np.zeros(cap * 5 * MEOW_HOURS)
    err = ctypes.create_string_buffer(512)
    n = lib().slg_meow_tracks(basin.h, ctypes.byref(m), cap, tr, rows.ctypes.data_as(ctypes.c_void_p), err, 512)
    _chk(min(n, 0), err)
    out = []
    for t in range(n):
        T = tr[t]
        r = rows[t * 5 * MEOW_HOURS:(t + 1) * 5 * MEOW_HOURS].reshape(5, MEOW_HOURS).copy()
        out.append(dict(year=T.year, month=T.month, day=T.day, hour_of_day=T.hour_of_day, minute=T.minute,
                        date_hour=T.date_hour, landfall_hour=T.landfall_hour, begin_hour=T.begin_hour,
                        end_hour=T.end_hour, ocean_datum=T.ocean_datum, lake_datum=T.lake_datum,
                        wind_factor=T.wind_factor, rmw_factor=T.rmw_factor, spline=T.spline,
                        hour=r[0], lat=r[1], lon=r[2], delp=r[3], rmw=r[4]))
    return out


class Batch:
    def __init__(self, basin: Basin, trks=None, precision: int = 32, sched_host: bool = False, tracks=None):
        """Runs on one basin from trk files (trks) or from tracks in memory
        (tracks: dicts as track_struct takes).  sched_host: the per-step storm
        values computed on the host (SLG_SCHED_HOST) instead of on the device."""
        L = lib()
        self.basin = basin
        self.h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        flags = SCHED_HOST if sched_host else 0
        if tracks is not None:
            self.n = len(tracks)
            held = [track_struct(t) for t in tracks]
            arr = (TrackC * self.n)(*[c for c, _ in held])
            _chk(L.slg_batch_open_tracks(basin.h, self.n, arr, precision, flags, ctypes.byref(self.h), err, 512),
                 err)
        else:
            self.n = len(trks)
            arr = (ctypes.c_char_p * self.n)(*[t.encode() for t in trks])
            _chk(L.slg_batch_open_opts(basin.h, self.n, arr, precision, flags, ctypes.byref(self.h), err, 512),
                 err)
        self.precision = precision
        self.nsteps = [L.slg_batch_nsteps(self.h, s) for s in range(self.n)]
        self._rexbuf = None

    def export(self, storm: int, init_path: str, rows_path: str, tables_path: str):
        L = lib()
        if L.slg_export_init(self.h, storm, init_path.encode()):
            raise SloshError("export_init failed")
        if L.slg_export_storm(self.h, storm, rows_path.encode(), tables_path.encode()):
            raise SloshError("export_storm failed")

    def set_path(self, path: int = PATH_FUSED, tile_rows: int = 16):
        if lib().slg_batch_set_path(self.h, path, tile_rows) != 0:
            raise SloshError("bad launch path %r / tile rows %r" % (path, tile_rows))

    def set_graph(self, steps_per_graph: int):
        if lib().slg_batch_set_graph(self.h, steps_per_graph) != 0:
            raise SloshError("bad graph length %r" % steps_per_graph)

    def set_streams(self, nstreams: int):
        """Step the batch in nstreams parts (1-4), each on its own stream."""
        if lib().slg_batch_set_streams(self.h, nstreams) != 0:
            raise SloshError("cannot step on %r streams (upload first; 1-4, at most one per storm)" % nstreams)

    @property
    def streams(self):
        return lib().slg_batch_streams(self.h)

    def autotune(self, steps_per_trial: int = 20):
        """Times fused 64x8/32x16/16x32 tiles and the staged path on real steps and
        keeps the fastest; returns (tile_rows, path, ms per step of each)."""
        rows, path = ctypes.c_int(), ctypes.c_int()
        ms = np.zeros(4)
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_autotune(self.h, steps_per_trial, ctypes.byref(rows), ctypes.byref(path),
                                      ms.ctypes.data_as(ctypes.c_void_p), err, 512), err)
        return rows.value, path.value, ms

    def set_tide(self, ft03: str, bhc: str, adj: str | None, mode: int, thresh: int = 0):
        """The gridded tide between steps (-f_tide V2 modes); before upload()."""
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_set_tide(self.h, ft03.encode(), bhc.encode(), adj.encode() if adj else None, mode,
                                      thresh, err, 512), err)

    def upload(self):
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_upload(self.h, err, 512), err)

    def schedule(self, storm: int, from_device: bool = True):
        """The per-step storm values of one run, [nsteps][NPAR] (double)."""
        out = np.zeros((self.nsteps[storm], NPAR))
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_schedule(self.h, storm, 1 if from_device else 0,
                                      out.ctypes.data_as(ctypes.c_void_p), err, 512), err)
        return out

    def set_storm_group(self, sg: int):
        """Storms whose tiles follow each other in the fused kernels' workgroup order."""
        if lib().slg_batch_set_storm_group(self.h, sg):
            raise SloshError("bad storm group %r (1-64, after upload)" % sg)

    @property
    def storm_group(self):
        return lib().slg_batch_storm_group(self.h)

    def set_replay_limits(self, evcap: int = 4096, budget: int = 0):
        """Lower the cascade replay's list capacity (it then sweeps every square)
        and visit budget (a run that exhausts it is flagged; its results refused)."""
        if lib().slg_batch_set_replay_limits(self.h, evcap, budget):
            raise SloshError("bad replay limits %r %r (upload first)" % (evcap, budget))

    def step(self, n: int):
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_step(self.h, n, err, 512), err)

    def step_timed(self, n: int):
        """n steps with HIP events around each kernel; returns (ms, launches) per
        kernel slot (KERNELS order)."""
        ms = np.zeros(len(KERNELS), dtype=np.float64)
        cnt = np.zeros(len(KERNELS), dtype=np.int32)
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_step_timed(self.h, n, ms.ctypes.data_as(ctypes.c_void_p),
                                        cnt.ctypes.data_as(ctypes.c_void_p), err, 512), err)
        return ms, cnt

    def sync(self):
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_sync(self.h, err, 512), err)

    @property
    def current_step(self):
        return lib().slg_batch_current_step(self.h)

    def get(self, storm: int, name: str) -> np.ndarray:
        out = np.zeros(self.basin.imxb * self.basin.jmxb, dtype=np.float64)
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_get(self.h, storm, name.encode(), out.ctypes.data_as(ctypes.c_void_p),
                                 err, 512), err)
        return out.reshape(self.basin.jmxb, self.basin.imxb)

    def flwsqr(self, storm: int) -> np.ndarray:
        """Transport across each sub-grid 1-D flow side of the basin."""
        out = np.zeros(max(self.basin.nsides, 1), dtype=np.float64)
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_get(self.h, storm, b"FLWSQR", out.ctypes.data_as(ctypes.c_void_p), err, 512),
             err)
        return out[:self.basin.nsides]

    def stats(self, storm: int):
        out = np.zeros(8, dtype=np.int32)
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_stats(self.h, storm, out.ctypes.data_as(ctypes.c_void_p), err, 512),
             err)
        return out

    def envelopes(self, want_env=True, want_mom=True):
        """Per-run envelopes (EnvSave int16 tenths, [nstorm, jmxb, imxb]) and
        their MOM ([jmxb, imxb]), computed on the device."""
        nj, ni = self.basin.jmxb, self.basin.imxb
        env = np.zeros((self.n, nj, ni), dtype=np.int16) if want_env else None
        mom = np.zeros((nj, ni), dtype=np.int16) if want_mom else None
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_envelopes(self.h, env.ctypes.data_as(ctypes.c_void_p) if want_env else None,
                                       mom.ctypes.data_as(ctypes.c_void_p) if want_mom else None,
                                       err, 512), err)
        return env, mom

    def track(self, storm: int, n: int):
        """What TMSTPM returns after step n (batches of up to 8 runs): ITIME,
        MHALT, DELT, storm lat, storm lon, DELP, size, speed, direction."""
        out = np.zeros(9)
        if lib().slg_batch_track(self.h, storm, n, out.ctypes.data_as(ctypes.c_void_p)):
            raise SloshError("no track values for storm %d step %d" % (storm, n))
        return out

    def start_date(self, storm: int):
        """year, month, day, hour, minute of the run's start (INITAL)."""
        out = np.zeros(5, dtype=np.int32)
        if lib().slg_batch_start_date(self.h, storm, out.ctypes.data_as(ctypes.c_void_p)):
            raise SloshError("no start date for storm %d" % storm)
        return [int(v) for v in out]

    def set_maxstat(self, on=True):
        """Keep the P-Surge windowed maxima of HB (the -f_stat mode)."""
        if lib().slg_batch_set_maxstat(self.h, 1 if on else 0):
            raise SloshError("cannot switch the windowed maxima (upload first)")

    def maxstat_reset(self):
        if lib().slg_batch_maxstat_reset(self.h):
            raise SloshError("windowed maxima are off")

    def maxstat_env(self):
        """EnvSave values of every run's window maxima, [nstorm, jmxb, imxb]."""
        env = np.zeros((self.n, self.basin.jmxb, self.basin.imxb), dtype=np.int16)
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_maxstat_env(self.h, env.ctypes.data_as(ctypes.c_void_p), err, 512), err)
        return env

    def set_maxstat2(self, on=True):
        """Keep the second window of maxima (the -envSave2 mode)."""
        if lib().slg_batch_set_maxstat2(self.h, 1 if on else 0):
            raise SloshError("cannot switch the second window (upload first)")

    def maxstat2_reset(self):
        if lib().slg_batch_maxstat2_reset(self.h):
            raise SloshError("the second window is off")

    def maxstat2_env(self):
        """EnvSave values of every run's second-window maxima, [nstorm, jmxb, imxb]."""
        env = np.zeros((self.n, self.basin.jmxb, self.basin.imxb), dtype=np.int16)
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_maxstat2_env(self.h, env.ctypes.data_as(ctypes.c_void_p), err, 512), err)
        return env

    def rex_frames(self, field="HHB", s0=0, ns=None):
        """rex v2 frame bit streams of storms s0..s0+ns-1, encoded on the
        device from HHB = min(HB, HMX) of the last step or from HMX; a list of
        bytes objects."""
        ns = self.n - s0 if ns is None else ns
        nj, ni = self.basin.jmxb, self.basin.imxb
        cap = (18 * (ni - 1) * (nj - 1) + 7) // 8
        # one staging buffer per batch, kept between calls (a frame of 18 bits
        # a cell is the longest possible; fresh pages of that size each call
        # cost more than the encoding)
        if self._rexbuf is None or self._rexbuf.size < ns * cap:
            self._rexbuf = np.empty(ns * cap, dtype=np.uint8)
        buf = self._rexbuf
        nb = np.zeros(ns, dtype=np.int32)
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_rex_frames(self.h, field.encode(), s0, ns, buf.ctypes.data_as(ctypes.c_void_p), cap,
                                        nb.ctypes.data_as(ctypes.c_void_p), err, 512), err)
        return [bytes(buf[k * cap:k * cap + nb[k]]) for k in range(ns)]

    def rex_points(self, xy, field="HHB", s0=0, ns=None):
        """rex values (tenths of a foot, 999 dry) at squares xy [(I, J), ...]
        (1-based) of storms s0..s0+ns-1, computed on the device from the same
        field as rex_frames; int16 [ns, npts]."""
        ns = self.n - s0 if ns is None else ns
        pts = np.ascontiguousarray(np.asarray(xy, dtype=np.int32).reshape(-1, 2))
        out = np.zeros((ns, len(pts)), dtype=np.int16)
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_rex_points(self.h, field.encode(), s0, ns, len(pts), pts.ctypes.data_as(ctypes.c_void_p),
                                        out.ctypes.data_as(ctypes.c_void_p), err, 512), err)
        return out

    def close(self):
        self._rexbuf = None
        if self.h:
            lib().slg_batch_close(self.h)
            self.h = ctypes.c_void_p()


class PointSeries:
    """rexout style-0 text of a point series (slg_pnt_*): the header, then one
    row per frame time (model clock, seconds since 1970) with the values of
    Batch.rex_points for one storm."""

    def __init__(self, path: str, abbrev: str, storm: str, rexfile: str, xy):
        self.pts = np.ascontiguousarray(np.asarray(xy, dtype=np.int32).reshape(-1, 2))
        self.h = ctypes.c_void_p()
        if lib().slg_pnt_open(path.encode(), abbrev.encode(), storm.encode(), rexfile.encode(), len(self.pts),
                              self.pts.ctypes.data_as(ctypes.c_void_p), ctypes.byref(self.h)):
            raise SloshError("cannot open point series %s" % path)

    def row(self, clock: float, vals):
        v = np.ascontiguousarray(vals, dtype=np.int16)
        if v.size != len(self.pts) or lib().slg_pnt_row(self.h, clock, v.ctypes.data_as(ctypes.c_void_p)):
            raise SloshError("point series row failed")

    def close(self):
        if self.h:
            lib().slg_pnt_close(self.h)
            self.h = ctypes.c_void_p()


def track_datums(trk: str, tide_mode: int = 0):
    """The env / rex trailer datums of a run from its trk file; tide_mode: those
    of a run with the gridded tide in that mode (2, 21, 22, 24)."""
    a, b = ctypes.c_float(), ctypes.c_float()
    if tide_mode:
        rc = lib().slg_track_datums_tide(trk.encode(), tide_mode, ctypes.byref(a), ctypes.byref(b))
    else:
        rc = lib().slg_track_datums(trk.encode(), ctypes.byref(a), ctypes.byref(b))
    if rc:
        raise SloshError("cannot read %s" % trk)
    return a.value, b.value


def track_env_comment(trk: str) -> str:
    """The comment the reference's EnvSave writes for a run of this trk file."""
    b = ctypes.create_string_buffer(256)
    if lib().slg_track_env_comment(trk.encode(), b, 256):
        raise SloshError("cannot read %s" % trk)
    return b.value.decode("latin-1")


def parse_tide(flag: str):
    """(mode, thresh) of an -f_tide V2 value as ParseTide reads it
    (usrparse.c:14-52): V2, V2.1[.x], V2.2[.x], V2.4[.x]."""
    if flag == "V2":
        return 2, 0
    for head, mode in (("V2.1", 21), ("V2.2", 22), ("V2.4", 24)):
        if flag.startswith(head):
            return mode, (int(flag[5:]) if len(flag) > 4 else 0)
    raise SloshError("not a gridded-tide V2 mode: %r" % flag)


def write_env(path: str, vals: np.ndarray, ht1: float, ht2: float, comment: str = "FileVersion=2013"):
    """env file in the reference's EnvSave layout; vals int16 [jmxb, imxb]."""
    v = np.ascontiguousarray(vals, dtype=np.int16)
    if lib().slg_write_env(path.encode(), v.shape[1], v.shape[0], comment.encode(),
                           v.ctypes.data_as(ctypes.c_void_p), ht1, ht2):
        raise SloshError("cannot write %s" % path)


class Rex:
    """A rex v2 file written as the reference's rex writer lays it out; the
    frames' bit streams come from Batch.rex_frames."""

    def __init__(self, path: str, comment: str, abbrev: str, imxb: int, jmxb: int):
        h = ctypes.c_void_p()
        if lib().slg_rex_open(path.encode(), comment.encode(), abbrev.encode(), imxb, jmxb, ctypes.byref(h)):
            raise SloshError("cannot open rex file %s" % path)
        self.h = h

    def frame(self, track6, clock: float, bits: bytes):
        t = np.ascontiguousarray(track6, dtype=np.float64)
        b = np.frombuffer(bits, dtype=np.uint8) if bits else np.zeros(1, np.uint8)
        if lib().slg_rex_frame(self.h, t.ctypes.data_as(ctypes.c_void_p), clock, b.ctypes.data_as(ctypes.c_void_p),
                               len(bits)):
            raise SloshError("rex frame write failed")

    def env(self, trk: str, bits: bytes):
        b = np.frombuffer(bits, dtype=np.uint8) if bits else np.zeros(1, np.uint8)
        if lib().slg_rex_env(self.h, trk.encode(), b.ctypes.data_as(ctypes.c_void_p), len(bits)):
            raise SloshError("rex envelope write failed (track file %s)" % trk)

    def close(self):
        if self.h:
            lib().slg_rex_close(self.h)
            self.h = None


class Server:
    """Storm runs on several basins, admitted as they arrive (continuous batching)."""

    def __init__(self, precision: int = 32, hbm_fraction: float = 0.5):
        self.h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_server_open(precision, hbm_fraction, ctypes.byref(self.h), err, 512), err)
        self.dims = {}

    def add_basin(self, dta: str, bnt_line: str, max_slots: int, max_steps: int, tables_per_slot: int = 64):
        bid = ctypes.c_int()
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_server_add_basin(self.h, dta.encode(), bnt_line.encode(), max_slots, max_steps,
                                        tables_per_slot, ctypes.byref(bid), err, 512), err)
        B = Basin(dta, bnt_line)
        self.dims[bid.value] = (B.jmxb, B.imxb, B.cells)
        B.close()
        return bid.value

    def slots(self, basin_id: int) -> int:
        return lib().slg_server_slots(self.h, basin_id)

    def set_order(self, hardest_first: bool):
        """Admit queued runs in submission order, or the most work first."""
        if lib().slg_server_set_order(self.h, 1 if hardest_first else 0):
            raise SloshError("bad admission order")

    def submit(self, basin_id: int, trk: str) -> int:
        rid = ctypes.c_int()
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_server_submit(self.h, basin_id, trk.encode(), ctypes.byref(rid), err, 512), err)
        return rid.value

    def pump(self, k: int) -> int:
        err = ctypes.create_string_buffer(512)
        rc = lib().slg_server_pump(self.h, k, err, 512)
        if rc < 0:
            _chk(rc, err)
        return rc

    def result(self, run_id: int, basin_id: int):
        """(state, env or None, latency or None); state 0 queued, 1 running, 2 done."""
        nj, ni, _ = self.dims[basin_id]
        env = np.zeros((nj, ni), dtype=np.int16)
        lat = ctypes.c_double()
        st = lib().slg_server_result(self.h, run_id, env.ctypes.data_as(ctypes.c_void_p), ctypes.byref(lat))
        if st == 3:
            err = ctypes.create_string_buffer(512)
            lib().slg_server_run_error(self.h, run_id, err, 512)
            raise SloshError("run %d failed: %s" % (run_id, err.value.decode(errors="replace")))
        return (st, env, lat.value) if st == 2 else (st, None, None)

    def tiles(self):
        """Tile rows each pool steps with (the pool's tuner choice once measured)."""
        out = np.zeros(64, dtype=np.int32)
        n = lib().slg_server_tiles(self.h, out.ctypes.data_as(ctypes.c_void_p), 64)
        return [int(x) for x in out[:n]]

    def stats(self):
        out = np.zeros(6, dtype=np.int64)
        lib().slg_server_stats(self.h, out.ctypes.data_as(ctypes.c_void_p))
        return dict(queued=int(out[0]), running=int(out[1]), done=int(out[2]), slots=int(out[3]),
                    bytes=int(out[4]), cell_steps=int(out[5]))

    def close(self):
        if self.h:
            lib().slg_server_close(self.h)
            self.h = ctypes.c_void_p()
