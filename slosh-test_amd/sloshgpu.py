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
        L.slg_basin_close.argtypes = [vp]
        L.slg_basin_close.restype = None
        L.slg_batch_open.argtypes = [vp, i, ctypes.POINTER(cp), ctypes.POINTER(vp), cp, i]
        L.slg_batch_open_prec.argtypes = [vp, i, ctypes.POINTER(cp), i, ctypes.POINTER(vp), cp, i]
        L.slg_batch_precision.argtypes = [vp]
        L.slg_batch_nsteps.argtypes = [vp, i]
        L.slg_batch_upload.argtypes = [vp, cp, i]
        L.slg_batch_step.argtypes = [vp, i, cp, i]
        L.slg_batch_sync.argtypes = [vp, cp, i]
        L.slg_batch_step_timed.argtypes = [vp, i, vp, vp, cp, i]
        L.slg_batch_set_path.argtypes = [vp, i, i]
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

    @property
    def cells(self):
        return (self.imxb - 1) * (self.jmxb - 1)

    def close(self):
        if self.h:
            lib().slg_basin_close(self.h)
            self.h = ctypes.c_void_p()


class Batch:
    def __init__(self, basin: Basin, trks, precision: int = 32):
        L = lib()
        self.basin = basin
        self.n = len(trks)
        arr = (ctypes.c_char_p * self.n)(*[t.encode() for t in trks])
        self.h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        _chk(L.slg_batch_open_prec(basin.h, self.n, arr, precision, ctypes.byref(self.h), err, 512), err)
        self.precision = precision
        self.nsteps = [L.slg_batch_nsteps(self.h, s) for s in range(self.n)]

    def export(self, storm: int, init_path: str, rows_path: str, tables_path: str):
        L = lib()
        if L.slg_export_init(self.h, storm, init_path.encode()):
            raise SloshError("export_init failed")
        if L.slg_export_storm(self.h, storm, rows_path.encode(), tables_path.encode()):
            raise SloshError("export_storm failed")

    def set_path(self, path: int = PATH_FUSED, tile_rows: int = 16):
        if lib().slg_batch_set_path(self.h, path, tile_rows) != 0:
            raise SloshError("bad launch path %r / tile rows %r" % (path, tile_rows))

    def upload(self):
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_upload(self.h, err, 512), err)

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

    def stats(self, storm: int):
        out = np.zeros(8, dtype=np.int32)
        err = ctypes.create_string_buffer(512)
        _chk(lib().slg_batch_stats(self.h, storm, out.ctypes.data_as(ctypes.c_void_p), err, 512),
             err)
        return out

    def close(self):
        if self.h:
            lib().slg_batch_close(self.h)
            self.h = ctypes.c_void_p()
