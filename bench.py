#!/usr/bin/env python3
"""Throughput of the SLOSH time-stepping core on one MI355X.

One "step" is one model time step (TMSTPC1 + TMSTPC2 + TMSTPM of the reference)
applied to every storm of the batch.  The headline workload is the batched MEOW
ensemble: 1024 storms (8 headings x 4 speeds x 4 pressure deficits x 2 RMW x 4
offsets) on one synthetic polar basin of 250 x 300 cells at dt = 30 s (48 h
runs, 5760 steps), all storms advancing together, fields resident in HBM, in
FP64: the reference's default precision, whose results the FP64 step
reproduces bit for bit (so its envelope error against the reference is 0).
The FP32 step is timed beside it, with its measured distance from the
reference.  Also measured:

  kernels  the headline batch again at one stream with HIP events around every
           kernel, over the same timed steps: per-kernel durations and their
           fractions of the HBM roofline;
  plain    one storm on the 250 x 300 elliptic basin (72 h spline track), the
           single-storm GPU path (fused kernels; and as hipGraph replays);
  uneven   4096 storms over six basins of 4 800 - 140 000 cells (polar,
           elliptic, hyperbolic), 24-72 h runs, one batch per basin on its own
           stream, all advancing together to the end of the longest run;
  serve    Poisson arrivals of 72 h forecast-track runs over the six basins,
           admitted into device slots as they arrive; p50/p99 latency per run;
  cpu      the reference's own FP64 build on this host's cores, one serial
           process per storm (its multi-process run pattern), over whole runs;
           and, since that run is the reference's own answer, the envelope of
           the polar 100x90 two-day run against the GPU's at both precisions.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

The K timed steps are spread over the whole run: when K is smaller than what
is left after the autotuner's trials and the warm-up, they are taken in up to
eight windows from there to the end (the steps between windows run untimed),
so the post-landfall steps, where the drying cascade costs most, are in the
sample.  Prints one JSON line.  Inputs are synthetic and seeded; rates are
storm cell steps per second (storm.cell.steps/s).
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "slosh-test_amd"))

import numpy as np  # noqa: E402

import synth  # noqa: E402
import sloshgpu  # noqa: E402

HBM_PEAK_GBS = 8000.0
TUNE_STEPS = 20          # autotuner: 4 candidates x 20 steps
# Algorithmic HBM bytes per storm.cell of one launch of each fused-path kernel
# in FP32 (FP64 doubles every real; the byte mask stays one byte).  Each element
# is read or written once by the launch; the basin's static fields are shared
# by every storm of the batch and are not counted per storm.
ALG_BYTES = {
    "contty": (16, 1),   # HB, UB, VB in; HB out (reals); U==0 mask out (byte)
    "replay": (0, 0),    # serial cascade: touches only the drying squares
    "stepB": (32, 1),    # HB in, HMX in/out, UB/VB in/out, HB out (reals); mask in (byte)
}
PROFILES = os.path.join(HERE, "profiles")


def alg_bytes(kernel, prec):
    reals, byts = ALG_BYTES[kernel]
    return reals * (2 if prec == 64 else 1) + byts


def b_alg_step(storms, prec):
    """Whole-step algorithmic bytes per storm.cell.step: HB, UB, VB, HMX each read
    and written once (8 reals) plus the static ZB, ZBM, Coriolis factor (3 reals)
    and a u32 of flags once per basin cell, shared by the storms of the batch."""
    rs = 8 if prec == 64 else 4
    return 8.0 * rs + (3.0 * rs + 4.0) / storms


def measured_traffic(kernel, prec, storms):
    """HBM bytes per launch of `kernel` of the headline batch at one stream,
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (copy-calibrated); None if there is no measurement for this configuration."""
    path = os.path.join(PROFILES, "meow1024_fp%d_hbm_traffic.md" % prec)
    if not os.path.exists(path):
        return None
    key = "traffic_bytes_per_launch %s storms=%d" % (kernel, storms)
    with open(path) as fh:
        for line in fh:
            if line.startswith(key):
                return float(line.split()[-1])
    return None


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def write_tracks(root, tracks, prefix):
    out = []
    for k, t in enumerate(tracks):
        p = os.path.join(root, "%s%04d.trk" % (prefix, k))
        with open(p, "w") as fh:
            fh.write(synth.trk_text(t))
        out.append(p)
    return out


def windows(start, end, k):
    """k timed steps between step start and end: one window when they fit, else
    up to eight windows spread evenly from start to end."""
    if k <= 0 or start >= end:
        return []
    if k >= end - start:
        return [(start, end)]
    n = min(8, k)
    sizes = [k // n + (1 if w < k % n else 0) for w in range(n)]
    free = (end - start) - k
    out, pos = [], start
    for w in range(n):
        gap = 0 if w == 0 else (free * w) // (n - 1) - (free * (w - 1)) // (n - 1)
        pos += gap
        out.append((pos, pos + sizes[w]))
        pos += sizes[w]
    return out


def advance_to(X, step):
    n = step - X.current_step
    if n > 0:
        X.step(n)


# ------------------------------------------------------------------ headline
def build_ensemble(root, nstorm):
    b = synth.config3_basin()
    tracks = synth.meow_tracks(b, n=nstorm)
    paths = synth.write_case(root, b, tracks[0], trk_name="s0000.trk")
    trks = [paths["trk"]] + write_tracks(root, tracks[1:], "s1_")
    return b, paths, trks


def batched_run(dta, bnt, trks, prec, steps, warmup, streams, events, path=None, rows=32, span=0, tide=None):
    """Open and upload the batch, autotune (or fix the path), warm up, then time
    `steps` steps in windows spread to the end of the run.  With events the
    timed steps carry HIP events around every kernel on the batch's stream
    (one stream), which give the per-kernel durations of those steps.  tide:
    (ft03, bhc, adj, mode, thresh) of the gridded tide between steps."""
    B = sloshgpu.Basin(dta, bnt)
    t0 = time.time()
    X = sloshgpu.Batch(B, trks, precision=prec)
    t_sched = time.time() - t0
    t_tide = None
    if tide is not None:
        t0 = time.time()
        X.set_tide(*tide)
        t_tide = time.time() - t0
    X.upload()
    tune = None
    if path is None:   # the online autotuner picks the launch configuration
        r, p, ms = X.autotune(TUNE_STEPS)
        tune = {"tile_rows": r, "path": "fused" if p == 0 else "staged",
                "ms_per_step": {"fused8": ms[0], "fused16": ms[1], "fused32": ms[2], "staged": ms[3]}}
        fused = p == sloshgpu.PATH_FUSED
    else:
        X.set_path(path, rows)
        fused = path == sloshgpu.PATH_FUSED
    if streams > 1 and X.n > 1 and fused and not events:
        X.set_streams(streams)
    X.step(warmup)
    X.sync()
    start, end = X.current_step, min(X.nsteps)
    if span > 0:
        end = min(end, start + span)
    wins = windows(start, end, steps)
    if sum(b - a for a, b in wins) != steps:
        raise SystemExit("bench.py: %d timed steps asked, the run has %d left after tuning and warm-up"
                         % (steps, end - start))
    dt = 0.0
    ms = np.zeros(len(sloshgpu.KERNELS))
    cnt = np.zeros(len(sloshgpu.KERNELS), dtype=np.int64)
    for a, b in wins:
        advance_to(X, a)
        X.sync()
        t0 = time.perf_counter()
        if events:
            m, c = X.step_timed(b - a)
            ms += m
            cnt += c
        else:
            X.step(b - a)
            X.sync()
        dt += time.perf_counter() - t0
    nstreams = X.streams
    X.close()
    cells = B.cells
    B.close()
    return dict(dt=dt, cells=cells, ms=ms, cnt=cnt, sched_s=t_sched, tune=tune, streams=nstreams, windows=wins,
                tide_s=t_tide)


def setup_bench(dta, bnt, trks, tracks, prec):
    """Open and upload of the ensemble three ways: the per-step storm values
    and the initial water computed on the host and uploaded (SLG_SCHED_HOST);
    generated on the device from the storm programs; and the same from tracks
    handed over in memory instead of trk files.  The basin's static fields are
    uploaded beforehand (untimed), so each line times the runs alone."""
    B = sloshgpu.Basin(dta, bnt)
    W = sloshgpu.Batch(B, trks[:1], precision=prec)
    W.upload()
    W.close()
    out = {}
    for name, kw in (("host", dict(trks=trks, sched_host=True)), ("device", dict(trks=trks)),
                     ("device_memory_tracks", dict(tracks=tracks))):
        t0 = time.perf_counter()
        X = sloshgpu.Batch(B, precision=prec, **kw)
        t1 = time.perf_counter()
        X.upload()
        X.sync()
        t2 = time.perf_counter()
        out[name] = {"open_s": t1 - t0, "upload_s": t2 - t1, "total_s": t2 - t0}
        X.close()
        log("setup %s: open %.2f s, upload %.2f s" % (name, t1 - t0, t2 - t1))
    B.close()
    out["storms"] = len(trks)
    return out


def kernel_roofline(r, prec, storms):
    """Per-kernel average launch time (HIP events on the batch's one stream) and
    achieved algorithmic GB/s against the HBM peak."""
    out = {}
    for i, k in enumerate(sloshgpu.KERNELS):
        if not r["cnt"][i] or k not in ALG_BYTES:
            continue
        avg = r["ms"][i] / r["cnt"][i]
        by = alg_bytes(k, prec) * storms * r["cells"]
        ach = by / (avg * 1e-3) / 1e9 if by else 0.0
        out[k] = {"avg_launch_ms": avg, "launches": int(r["cnt"][i]), "alg_bytes_per_launch": by,
                  "achieved_GBs": ach, "frac": ach / HBM_PEAK_GBS,
                  "traffic": measured_traffic(k, prec, storms)}
    return out


# ----------------------------------------------------------------------- rex
def rex_frames_bench(dta, bnt, trks, prec, reps=5):
    """rex v2 frames of every storm of the ensemble encoded on the device
    (slg_batch_rex_frames: HHB of the last step, run-length coded, copied back)
    after 20 staged steps; time per frame set, over reps calls after one
    untimed call."""
    B = sloshgpu.Basin(dta, bnt)
    X = sloshgpu.Batch(B, trks, precision=prec)
    X.upload()
    X.set_path(sloshgpu.PATH_STAGED, 16)
    X.step(20)
    X.sync()
    fr = X.rex_frames("HHB")
    t0 = time.perf_counter()
    for _ in range(reps):
        fr = X.rex_frames("HHB")
    dt = (time.perf_counter() - t0) / reps
    n = X.n
    cells = (B.imxb - 1) * (B.jmxb - 1)
    X.close()
    B.close()
    nbytes = sum(len(f) for f in fr)
    return {"storms": n, "ms_per_frame_set": dt * 1e3, "frames_per_s": n / dt,
            "bytes_out_per_frame": nbytes / n, "cells_per_frame": cells,
            "note": "HHB = min(HB, HMX) per cell to rex v2 run-length records, copied to the host; "
                    "includes the host copy"}


# -------------------------------------------------------------------- uneven
def uneven(root, nstorm, prec, graph=20):
    """nstorm storms over the six basins (share proportional to cells), one
    batch per basin on its own stream; each batch's launch shape picked by the
    online autotuner and its step loop replayed as hipGraphs of `graph` steps;
    the batches advance together, `graph` steps at a time, to the end of every
    run.  Timed from after the tuning to the end (runs that end early stop
    costing anything)."""
    basins = synth.config4_basins()
    cells = np.array([(b.imxb - 1) * (b.jmxb - 1) for b in basins], dtype=float)
    share = np.floor(nstorm * cells / cells.sum()).astype(int)
    share[-1] += nstorm - share.sum()
    rng = np.random.default_rng(4096)
    runs, tunes = [], []
    t0 = time.time()
    for b, n in zip(basins, share):
        d = os.path.join(root, "c4_" + b.name)
        tracks = [synth.random_track(b, rng) for _ in range(n)]
        case = synth.write_case(d, b, tracks[0])
        trks = [case["trk"]] + write_tracks(d, tracks[1:], "t")
        B = sloshgpu.Basin(case["dta"], case["bntline"])
        X = sloshgpu.Batch(B, trks, precision=prec)
        X.upload()
        r, p, ms = X.autotune(10)
        tunes.append({"tile_rows": r, "path": "fused" if p == 0 else "staged"})
        if graph and p == sloshgpu.PATH_FUSED:
            X.set_graph(graph)
        runs.append((B, X, n))
    setup = time.time() - t0
    for B, X, n in runs:
        X.sync()
    # work still ahead: every storm from the step the tuner left it at to its end
    work = 0
    for B, X, n in runs:
        work += sum(max(0, s - X.current_step) for s in X.nsteps) * B.cells
    chunk = max(1, graph)
    t0 = time.perf_counter()
    rounds = 0
    while True:
        active = [(B, X) for B, X, n in runs if X.current_step < max(X.nsteps)]
        if not active:
            break
        for B, X in active:
            X.step(min(chunk, max(X.nsteps) - X.current_step))
        rounds += 1
        if rounds % 50 == 0:   # a progress line every 1000 steps (the host runs ahead by at most that)
            active[0][1].sync()
            log("uneven: step %d, %d of 6 basins running, %.1f s" % (rounds * chunk, len(active),
                                                                     time.perf_counter() - t0))
    for B, X, n in runs:
        X.sync()
    dt = time.perf_counter() - t0
    out = {"value": work / dt, "unit": "storm.cell.steps/s", "storms": int(nstorm), "precision": prec,
           "timed": "whole runs after the tuner's 40 steps", "wall_s": dt, "setup_s": setup,
           "graph_steps": graph,
           "basins": [{"name": b.name, "grid": {" ": "polar", "$": "elliptic", "+": "hyperbolic"}[b.gtype],
                       "cells": int(c), "storms": int(n), "tuned": t}
                      for b, c, n, t in zip(basins, cells, share, tunes)]}
    for B, X, n in runs:
        X.close()
        B.close()
    return out


# --------------------------------------------------------------------- serve
def plain_ms_per_step(root, basins, prec, steps=8600):
    """Per-step time of the plain single-storm GPU path (fused kernels, one
    storm) on each basin, over a whole 72 h forecast-track run (8640 steps)."""
    rng = np.random.default_rng(77)
    out = []
    for b in basins:
        d = os.path.join(root, "pl_" + b.name)
        case = synth.write_case(d, b, synth.forecast_track(b, rng, hours=72))
        r = batched_run(case["dta"], case["bntline"], [case["trk"]], prec, steps, 20, 1, False,
                        sloshgpu.PATH_FUSED, 32)
        out.append(r["dt"] * 1e3 / steps)
    return out


def fifo_latencies(arrivals, service):
    """Latency of each request through one FIFO server (the plain path serves
    one storm run at a time): start = max(arrival, previous finish)."""
    fin, lat = 0.0, []
    for t, s in zip(arrivals, service):
        fin = max(fin, t) + s
        lat.append(fin - t)
    return np.array(lat)


def serve(root, rate, seconds, hours, slots, pump_steps, prec, plain_ms=None, seed=55):
    """Poisson arrivals of forecast-track runs (basin chosen in proportion to its
    cells) admitted into device slots as they arrive; per-run latency from
    submission to envelope.  Offered load = the arrivals' work per second of
    the arrival window; served = work done per second of wall time until the
    last run's envelope is out."""
    basins = synth.config4_basins()
    cells = np.array([(b.imxb - 1) * (b.jmxb - 1) for b in basins], dtype=float)
    rng = np.random.default_rng(seed)
    S = sloshgpu.Server(precision=prec, hbm_fraction=0.5)
    ids, cases = [], []
    maxsteps = int(hours * 3600 / 30) + 400
    for b in basins:
        d = os.path.join(root, "c5_%g_%s" % (rate, b.name))
        case = synth.write_case(d, b, synth.forecast_track(b, rng, hours=hours))
        ids.append(S.add_basin(case["dta"], case["bntline"], max_slots=slots, max_steps=maxsteps,
                               tables_per_slot=160))
        cases.append(d)
    n = int(rng.poisson(rate * seconds))
    times = np.sort(rng.uniform(0, seconds, n))
    which = rng.choice(len(basins), size=n, p=cells / cells.sum())
    trks = []
    for k in range(n):
        t = synth.forecast_track(basins[which[k]], rng, hours=hours)
        p = os.path.join(cases[which[k]], "a%05d.trk" % k)
        with open(p, "w") as fh:
            fh.write(synth.trk_text(t))
        trks.append(p)
    submitted = []
    t0 = time.perf_counter()
    nxt, last, pumps, t_pump = 0, -1, 0, 0.0
    done_at_window = None
    while True:
        now = time.perf_counter() - t0
        while nxt < n and times[nxt] <= now:
            submitted.append((S.submit(ids[which[nxt]], trks[nxt]), ids[which[nxt]]))
            nxt += 1
        tp = time.perf_counter()
        S.pump(pump_steps)
        t_pump += time.perf_counter() - tp
        pumps += 1
        st = S.stats()
        if done_at_window is None and now >= seconds:
            done_at_window = (st["done"], st["cell_steps"])
        if st["done"] // 50 != last:
            last = st["done"] // 50
            log("serve %g/s: t=%.1fs submitted %d done %d running %d" % (rate, now, nxt, st["done"], st["running"]))
        if nxt == n and st["done"] >= n:
            break
        if now > seconds + 900:
            raise RuntimeError("serving did not drain: %r" % st)
        if st["running"] == 0 and st["queued"] == 0 and nxt < n:
            time.sleep(max(0.0, min(0.005, times[nxt] - now)))
    wall = time.perf_counter() - t0
    lat = np.array([S.result(r, b)[2] for r, b in submitted])
    st = S.stats()
    slots_used = [S.slots(i) for i in ids]
    tiles = S.tiles()
    S.close()
    offered = st["cell_steps"] / seconds
    served = st["cell_steps"] / wall
    out = {"rate_per_s": rate, "precision": prec, "arrival_window_s": seconds, "runs": n, "horizon_h": hours,
           "slots_per_basin": slots_used, "tile_rows_per_basin": tiles, "pump_steps": pump_steps, "pumps": pumps,
           "host_s_in_pump": t_pump,
           "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
           "mean_latency_s": float(lat.mean()), "wall_s": wall,
           "offered": offered, "value": served, "unit": "storm.cell.steps/s", "pool_bytes": st["bytes"],
           "over_capacity": bool(wall > 1.5 * seconds)}
    if plain_ms is not None:
        # the same arrivals served one run at a time by the plain single-storm path
        # (per-run time = its steps x the plain path's measured step time on its basin)
        svc = [(hours * 120) * plain_ms[w] * 1e-3 for w in which]
        pl = fifo_latencies(times, svc)
        out["plain_fifo"] = {"p50_latency_s": float(np.percentile(pl, 50)),
                             "p99_latency_s": float(np.percentile(pl, 99)),
                             "service_s_per_run_by_basin": [hours * 120 * m * 1e-3 for m in plain_ms],
                             "utilisation": float(sum(svc) / seconds)}
    return out


# ----------------------------------------------------------------------- cpu
def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return None


def cpu_baseline(root, dta, bnt_path, trks, nsteps, procs):
    """The reference FP64 build (oracle/_ref/refdrv64), one serial process per
    storm as the reference's multi-process runner does, on `procs` cores; each
    process runs one whole storm; the init-only time is subtracted."""
    exe = os.path.join(HERE, "oracle", "_ref", "refdrv64")
    if not os.path.exists(exe):
        return None

    def launch(n):
        ps = []
        for k in range(procs):
            out = tempfile.mkdtemp(prefix="cpu%02d_" % k, dir=root)
            ps.append(subprocess.Popen([exe, bnt_path, dta, trks[k], out, str(n), ""],
                                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
        t0 = time.perf_counter()
        for p in ps:
            p.wait()
        return time.perf_counter() - t0

    t_init = launch(0)
    t_run = launch(nsteps)
    return max(t_run - t_init, 1e-9), t_init


def read_rec(path):
    """Tagged records of the reference driver's dumps: name[16] type[1] pad[3] n(int64) data."""
    import struct
    out, data, p = {}, open(path, "rb").read(), 0
    dts = {"d": np.float64, "i": np.int32, "h": np.int16, "c": np.uint8}
    while p < len(data):
        name = data[p:p + 16].split(b"\0")[0].decode()
        t = chr(data[p + 16])
        n = struct.unpack("<q", data[p + 20:p + 28])[0]
        p += 28
        out[name] = np.frombuffer(data, dtype=dts[t], count=n, offset=p)
        p += n * np.dtype(dts[t]).itemsize
    return out


def polar_c1_gpu(root):
    """The polar 100x90 two-day run (2880 steps) on the GPU at both precisions:
    its envelope HMX [jmxb][imxb] and ZB."""
    b, t = synth.config1()
    case = synth.write_case(os.path.join(root, "c1"), b, t)
    bnt = os.path.join(root, "c1", "bnt.txt")
    with open(bnt, "w") as fh:
        fh.write(case["bntline"] + "\n")
    out = {"case": case, "bnt_path": bnt}
    B = sloshgpu.Basin(case["dta"], case["bntline"])
    out["zb"] = B.field("ZB")
    for prec in (32, 64):
        X = sloshgpu.Batch(B, [case["trk"]], precision=prec)
        X.upload()
        X.step(X.nsteps[0])
        X.sync()
        out[prec] = X.get(0, "HMX")
        X.close()
    B.close()
    return out


def accuracy_vs_reference(c1, root):
    """The reference FP64 build's own envelope of the polar 100x90 run against
    the GPU's: max |HMX difference| over squares wet in both, and wet/dry
    agreement of the EnvSave values (999 where HMX + ZB == 0)."""
    exe = os.path.join(HERE, "oracle", "_ref", "refdrv64")
    if not os.path.exists(exe):
        return None
    outdir = tempfile.mkdtemp(prefix="c1ref_", dir=root)
    case = c1["case"]
    r = subprocess.run([exe, c1["bnt_path"], case["dta"], case["trk"], outdir, "-1", ""],
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        return None
    zb = c1["zb"]
    ref = read_rec(os.path.join(outdir, "final.bin"))["HMX"].reshape(zb.shape)
    res = {"workload": "polar 100x90 basin, 48 h straight track, dt=60 s (2880 steps), reference FP64 build"}
    for prec in (64, 32):
        g = c1[prec]
        sq = np.zeros_like(zb, dtype=bool)
        sq[:-1, :-1] = True
        wet_g, wet_r = (g + zb != 0) & sq, (ref + zb != 0) & sq
        both = wet_g & wet_r
        d = np.abs(g - ref)[both]
        res["f%d" % prec] = {"max_abs_ft": float(d.max()) if d.size else 0.0,
                             "within_0.01ft": float((d <= 0.01).mean()) if d.size else 1.0,
                             "wet_dry_agreement": float((wet_g == wet_r)[sq].mean()),
                             "wet_dry_flips": int((wet_g != wet_r).sum()), "squares": int(sq.sum()),
                             "bit_identical": bool(np.array_equal(g, ref))}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5600,
                    help="timed steps of the batched run (48 h runs = 5760 steps; the autotuner's trials "
                         "take 80 and the warm-up W of them); fewer are spread over the run in windows")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--span", type=int, default=0,
                    help="take the timed windows from the first SPAN steps after the warm-up only (0: to the end of "
                         "the run); for short profiling runs")
    ap.add_argument("--parts", default="batched,kernels,fp32,tide,setup,rex,plain,uneven,serve,cpu",
                    help="comma list of batched, kernels, fp32, tide, setup, rex, plain, uneven, serve, cpu "
                         "(and copy: the counter calibration copy)")
    ap.add_argument("--precision", type=int, default=64, choices=[32, 64],
                    help="arithmetic of the headline (64: the reference's default build, bit for bit)")
    ap.add_argument("--storms", type=int, default=1024)
    ap.add_argument("--plain-steps", type=int, default=8600, help="timed steps of the plain run (72 h = 8640)")
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="reference processes (one storm each) for the CPU baseline; 16 is one GPU's share "
                         "of the GPU box's cores")
    ap.add_argument("--path", choices=["auto", "fused", "staged"], default="auto")
    ap.add_argument("--tile-rows", type=int, default=32)
    ap.add_argument("--streams", type=int, default=3, choices=[1, 2, 3, 4],
                    help="step the batched ensemble in this many parts of storms, each on its own stream")
    ap.add_argument("--uneven-storms", type=int, default=4096)
    ap.add_argument("--serve-rates", default="5,20,50")
    ap.add_argument("--serve-seconds", type=float, default=15.0)
    ap.add_argument("--serve-hours", type=int, default=72)
    ap.add_argument("--serve-slots", type=int, default=512)
    ap.add_argument("--serve-pump", type=int, default=64)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(storms would shard across GPUs, see the design notes)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    parts = set(a.parts.split(","))
    prec = a.precision

    root = tempfile.mkdtemp(prefix="slosh_bench_")
    out = {
        "metric": "storm.cell.steps/s (batched MEOW ensemble; plain single storm beside it) @ envelope error "
                  "vs the reference",
        "value": None,
        "unit": "storm.cell.steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f%d" % prec,
        "data": "synthetic (seeded basins and tracks in SLOSH dta/trk formats)",
        "config": {"workload": "batched MEOW ensemble, %d storms, polar basin 250x300 cells, dt=30 s, "
                               "48 h runs (5760 steps), FP%d" % (a.storms, prec),
                   "storms": a.storms, "parallelism": "single GPU"},
        "roofline": None,
        "cpu_baseline": None,
    }
    try:
        if "copy" in parts:
            # a plain device copy of known size (1 GiB read, 1 GiB written): the
            # reference point that scales the FETCH_SIZE / WRITE_SIZE counters
            import ctypes
            hip = ctypes.CDLL("libamdhip64.so")
            src_buf, dst_buf, nb = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_size_t(1 << 30)
            ok = hip.hipMalloc(ctypes.byref(src_buf), nb) == 0 and hip.hipMalloc(ctypes.byref(dst_buf), nb) == 0
            ok = ok and hip.hipMemset(src_buf, 1, nb) == 0 and hip.hipMemcpy(dst_buf, src_buf, nb, 3) == 0
            ok = ok and hip.hipDeviceSynchronize() == 0
            hip.hipFree(src_buf)
            hip.hipFree(dst_buf)
            if not ok:
                raise RuntimeError("calibration copy failed")
            log("copy: 1 GiB device-to-device copy done")
        b, paths, trks = build_ensemble(root, a.storms)
        bnt_path = os.path.join(root, "bnt.txt")
        with open(bnt_path, "w") as fh:
            fh.write(paths["bntline"] + "\n")
        cells = (b.imxb - 1) * (b.jmxb - 1)
        out["config"]["cells"] = cells
        path = None if a.path == "auto" else (sloshgpu.PATH_FUSED if a.path == "fused" else sloshgpu.PATH_STAGED)
        rate = {}
        if "batched" in parts:
            log("batched %d storms on %dx%d polar basin, FP%d, %d streams" % (a.storms, b.imxb - 1, b.jmxb - 1,
                                                                              prec, a.streams))
            r = batched_run(paths["dta"], paths["bntline"], trks, prec, a.steps, a.warmup, a.streams, False,
                            path, a.tile_rows, a.span)
            rate[prec] = a.storms * r["cells"] * a.steps / r["dt"]
            log("batched %.3e storm.cell.steps/s (%s)" % (rate[prec], r["tune"]))
            out.update(value=rate[prec], ms_per_step=r["dt"] * 1e3 / a.steps)
            out["config"]["timed_windows"] = r["windows"]
            out["autotune"] = r["tune"]
            out["streams"] = r["streams"]
            out["schedule_s"] = r["sched_s"]
            step_bytes = b_alg_step(a.storms, prec)
            ach = rate[prec] * step_bytes / 1e9
            out["roofline"] = {"bound": "hbm", "kernel": "whole step (CONTTY + replay + stepB + advance)",
                               "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                               "traffic": None, "alg_bytes_per_storm_cell_step": step_bytes}
        if "kernels" in parts:
            log("kernels: the same batch at one stream with events, FP%d" % prec)
            rk = batched_run(paths["dta"], paths["bntline"], trks, prec, a.steps, a.warmup, 1, True,
                             path, a.tile_rows, a.span)
            kr = kernel_roofline(rk, prec, a.storms)
            one = a.storms * rk["cells"] * a.steps / rk["dt"]
            out["kernels"] = {"streams": 1, "value": one, "ms_per_step": rk["dt"] * 1e3 / a.steps,
                              "windows": rk["windows"], "per_kernel": kr}
            if kr:
                dom = max(kr, key=lambda k: kr[k]["avg_launch_ms"] * kr[k]["launches"])
                out["kernels"]["dominant"] = dom
                if out["roofline"] is not None:
                    tr = [kr[k]["traffic"] for k in kr if kr[k]["alg_bytes_per_launch"]]
                    if tr and all(x is not None for x in tr):
                        out["roofline"]["traffic"] = sum(tr)
                        out["roofline"]["traffic_note"] = ("HBM bytes of one step's CONTTY + stepB launches "
                                                           "(rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)")
        if "fp32" in parts and prec == 64:
            log("batched FP32, %d streams" % a.streams)
            r32 = batched_run(paths["dta"], paths["bntline"], trks, 32, a.steps, a.warmup, a.streams, False,
                              path, a.tile_rows, a.span)
            rate[32] = a.storms * r32["cells"] * a.steps / r32["dt"]
            out["fp32"] = {"value": rate[32], "ms_per_step": r32["dt"] * 1e3 / a.steps,
                           "windows": r32["windows"], "tune": r32["tune"],
                           "step_roofline_frac": rate[32] * b_alg_step(a.storms, 32) / (HBM_PEAK_GBS * 1e9)}
            log("batched FP32 %.3e storm.cell.steps/s" % rate[32])
        if "tide" in parts:
            # the same ensemble with the gridded tide (-f_tide V2.1.-5: STATICHT +
            # tide on squares 290 ft deep, none above 5 ft of land) between steps
            td = synth.write_tide_db(os.path.join(root, "tide"), b, "cfg3", seed=9, adj=True)
            tide = (os.path.join(td, "ft03.dta"), os.path.join(td, "cfg3.bhc"), os.path.join(td, "cfg3.adj"), 21, -5)
            log("batched with the gridded tide, FP%d" % prec)
            rt = batched_run(paths["dta"], paths["bntline"], trks, prec, a.steps, a.warmup, a.streams, False,
                             path, a.tile_rows, a.span, tide=tide)
            vt = a.storms * rt["cells"] * a.steps / rt["dt"]
            out["tide"] = {"workload": "the batched ensemble with -f_tide V2.1.-5 (37 constituents)",
                           "value": vt, "ms_per_step": rt["dt"] * 1e3 / a.steps, "set_tide_s": rt["tide_s"],
                           "tune": rt["tune"], "windows": rt["windows"], "streams": rt["streams"]}
            if prec in rate:
                out["tide"]["vs_without_tide"] = vt / rate[prec]
            log("tide: %.3e storm.cell.steps/s, set_tide %.2f s" % (vt, rt["tide_s"]))
        if "setup" in parts:
            tracks = [synth.track_dict(t) for t in synth.meow_tracks(b, n=a.storms)]
            out["setup"] = setup_bench(paths["dta"], paths["bntline"], trks, tracks, prec)
        if "rex" in parts:
            rx = rex_frames_bench(paths["dta"], paths["bntline"], trks, prec)
            log("rex: %d frames in %.2f ms" % (rx["storms"], rx["ms_per_frame_set"]))
            out["rex"] = rx
        if "plain" in parts:
            # plain single storm: elliptic 250 x 300, spline track (fused; and graphs)
            b2, t2 = synth.config2()
            p2 = synth.write_case(os.path.join(root, "plain"), b2, t2)
            out["plain"] = {"workload": "one storm, elliptic basin 250x300, dt=30 s, 72 h spline track"}
            for pp in sorted(set([prec, 32]) if "fp32" in parts else [prec]):
                rp = batched_run(p2["dta"], p2["bntline"], [p2["trk"]], pp, a.plain_steps, a.warmup, 1, False,
                                 sloshgpu.PATH_FUSED, a.tile_rows)
                plain_rate = rp["cells"] * a.plain_steps / rp["dt"]
                # the same with hipGraphs of 32 steps
                B2 = sloshgpu.Basin(p2["dta"], p2["bntline"])
                X2 = sloshgpu.Batch(B2, [p2["trk"]], precision=pp)
                X2.upload()
                X2.set_path(sloshgpu.PATH_FUSED, a.tile_rows)
                X2.set_graph(32)
                X2.step(a.warmup)
                X2.sync()
                t0 = time.perf_counter()
                X2.step(a.plain_steps)
                X2.sync()
                gdt = time.perf_counter() - t0
                X2.close()
                B2.close()
                graph_rate = rp["cells"] * a.plain_steps / gdt
                log("plain FP%d %.3e storm.cell.steps/s; with hipGraphs %.3e" % (pp, plain_rate, graph_rate))
                out["plain"]["f%d" % pp] = {"value": plain_rate, "ms_per_step": rp["dt"] * 1e3 / a.plain_steps,
                                            "graph_value": graph_rate, "graph_ms_per_step": gdt * 1e3 / a.plain_steps}
                if pp in rate:
                    out["plain"]["f%d" % pp]["batched_over_plain"] = rate[pp] / max(plain_rate, graph_rate)
            if prec in rate:
                out["batched_over_plain"] = out["plain"]["f%d" % prec]["batched_over_plain"]
        if "uneven" in parts:
            un = uneven(root, a.uneven_storms, prec)
            log("uneven %.3e storm.cell.steps/s" % un["value"])
            out["uneven"] = un
        if "serve" in parts:
            basins = synth.config4_basins()
            pm = plain_ms_per_step(root, basins, prec)
            log("plain ms/step by basin: %s" % ", ".join("%.3f" % m for m in pm))
            out["serve"] = []
            for k, lam in enumerate(float(x) for x in a.serve_rates.split(",") if x):
                sv = serve(root, lam, a.serve_seconds, a.serve_hours, a.serve_slots, a.serve_pump, prec, pm,
                           seed=55 + k)
                log("serve %g/s: p50 %.2fs p99 %.2fs offered %.3e served %.3e (plain path FIFO: p50 %.1fs p99 %.1fs)"
                    % (lam, sv["p50_latency_s"], sv["p99_latency_s"], sv["offered"], sv["value"],
                       sv["plain_fifo"]["p50_latency_s"], sv["plain_fifo"]["p99_latency_s"]))
                out["serve"].append(sv)
        if "cpu" in parts:
            c1 = polar_c1_gpu(root) if ("fp32" in parts or prec == 32) else None
            procs = max(1, a.cpu_procs)
            # a stratified sample of the ensemble: every (storms/procs)-th storm,
            # so headings, speeds, deficits, sizes and offsets are all in it
            stride = max(1, a.storms // procs)
            sample = [trks[(k * stride) % len(trks)] for k in range(procs)]
            nsteps = 5760
            res = cpu_baseline(root, paths["dta"], bnt_path, sample, nsteps, procs)
            if res is not None:
                t_cpu, t_init = res
                out["cpu_baseline"] = {
                    "value": procs * cells * nsteps / t_cpu,
                    "unit": "storm.cell.steps/s", "cores": procs, "kind": "reference",
                    "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
                    "sample": "%d of the %d ensemble storms (every %d-th), each a whole 48 h run (%d steps) on "
                              "the 250x300 basin, reference FP64 build, one serial process per storm on %d "
                              "cores, init (%.2fs) excluded" % (procs, a.storms, stride, nsteps, procs, t_init)}
                if rate.get(prec):
                    out["cpu_baseline"]["gpu_over_cpu"] = rate[prec] / out["cpu_baseline"]["value"]
            if c1 is not None:
                acc = accuracy_vs_reference(c1, root)
                if acc:
                    out["accuracy_vs_reference"] = acc
                    log("accuracy: %r" % acc)
        print(json.dumps(out), flush=True)
    finally:
        shutil.rmtree(root, ignore_errors=True)


if __name__ == "__main__":
    main()
