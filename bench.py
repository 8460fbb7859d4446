#!/usr/bin/env python3
"""Throughput of the SLOSH time-stepping core on one MI355X.

One "step" is one model time step (TMSTPC1 + TMSTPC2 + TMSTPM of the reference)
applied to every storm of the batch.  The headline workload is the batched MEOW
ensemble: 1024 storms (8 headings x 4 speeds x 4 pressure deficits x 2 RMW x 4
offsets) on one synthetic polar basin of 250 x 300 cells at dt = 30 s, all
storms advancing together, fields resident in HBM, FP32.  Beside it:

  plain    one storm on the 250 x 300 elliptic basin (72 h spline track), the
           single-storm GPU path (fused kernels; and as hipGraph replays);
  uneven   4096 storms over six basins of 4 800 - 140 000 cells (polar,
           elliptic, hyperbolic), 24-72 h runs, one batch per basin on its own
           stream, all advancing together;
  serve    Poisson arrivals of forecast-track runs over the six basins, admitted
           into device slots as they arrive; p50/p99 latency per run;
  cpu      the reference's own FP64 build on this host's cores, one serial
           process per storm (its multi-process run pattern).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  Inputs are synthetic and seeded; rates are storm cell
steps per second (storm.cell.steps/s).
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
# Algorithmic HBM bytes per storm.cell of one launch of each kernel (FP32 fields,
# each element read or written once by the launch; the basin's static fields are
# shared by every storm of the batch and are not counted per storm):
ALG_BYTES = {
    # fused path
    "contty": 17,   # HB in, UB, VB in; HB out, U==0 mask byte out
    "replay": 0,    # serial cascade: touches only the drying squares
    "stepB": 33,    # HB in, mask byte, HMX in/out, UB/VB in/out, HB out
    # staged path (one kernel per routine)
    "pflag": 9,     # UB, HB in, flag byte out
    "hclass": 6,    # HB, flag byte in, class byte out
    "smooth": 9,    # HB in/out, class byte
    "hmxsv": 12,    # HB, HMX in/out
    "filter": 8,    # HB in/out
    "momntm": 20,   # HB, UB/VB in/out
}


def b_alg_step(storms):
    """Whole-step algorithmic bytes per storm.cell.step: HB, UB, VB, HMX each read
    and written once (32 B) plus the static ZB, ZBM, Coriolis factor and flags
    (16 B) once per basin cell, shared by the storms of the batch."""
    return 32.0 + 16.0 / storms


def measured_traffic(kernel, storms, per_launch):
    """HBM bytes per launch of `kernel` at this workload (per_launch of the
    ensemble's storms in one launch), from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of the 1024-storm launch (copy-calibrated, scaled per
    storm); None if there is no measurement for it."""
    path = os.path.join(HERE, "profiles", "stepB_meow1024_hbm_traffic.md")
    if kernel != "stepB" or storms != 1024 or not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith("traffic_bytes_per_launch stepB"):
                return float(line.split()[-1]) * per_launch / 1024.0
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


# ------------------------------------------------------------------ headline
def build_ensemble(root, nstorm):
    b = synth.config3_basin()
    tracks = synth.meow_tracks(b, n=nstorm)
    paths = synth.write_case(root, b, tracks[0], trk_name="s0000.trk")
    trks = [paths["trk"]] + write_tracks(root, tracks[1:], "s1_")
    return b, paths, trks


def gpu_run(dta, bnt, trks, steps, warmup, path=None, rows=16, graph=0, events=True, streams=1):
    """Open, upload, (autotune,) warm up, then time `steps` steps of the batch.
    Without graphs the timed steps carry HIP events around every kernel (on
    the batch's stream), which give the per-kernel durations of the same
    region; the steps must stay within every storm's run."""
    B = sloshgpu.Basin(dta, bnt)
    t0 = time.time()
    X = sloshgpu.Batch(B, trks)
    t_sched = time.time() - t0
    X.upload()
    tune = None
    if path is None:   # the online autotuner picks the launch configuration
        r, p, ms = X.autotune(20)
        tune = {"tile_rows": r, "path": "fused" if p == 0 else "staged",
                "ms_per_step": {"fused8": ms[0], "fused16": ms[1], "fused32": ms[2], "staged": ms[3]}}
    else:
        X.set_path(path, rows)
    if streams > 1 and X.n > 1 and (path or sloshgpu.PATH_FUSED) == sloshgpu.PATH_FUSED \
            and (tune is None or tune["path"] == "fused"):
        X.set_streams(streams)
    X.step(warmup)
    X.sync()
    left = min(X.nsteps) - X.current_step
    if steps > left:
        raise SystemExit("bench.py: %d timed steps asked, the shortest run has %d left after tuning "
                         "and warm-up" % (steps, left))
    ms = cnt = None
    if graph or not events:
        if graph:
            X.set_graph(graph)
        t0 = time.perf_counter()
        X.step(steps)
        X.sync()
        dt = time.perf_counter() - t0
    else:
        t0 = time.perf_counter()
        ms, cnt = X.step_timed(steps)
        dt = time.perf_counter() - t0
    nstreams = X.streams
    X.close()
    cells = B.cells
    B.close()
    return dict(dt=dt, cells=cells, ms=ms, cnt=cnt, sched_s=t_sched, tune=tune, streams=nstreams)


# ----------------------------------------------------------------------- rex
def rex_frames_bench(dta, bnt, trks, reps=5):
    """rex v2 frames of every storm of the ensemble encoded on the device
    (slg_batch_rex_frames: HHB of the last step, run-length coded, copied back)
    after 20 staged steps; time per frame set, over reps calls after one
    untimed call."""
    B = sloshgpu.Basin(dta, bnt)
    X = sloshgpu.Batch(B, trks)
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
            "note": "HHB = min(HB, HMX) per cell (8 B read in FP32) to rex v2 run-length records, "
                    "copied to the host; includes the host copy"}


# -------------------------------------------------------------------- uneven
def uneven(root, nstorm, steps, warmup, graph=20):
    """nstorm storms over the six basins (share proportional to cells), one
    batch per basin on its own stream; each batch's launch shape picked by the
    online autotuner and its step loop replayed as hipGraphs of `graph` steps;
    the batches advance together, `graph` steps at a time."""
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
        X = sloshgpu.Batch(B, trks)
        X.upload()
        r, p, ms = X.autotune(10)
        tunes.append({"tile_rows": r, "path": "fused" if p == 0 else "staged"})
        if graph:
            X.set_graph(graph)
        runs.append((B, X, n))
    setup = time.time() - t0
    chunk = max(1, graph)
    for _ in range(max(1, warmup // chunk)):
        for B, X, n in runs:
            X.step(chunk)
    for B, X, n in runs:
        X.sync()
    nchunks = max(1, steps // chunk)
    t0 = time.perf_counter()
    for _ in range(nchunks):
        for B, X, n in runs:
            X.step(chunk)
    for B, X, n in runs:
        X.sync()
    dt = time.perf_counter() - t0
    steps = nchunks * chunk
    work = sum(n * B.cells for B, X, n in runs) * steps
    out = {"value": work / dt, "unit": "storm.cell.steps/s", "storms": int(nstorm), "steps": steps,
           "ms_per_step": dt * 1e3 / steps, "setup_s": setup, "graph_steps": graph,
           "basins": [{"name": b.name, "grid": {" ": "polar", "$": "elliptic", "+": "hyperbolic"}[b.gtype],
                       "cells": int(c), "storms": int(n), "tuned": t}
                      for b, c, n, t in zip(basins, cells, share, tunes)]}
    for B, X, n in runs:
        X.close()
        B.close()
    return out


# --------------------------------------------------------------------- serve
def plain_ms_per_step(root, basins, steps=8600):
    """Per-step time of the plain single-storm GPU path (fused kernels, one
    storm) on each basin, over a whole 72 h forecast-track run (8640 steps)."""
    rng = np.random.default_rng(77)
    out = []
    for b in basins:
        d = os.path.join(root, "pl_" + b.name)
        case = synth.write_case(d, b, synth.forecast_track(b, rng, hours=72))
        r = gpu_run(case["dta"], case["bntline"], [case["trk"]], steps, 20, sloshgpu.PATH_FUSED, 16, events=False)
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


def serve(root, rate, seconds, hours, slots, pump_steps, plain_ms=None, seed=55):
    """Poisson arrivals of forecast-track runs (basin chosen in proportion to its
    cells) admitted into device slots as they arrive; per-run latency from
    submission to envelope."""
    basins = synth.config4_basins()
    cells = np.array([(b.imxb - 1) * (b.jmxb - 1) for b in basins], dtype=float)
    rng = np.random.default_rng(seed)
    S = sloshgpu.Server(precision=32, hbm_fraction=0.5)
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
    S.close()
    out = {"rate_per_s": rate, "arrival_window_s": seconds, "runs": n, "horizon_h": hours,
           "slots_per_basin": slots_used, "pump_steps": pump_steps, "pumps": pumps,
           "host_s_in_pump": t_pump,
           "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
           "mean_latency_s": float(lat.mean()), "wall_s": wall,
           "value": st["cell_steps"] / wall, "unit": "storm.cell.steps/s", "pool_bytes": st["bytes"]}
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
def cpu_baseline(root, dta, bnt_path, trks, nsteps, procs):
    """The reference FP64 build (oracle/_ref/refdrv64), one serial process per
    storm as the reference's multi-process runner does, on `procs` cores."""
    exe = os.path.join(HERE, "oracle", "_ref", "refdrv64")
    if not os.path.exists(exe):
        return None

    def launch(n):
        ps = []
        for k in range(procs):
            out = tempfile.mkdtemp(prefix="cpu%02d_" % k, dir=root)
            ps.append(subprocess.Popen([exe, bnt_path, dta, trks[k % len(trks)], out, str(n), ""],
                                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
        t0 = time.perf_counter()
        for p in ps:
            p.wait()
        return time.perf_counter() - t0

    t_init = launch(0)
    t_run = launch(nsteps)
    return max(t_run - t_init, 1e-9), t_init


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5600,
                    help="timed steps of the batched run (48 h runs = 5760 steps; the autotuner's trials "
                         "take 80 and the warm-up W of them)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--parts", default="batched,rex,plain,uneven,serve,cpu",
                    help="comma list of batched, rex, plain, uneven, serve, cpu (and copy: the counter "
                         "calibration copy)")
    ap.add_argument("--storms", type=int, default=1024)
    ap.add_argument("--plain-steps", type=int, default=8600, help="timed steps of the plain run (72 h = 8640)")
    ap.add_argument("--cpu-steps", type=int, default=2000)
    ap.add_argument("--path", choices=["auto", "fused", "staged"], default="auto")
    ap.add_argument("--tile-rows", type=int, default=16)
    ap.add_argument("--streams", type=int, default=3, choices=[1, 2, 3, 4],
                    help="step the batched ensemble in this many parts of storms, each on its own stream")
    ap.add_argument("--uneven-storms", type=int, default=4096)
    ap.add_argument("--uneven-steps", type=int, default=400)
    ap.add_argument("--serve-rates", default="5,20,50")
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--serve-hours", type=int, default=72)
    ap.add_argument("--serve-slots", type=int, default=512)
    ap.add_argument("--serve-pump", type=int, default=64)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(storms would shard across GPUs, see the design notes)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    parts = set(a.parts.split(","))

    root = tempfile.mkdtemp(prefix="slosh_bench_")
    out = {
        "metric": "storm.cell.steps/s (batched MEOW ensemble; plain single storm beside it)",
        "value": None,
        "unit": "storm.cell.steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded basins and tracks in SLOSH dta/trk formats)",
        "config": {"workload": "batched MEOW ensemble, %d storms, polar basin 250x300 cells, "
                               "dt=30 s (48 h runs)" % a.storms,
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
        rate = None
        if "batched" in parts:
            log("batched %d storms on %dx%d polar basin" % (a.storms, b.imxb - 1, b.jmxb - 1))
            path = None if a.path == "auto" else (sloshgpu.PATH_FUSED if a.path == "fused"
                                                  else sloshgpu.PATH_STAGED)
            r = gpu_run(paths["dta"], paths["bntline"], trks, a.steps, a.warmup, path, a.tile_rows,
                        streams=a.streams)
            rate = a.storms * r["cells"] * a.steps / r["dt"]
            log("batched %.3e storm.cell.steps/s (%s)" % (rate, r["tune"]))
            # dominant kernel, timed with HIP events on the batch's stream
            avg = {k: r["ms"][i] / r["cnt"][i] for i, k in enumerate(sloshgpu.KERNELS) if r["cnt"][i]}
            dom = max(avg, key=lambda k: r["ms"][sloshgpu.KERNELS.index(k)])
            # with two streams the events time the first half's launches
            per_launch = a.storms // r["streams"]
            dom_bytes = ALG_BYTES[dom] * per_launch * r["cells"]
            achieved = dom_bytes / (avg[dom] * 1e-3) / 1e9
            out.update(value=rate, ms_per_step=r["dt"] * 1e3 / a.steps)
            out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                               "traffic": measured_traffic(dom, a.storms, per_launch),
                               "alg_bytes_per_launch": dom_bytes, "avg_launch_ms": avg[dom]}
            out["autotune"] = r["tune"]
            out["streams"] = r["streams"]
            out["roofline"]["storms_per_launch"] = per_launch
            out["kernel_avg_ms"] = avg
            out["step_roofline_frac"] = rate * b_alg_step(a.storms) / (HBM_PEAK_GBS * 1e9)
        if "rex" in parts:
            rx = rex_frames_bench(paths["dta"], paths["bntline"], trks)
            log("rex: %d frames in %.2f ms" % (rx["storms"], rx["ms_per_frame_set"]))
            out["rex"] = rx
        if "plain" in parts:
            # plain single storm: elliptic 250 x 300, spline track (fused; and graphs)
            b2, t2 = synth.config2()
            p2 = synth.write_case(os.path.join(root, "plain"), b2, t2)
            rp = gpu_run(p2["dta"], p2["bntline"], [p2["trk"]], a.plain_steps, a.warmup, sloshgpu.PATH_FUSED, 16,
                         events=False)
            plain_rate = rp["cells"] * a.plain_steps / rp["dt"]
            rg = gpu_run(p2["dta"], p2["bntline"], [p2["trk"]], a.plain_steps, a.warmup, sloshgpu.PATH_FUSED, 16,
                         graph=32)
            graph_rate = rg["cells"] * a.plain_steps / rg["dt"]
            log("plain %.3e storm.cell.steps/s; with hipGraphs %.3e" % (plain_rate, graph_rate))
            best = max(plain_rate, graph_rate)
            out["plain"] = {"value": plain_rate, "workload": "one storm, elliptic basin 250x300, "
                            "dt=30 s, 72 h spline track", "ms_per_step": rp["dt"] * 1e3 / a.plain_steps,
                            "graph_value": graph_rate, "graph_ms_per_step": rg["dt"] * 1e3 / a.plain_steps}
            if rate:
                out["batched_over_plain"] = rate / best
        if "uneven" in parts:
            un = uneven(root, a.uneven_storms, a.uneven_steps, 20)
            log("uneven %.3e storm.cell.steps/s" % un["value"])
            out["uneven"] = un
        if "serve" in parts:
            basins = synth.config4_basins()
            pm = plain_ms_per_step(root, basins)
            log("plain ms/step by basin: %s" % ", ".join("%.3f" % m for m in pm))
            out["serve"] = []
            for k, lam in enumerate(float(x) for x in a.serve_rates.split(",") if x):
                sv = serve(root, lam, a.serve_seconds, a.serve_hours, a.serve_slots, a.serve_pump, pm,
                           seed=55 + k)
                log("serve %g/s: p50 %.2fs p99 %.2fs (plain path FIFO: p50 %.1fs p99 %.1fs)"
                    % (lam, sv["p50_latency_s"], sv["p99_latency_s"], sv["plain_fifo"]["p50_latency_s"],
                       sv["plain_fifo"]["p99_latency_s"]))
                out["serve"].append(sv)
        if "cpu" in parts:
            procs = max(1, min(16, os.cpu_count() or 1))
            res = cpu_baseline(root, paths["dta"], bnt_path, trks, a.cpu_steps, procs)
            if res is not None:
                t_cpu, t_init = res
                out["cpu_baseline"] = {
                    "value": procs * cells * a.cpu_steps / t_cpu,
                    "unit": "storm.cell.steps/s", "cores": procs, "kind": "reference",
                    "sample": "%d ensemble storms x %d steps (of 5760) on the 250x300 basin, reference "
                              "FP64 build, one serial process per storm, init (%.2fs) excluded"
                              % (procs, a.cpu_steps, t_init)}
        print(json.dumps(out), flush=True)
    finally:
        shutil.rmtree(root, ignore_errors=True)


if __name__ == "__main__":
    main()
