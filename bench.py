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


def gpu_run(dta, bnt, trks, steps, warmup, timed_steps, path=None, rows=16, graph=0):
    """Open, upload, (autotune,) warm up, then time `steps` steps of the batch."""
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
    if graph:
        X.set_graph(graph)
    X.step(warmup)
    X.sync()
    t0 = time.perf_counter()
    X.step(steps)
    X.sync()
    dt = time.perf_counter() - t0
    X.set_graph(0)
    ms, cnt = X.step_timed(timed_steps)
    X.close()
    cells = B.cells
    B.close()
    return dict(dt=dt, cells=cells, ms=ms, cnt=cnt, sched_s=t_sched, tune=tune)


# -------------------------------------------------------------------- uneven
def uneven(root, nstorm, steps, warmup):
    """nstorm storms over the six basins (share proportional to cells), one
    batch per basin on its own stream, all stepped together."""
    basins = synth.config4_basins()
    cells = np.array([(b.imxb - 1) * (b.jmxb - 1) for b in basins], dtype=float)
    share = np.floor(nstorm * cells / cells.sum()).astype(int)
    share[-1] += nstorm - share.sum()
    rng = np.random.default_rng(4096)
    runs = []
    t0 = time.time()
    for b, n in zip(basins, share):
        d = os.path.join(root, "c4_" + b.name)
        tracks = [synth.random_track(b, rng) for _ in range(n)]
        case = synth.write_case(d, b, tracks[0])
        trks = [case["trk"]] + write_tracks(d, tracks[1:], "t")
        B = sloshgpu.Basin(case["dta"], case["bntline"])
        X = sloshgpu.Batch(B, trks)
        X.upload()
        runs.append((B, X, n))
    setup = time.time() - t0
    for _ in range(warmup):
        for B, X, n in runs:
            X.step(1)
    for B, X, n in runs:
        X.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        for B, X, n in runs:
            X.step(1)
    for B, X, n in runs:
        X.sync()
    dt = time.perf_counter() - t0
    work = sum(n * B.cells for B, X, n in runs) * steps
    out = {"value": work / dt, "unit": "storm.cell.steps/s", "storms": int(nstorm), "steps": steps,
           "ms_per_step": dt * 1e3 / steps, "setup_s": setup,
           "basins": [{"name": b.name, "grid": {" ": "polar", "$": "elliptic", "+": "hyperbolic"}[b.gtype],
                       "cells": int(c), "storms": int(n)} for b, c, n in zip(basins, cells, share)]}
    for B, X, n in runs:
        X.close()
        B.close()
    return out


# --------------------------------------------------------------------- serve
def serve(root, rate, seconds, hours, slots, pump_steps):
    """Poisson arrivals of forecast-track runs (basin chosen in proportion to its
    cells) admitted into device slots as they arrive; per-run latency from
    submission to envelope."""
    basins = synth.config4_basins()
    cells = np.array([(b.imxb - 1) * (b.jmxb - 1) for b in basins], dtype=float)
    rng = np.random.default_rng(55)
    S = sloshgpu.Server(precision=32, hbm_fraction=0.5)
    ids, cases = [], []
    for b in basins:
        d = os.path.join(root, "c5_" + b.name)
        case = synth.write_case(d, b, synth.forecast_track(b, rng, hours=hours))
        maxsteps = int(hours * 3600 / 30) + 400
        ids.append(S.add_basin(case["dta"], case["bntline"], max_slots=slots, max_steps=maxsteps,
                               tables_per_slot=160))
        cases.append(d)
    # arrivals
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
    nxt = 0
    last = -1
    while True:
        now = time.perf_counter() - t0
        while nxt < n and times[nxt] <= now:
            submitted.append((S.submit(ids[which[nxt]], trks[nxt]), ids[which[nxt]]))
            nxt += 1
        S.pump(pump_steps)
        st = S.stats()
        if st["done"] != last and st["done"] % 20 == 0:
            last = st["done"]
            log("serve: t=%.1fs submitted %d done %d running %d" % (now, nxt, st["done"], st["running"]))
        if nxt == n and st["done"] >= n:
            break
        if now > seconds + 900:
            raise RuntimeError("serving did not drain: %r" % st)
        if st["running"] == 0 and st["queued"] == 0 and nxt < n:
            time.sleep(max(0.0, min(0.005, times[nxt] - now)))
    wall = time.perf_counter() - t0
    lat = np.array([S.result(r, b)[2] for r, b in submitted])
    st = S.stats()
    S.close()
    return {"rate_per_s": rate, "arrival_window_s": seconds, "runs": n, "horizon_h": hours,
            "slots_per_basin": slots, "pump_steps": pump_steps,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "wall_s": wall,
            "value": st["cell_steps"] / wall, "unit": "storm.cell.steps/s", "pool_bytes": st["bytes"]}


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
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--storms", type=int, default=1024)
    ap.add_argument("--plain-steps", type=int, default=1000)
    ap.add_argument("--timed-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--path", choices=["auto", "fused", "staged"], default="auto")
    ap.add_argument("--tile-rows", type=int, default=16)
    ap.add_argument("--uneven-storms", type=int, default=4096)
    ap.add_argument("--uneven-steps", type=int, default=60)
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=5.0)
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--serve-hours", type=int, default=72)
    ap.add_argument("--serve-slots", type=int, default=64)
    ap.add_argument("--serve-pump", type=int, default=64)
    ap.add_argument("--no-serve", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(storms would shard across GPUs, see the design notes)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    root = tempfile.mkdtemp(prefix="slosh_bench_")
    try:
        b, paths, trks = build_ensemble(root, a.storms)
        bnt_path = os.path.join(root, "bnt.txt")
        with open(bnt_path, "w") as fh:
            fh.write(paths["bntline"] + "\n")
        log("batched %d storms on %dx%d polar basin" % (a.storms, b.imxb - 1, b.jmxb - 1))
        path = None if a.path == "auto" else (sloshgpu.PATH_FUSED if a.path == "fused" else sloshgpu.PATH_STAGED)
        r = gpu_run(paths["dta"], paths["bntline"], trks, a.steps, a.warmup, a.timed_steps, path, a.tile_rows)
        rate = a.storms * r["cells"] * a.steps / r["dt"]
        log("batched %.3e storm.cell.steps/s (%s)" % (rate, r["tune"]))

        # plain single storm: elliptic 250 x 300, spline track (fused; and graphs)
        root2 = os.path.join(root, "plain")
        b2, t2 = synth.config2()
        p2 = synth.write_case(root2, b2, t2)
        rp = gpu_run(p2["dta"], p2["bntline"], [p2["trk"]], a.plain_steps, a.warmup, a.timed_steps,
                     sloshgpu.PATH_FUSED, 16)
        plain_rate = rp["cells"] * a.plain_steps / rp["dt"]
        rg = gpu_run(p2["dta"], p2["bntline"], [p2["trk"]], a.plain_steps, a.warmup, a.timed_steps,
                     sloshgpu.PATH_FUSED, 16, graph=32)
        graph_rate = rg["cells"] * a.plain_steps / rg["dt"]
        log("plain %.3e storm.cell.steps/s; with hipGraphs %.3e" % (plain_rate, graph_rate))

        un = None
        if not a.no_uneven:
            un = uneven(root, a.uneven_storms, a.uneven_steps, 5)
            log("uneven %.3e storm.cell.steps/s" % un["value"])
        sv = None
        if not a.no_serve:
            sv = serve(root, a.serve_rate, a.serve_seconds, a.serve_hours, a.serve_slots, a.serve_pump)
            log("serve p50 %.2fs p99 %.2fs" % (sv["p50_latency_s"], sv["p99_latency_s"]))

        # dominant kernel of the batched run, timed with HIP events on its stream
        avg = {k: (r["ms"][i] / r["cnt"][i] if r["cnt"][i] else 0.0)
               for i, k in enumerate(sloshgpu.KERNELS) if r["cnt"][i]}
        dom = max(avg, key=lambda k: avg[k] * (r["cnt"][sloshgpu.KERNELS.index(k)] or 0))
        dom_bytes = ALG_BYTES[dom] * a.storms * r["cells"]
        achieved = dom_bytes / (avg[dom] * 1e-3) / 1e9
        step_ms = r["dt"] * 1e3 / a.steps
        cpu = None
        if not a.no_cpu:
            procs = max(1, min(16, os.cpu_count() or 1))
            res = cpu_baseline(root, paths["dta"], bnt_path, trks, a.cpu_steps, procs)
            if res is not None:
                t_cpu, t_init = res
                cpu = {"value": procs * r["cells"] * a.cpu_steps / t_cpu,
                       "unit": "storm.cell.steps/s", "cores": procs, "kind": "reference",
                       "sample": "%d ensemble storms x %d steps (of 5760) on the 250x300 "
                                 "basin, reference FP64 build, one serial process per storm, "
                                 "init (%.2fs) excluded" % (procs, a.cpu_steps, t_init)}
        out = {
            "metric": "storm.cell.steps/s (batched MEOW ensemble; plain single storm beside it)",
            "value": rate,
            "unit": "storm.cell.steps/s",
            "n_gpus": 1,
            "steps": a.steps,
            "warmup": a.warmup,
            "ms_per_step": step_ms,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "fp32",
            "data": "synthetic (seeded basins and tracks in SLOSH dta/trk formats)",
            "config": {"workload": "batched MEOW ensemble, %d storms, polar basin %dx%d cells, "
                                   "dt=30 s (48 h runs)" % (a.storms, b.imxb - 1, b.jmxb - 1),
                       "storms": a.storms, "cells": r["cells"], "parallelism": "single GPU"},
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": None,
                         "alg_bytes_per_launch": dom_bytes,
                         "avg_launch_ms": avg[dom]},
            "cpu_baseline": cpu,
            "autotune": r["tune"],
            "kernel_avg_ms": avg,
            "step_roofline_frac": rate * b_alg_step(a.storms) / (HBM_PEAK_GBS * 1e9),
            "plain": {"value": plain_rate, "workload": "one storm, elliptic basin 250x300, "
                      "dt=30 s, 72 h spline track", "ms_per_step": rp["dt"] * 1e3 / a.plain_steps,
                      "graph_value": graph_rate, "graph_ms_per_step": rg["dt"] * 1e3 / a.plain_steps},
            "batched_over_plain": rate / plain_rate,
            "batched_over_plain_graph": rate / graph_rate,
            "uneven": un,
            "serve": sv,
        }
        print(json.dumps(out), flush=True)
    finally:
        shutil.rmtree(root, ignore_errors=True)


if __name__ == "__main__":
    main()
