#!/usr/bin/env python3
"""Throughput of the SLOSH time-stepping core on one MI355X.

One "step" is one model time step (TMSTPC1 + TMSTPC2 + TMSTPM of the reference)
applied to every storm of the batch.  The headline workload is the batched MEOW
ensemble: 1024 storms (8 headings x 4 speeds x 4 pressure deficits x 2 RMW x 4
offsets) on one synthetic polar basin of 250 x 300 cells at dt = 30 s, all
storms advancing together, fields resident in HBM.  The plain path (one storm on
the 250 x 300 elliptic basin, 72 h spline track) is timed beside it, and the
reference's own serial FP64 build on this host's cores gives the CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--storms S]

Prints one JSON line.  Inputs are synthetic and seeded; values are storm
cell steps per second (storm.cell.steps/s).
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
# Whole-step algorithmic bytes per storm.cell.step: HB, UB, VB, HMX each read
# and written once (32 B) plus the static ZB, ZBM, Coriolis factor and flags
# (16 B) once per basin cell, shared by the S storms of the batch.
def b_alg_step(storms):
    return 32.0 + 16.0 / storms


def build_ensemble(root, nstorm):
    b = synth.config3_basin()
    tracks = synth.meow_tracks(b, n=nstorm)
    paths = synth.write_case(root, b, tracks[0], trk_name="s0000.trk")
    trks = [paths["trk"]]
    for k, t in enumerate(tracks[1:], start=1):
        p = os.path.join(root, "s%04d.trk" % k)
        with open(p, "w") as fh:
            fh.write(synth.trk_text(t))
        trks.append(p)
    return b, paths, trks


def gpu_run(dta, bnt, trks, steps, warmup, timed_steps, path=0, rows=16):
    B = sloshgpu.Basin(dta, bnt)
    t0 = time.time()
    X = sloshgpu.Batch(B, trks)
    t_sched = time.time() - t0
    X.set_path(path, rows)
    X.upload()
    X.step(warmup)
    X.sync()
    t0 = time.perf_counter()
    X.step(steps)
    X.sync()
    dt = time.perf_counter() - t0
    ms, cnt = X.step_timed(timed_steps)
    stats = X.stats(0)
    X.close()
    cells = B.cells
    B.close()
    return dict(dt=dt, cells=cells, ms=ms, cnt=cnt, sched_s=t_sched, stats=stats)


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
    ap.add_argument("--path", choices=["fused", "staged"], default="fused")
    ap.add_argument("--tile-rows", type=int, default=16)
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
        print("bench: batched %d storms on %dx%d polar basin" % (a.storms, b.imxb - 1, b.jmxb - 1),
              file=sys.stderr, flush=True)
        path = sloshgpu.PATH_FUSED if a.path == "fused" else sloshgpu.PATH_STAGED
        r = gpu_run(paths["dta"], paths["bntline"], trks, a.steps, a.warmup, a.timed_steps, path,
                    a.tile_rows)
        rate = a.storms * r["cells"] * a.steps / r["dt"]
        print("bench: batched %.3e storm.cell.steps/s" % rate, file=sys.stderr, flush=True)

        # plain single storm: configuration 2 (elliptic 250 x 300, spline track)
        root2 = os.path.join(root, "plain")
        b2, t2 = synth.config2()
        p2 = synth.write_case(root2, b2, t2)
        rp = gpu_run(p2["dta"], p2["bntline"], [p2["trk"]], a.plain_steps, a.warmup, a.timed_steps,
                     path, a.tile_rows)
        plain_rate = rp["cells"] * a.plain_steps / rp["dt"]
        print("bench: plain %.3e storm.cell.steps/s" % plain_rate, file=sys.stderr, flush=True)

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
            "plain": {"value": plain_rate, "workload": "one storm, elliptic basin 250x300, "
                      "dt=30 s, 72 h spline track", "ms_per_step": rp["dt"] * 1e3 / a.plain_steps},
            "batched_over_plain": rate / plain_rate,
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": None,
                         "alg_bytes_per_launch": dom_bytes,
                         "avg_launch_ms": avg[dom]},
            "step_roofline_frac": rate * b_alg_step(a.storms) / (HBM_PEAK_GBS * 1e9),
            "path": a.path, "tile_rows": a.tile_rows,
            "kernel_avg_ms": avg,
            "cpu_baseline": cpu,
        }
        print(json.dumps(out), flush=True)
    finally:
        shutil.rmtree(root, ignore_errors=True)


if __name__ == "__main__":
    main()
