"""Benchmark: tracked segments/s and seconds per IMC step on one MI355X.

Workload (N = 1): the 64^3 gray IMC cube (storm_amd.problems.c2_cube):
sigma_a = sigma0 * max(T, 0.1 keV)^-3 with sigma0 log-uniform in [1, 100] per
cell, cv proportional to T, T0 = 1e-3 keV, a 1 keV source on the x- face and
escape on every face, ~1e7 packets per step, dt = 1e-11 s, census resident in
HBM.  One "step" = one full IMC step (preStep, sourcing, transport, comb,
postStep).  W untimed steps, then K timed steps; `value` is the whole-job
segments/s of the mixed-precision engine.  The plain FP64 engine runs the
same W + K steps for comparison, and the FP64 CPU oracle is timed on a
bounded sample of the same workload (`cpu_baseline`).

Roofline: no MFMA (nothing is a dense contraction), so the roofline is HBM
bandwidth against the algorithmic bytes of the metric's model: 144 B per
segment (a 48 B packet read and written, a 16 B cell record, two 16 B
tallies).  achieved = segments * 144 B / transport time (HIP events on the
engine's stream around each launch), over the timed steps.  The history
kernel keeps a packet in registers for its whole history, so its measured
HBM traffic per segment is far below the model (rocprof FETCH_SIZE /
WRITE_SIZE).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "tracked segments/s and s per IMC step, mixed vs FP64 GPU at matched tally error"
B_SEG_MIXED = 144.0  # bytes per segment: 2 * 48 B packet + 16 B cell + 2 * 16 B tally
B_SEG_FP64 = 224.0   # 2 * 80 B packet + 32 B cell + 2 * 16 B tally
HBM_PEAK_GBS = 8000.0
# HBM bytes per transport launch, measured by rocprofv3 (FETCH_SIZE and
# WRITE_SIZE, scaled by the same counters on a plain copy kernel) on this
# benchmark's own command; the summary holds the figures and the command.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_transport_c2_cube_counters.md")


def measured_traffic():
    """(bytes per launch, source) from the committed counter summary, or (None, None)."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1]), os.path.relpath(TRAFFIC_PROFILE, ROOT)
    except OSError:
        pass
    return None, None


def run_engine(precision, steps, warmup, seed, log, schedule=None, copy_mib=0):
    from storm_amd import problems as P
    from storm_amd.engine import GpuEngine

    spec = P.c2_cube(seed=seed)
    with GpuEngine(spec, precision) as e:
        if schedule is not None:
            e.set_tuning(schedule=schedule)
        for s in range(warmup):
            e.step()
            log(f"warmup step {s + 1}/{warmup}: {e.counters()['segments']} segments")
        seg = 0
        t_tr = 0.0
        launches = 0
        packets = 0
        t0 = time.perf_counter()
        for s in range(steps):
            e.step()
            c = e.counters()
            seg += c["segments"]
            t_tr += c["seconds_transport"]
            launches += c["transport_launches"]
            packets += c["start_particles"]
            log(f"step {s + 1}/{steps}: {c['segments']} segments, {c['start_particles']} packets, "
                f"transport {c['seconds_transport'] * 1e3:.2f} ms, step {c['seconds_step'] * 1e3:.2f} ms")
        wall = time.perf_counter() - t0
    return dict(segments=seg, wall=wall, transport_s=t_tr, launches=launches, packets=packets,
                name=spec.name, N=spec.ncells)


def cpu_baseline(log):
    """FP64 oracle on the host cores: one independent history batch per thread
    (different seeds of the same workload, 3/20 of the packets, 2 steps),
    aggregate segments over the wall time of the batch."""
    import concurrent.futures as cf
    from oracle import pyoracle
    from storm_amd import problems as P

    threads = max(1, min(16, os.cpu_count() or 1))
    specs = [P.c2_cube(seed=1, packets_scale=0.15, steps=2).with_seed(1 + t) for t in range(threads)]
    pyoracle.lib()  # build / load outside the timed region
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        runs = list(ex.map(pyoracle.run, specs))
    wall = time.perf_counter() - t0
    seg = float(sum(r["ledger"]["segments"].sum() for r in runs))
    log(f"cpu baseline: {seg:.0f} segments on {threads} threads in {wall:.2f} s")
    return dict(value=seg / wall, unit="segments/s", cores=threads, kind="port",
                sample=f"64^3 cube, first 2 IMC steps at 3/20 of the packets (~1e6 per step), FP64 oracle, "
                       f"{threads} independent seeds on {threads} threads")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--schedule", choices=("pass", "refill"), default="pass",
                    help="transport schedule of the mixed engine")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the run, one plain copy kernel over this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(packets would shard across GPUs with one integer all-reduce of the tallies per step)",
              file=sys.stderr)
        sys.exit(2)

    def log(msg):
        if not args.quiet:
            print(msg, file=sys.stderr, flush=True)

    from storm_amd import capi
    from storm_amd.engine import device_info

    log("device: " + device_info())
    sched = capi.SCHEDULE_PASS if args.schedule == "pass" else capi.SCHEDULE_REFILL
    log(f"mixed engine ({args.schedule} schedule)")
    mixed = run_engine(capi.MIXED, args.steps, args.warmup, args.seed, log, schedule=sched)
    value = mixed["segments"] / mixed["wall"]
    ms_per_step = mixed["wall"] / args.steps * 1e3
    achieved = mixed["segments"] * B_SEG_MIXED / mixed["transport_s"] / 1e9
    traffic, traffic_src = measured_traffic() if args.schedule == "pass" else (None, None)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "segments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "c2_cube: 64^3 gray IMC, sigma_a=sigma0*max(T,0.1keV)^-3, ~1e7 packets/step, dt=1e-11 s",
            "cells": mixed["N"],
            "packets_per_step": mixed["packets"] / args.steps,
            "segments_per_step": mixed["segments"] / args.steps,
            "arithmetic": "f32 cell-local tracking, 64-bit fixed-point tallies, f64 material update",
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": "k_transport_w6<float>" if args.schedule == "pass" else "k_transport_refill<float, 0u>",
            "bytes_per_segment": B_SEG_MIXED,
        },
    }
    if not args.no_fp64:
        log("plain FP64 engine (one packet per thread)")
        f64 = run_engine(capi.FP64, args.steps, args.warmup, args.seed, log)
        out["fp64_gpu"] = {
            "value": f64["segments"] / f64["wall"],
            "ms_per_step": f64["wall"] / args.steps * 1e3,
            "transport_segments_per_s": f64["segments"] / f64["transport_s"],
            "form": "plain: FP64 tracking and tallies, one packet per thread",
        }
        out["speedup_vs_fp64"] = value / out["fp64_gpu"]["value"]

    out["transport_segments_per_s"] = mixed["segments"] / mixed["transport_s"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(log)
    if args.copy_mib > 0:
        copy_calibration(args.copy_mib, log)
    print(json.dumps(out))


def copy_calibration(mib, log):
    """One run of the library's plain copy kernel over `mib` MiB (reads and
    writes each byte once: a kernel of known traffic, to scale the HBM byte
    counters of a profiler run)."""
    from storm_amd.engine import calibration_copy
    ms = calibration_copy(mib << 20)
    log(f"calibration copy: {mib} MiB in, {mib} MiB out, {ms:.3f} ms ({2 * (mib << 20) / ms / 1e6:.0f} GB/s)")


if __name__ == "__main__":
    main()
