"""Benchmark: tracked segments/s and seconds per IMC step on one MI355X,
mixed precision against plain FP64 at matched per-cell tally error.

Workload (N = 1): the 64^3 gray IMC cube as stated (storm_amd.problems.c2_cube
with literal=True): sigma_a = sigma0 * T^-3 with sigma0 log-uniform in
[1, 100] per cell, cv proportional to T, T0 = 1e-3 keV, a 1 keV source on the
x- face, escape on every face, >= 1e7 packets per step, dt = 1e-11 s, 50 IMC
steps, the random walk on at the reference's default thresholds (the cold
cube is ~1e7 mean free paths per cell), census resident in HBM.  One "step" =
one full IMC step (preStep, sourcing, transport, comb, postStep).

Timed region: seed 1 of the mixed engine, W untimed steps then K timed steps
(`value` = segments/s over those K steps, wall clock around each host-
synchronised step).  Every run then continues, untimed by `value`, to the
problem's 50 steps: S seeds per form (mixed and FP64) give per-cell means and
variances of the time-averaged radiation energy density at step 50.  The
matched-error comparison:
  rel_err = median over cells of the one-run relative standard deviation,
  FOM     = 1 / (rel_err^2 * t_run), t_run = wall time of a 50-step run,
and the per-cell agreement of the two forms (|z| <= 3, pooled variance).

Roofline: no MFMA (nothing is a dense contraction), so the roofline is HBM
bandwidth against the metric's algorithmic bytes, 144 B per segment (a 48 B
packet read and written, a 16 B cell record, two 16 B tallies).  The mixed
engine runs two transport kernels (the pass and lane-refill schedules,
switched as the problem evolves); the roofline is that of the one that ran
the timed window: achieved = segments * 144 B / transport time over its
launches in this process (HIP events on the engine's stream, attributed by
the schedule each step ran), so its average launch duration is the one a
profiler's kernel statistics of this command report; `by_kernel` gives both.  The history kernel keeps a packet in registers for its whole
history, so its measured HBM traffic (rocprof FETCH_SIZE / WRITE_SIZE,
`traffic`) is far below the model.

Further lines of figures, a few steps each: `c3` runs the two-material step
slab (random walk and DDMC, 2e7 packets per step); `c4` the 64-group problem
on a Voronoi tessellation of 2e5 points (5e7 packets per step), mixed and
FP64; `c5` 256 independent slab instances as one batched mixed engine against
one FP64 engine per instance.  `cpu_baseline` times the FP64 CPU oracle on
host cores on a bounded sample of the C2 workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "tracked segments/s and s per IMC step, mixed vs FP64 GPU at matched tally error"
B_SEG_MIXED = 144.0  # bytes per segment: 2 * 48 B packet + 16 B cell + 2 * 16 B tally
B_SEG_FP64 = 224.0   # 2 * 80 B packet + 32 B cell + 2 * 16 B tally
HBM_PEAK_GBS = 8000.0
PROBLEM_STEPS = 50
# C2 packet knobs: 2400 boundary packets per x- face (9.8e6), >= 2 emitted per
# cell (5.2e5), comb floor 8 per cell: >= 1e7 packets in every step
C2_KNOBS = dict(source_per_face=2400, comb_nmin=8)
# HBM bytes per transport launch, measured by rocprofv3 (FETCH_SIZE and
# WRITE_SIZE, scaled by the same counters on a plain copy kernel) on this
# benchmark's own default command; the summary holds the figures and the command.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_transport_refill_c2_literal_counters.md")


# the mixed engine's Cartesian transport kernels with the random-walk / DDMC
# features, by schedule (storm_gpu_counters.transport_schedule)
KERNEL_NAME = {0: "k_transport<float, 1u, MeshView<float>>", 1: "k_transport_refill<float, 1u, MeshView<float>>"}
SCHEDULE_NAME = {0: "pass", 1: "refill"}


class KernelTotals:
    """Segments, transport seconds and launches per transport kernel."""

    def __init__(self):
        self.tot = {}

    def add(self, schedule, segments, seconds, launches):
        a = self.tot.setdefault(schedule, [0.0, 0.0, 0])
        a[0] += segments
        a[1] += seconds
        a[2] += launches

    def merge(self, other):
        for k, (sg, t, n) in other.tot.items():
            self.add(k, sg, t, n)

    def segments(self):
        return sum(a[0] for a in self.tot.values())

    def seconds(self):
        return sum(a[1] for a in self.tot.values())

    def names(self):
        return " + ".join(KERNEL_NAME[k] for k in sorted(self.tot))

    def report(self):
        return {KERNEL_NAME[k]: {"launches": n, "segments": sg, "seconds": t, "avg_launch_us": t / n * 1e6,
                                 "achieved_GBps": sg * B_SEG_MIXED / t / 1e9}
                for k, (sg, t, n) in sorted(self.tot.items())}


def measured_traffic(kernel):
    """(bytes per launch of `kernel`, source) from the committed counter
    summary (a line "traffic_bytes_per_launch <kernel>: <bytes>"), or (None, None)."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch") and kernel in line:
                    return float(line.rsplit(":", 1)[1]), os.path.relpath(TRAFFIC_PROFILE, ROOT)
    except OSError:
        pass
    return None, None


def c2_spec(seed):
    from storm_amd import problems as P
    return P.c2_cube(seed=seed, literal=True, **C2_KNOBS)


def run_seed(precision, seed, n_steps, snap_step, timed=None, log=None):
    """One run of the C2 cube; per-step wall times and counters, the cells at
    `snap_step`.  `timed` = (W, K): log the steps of that window."""
    from storm_amd.engine import GpuEngine

    spec = c2_spec(seed)
    steps = []
    snap = None
    with GpuEngine(spec, precision) as e:
        for s in range(n_steps):
            t0 = time.perf_counter()
            e.step()
            wall = time.perf_counter() - t0
            c = e.counters()
            steps.append(dict(wall=wall, segments=c["segments"], packets=c["start_particles"],
                              t_tr=c["seconds_transport"], launches=c["transport_launches"],
                              rw=c["random_walk_steps"], schedule=c["transport_schedule"]))
            if log and timed:
                W, K = timed
                tag = "warmup" if s < W else ("timed" if s < W + K else "untimed")
                log(f"seed {seed} step {s + 1}/{n_steps} ({tag}): {c['segments']} segments, {c['start_particles']} "
                    f"packets, {c['random_walk_steps']} walk steps, transport {c['seconds_transport'] * 1e3:.2f} ms "
                    f"({SCHEDULE_NAME[c['transport_schedule']]}), step {wall * 1e3:.2f} ms")
            if s + 1 == snap_step:
                snap = e.cells()
    return steps, snap


def ensemble(precision, seeds, n_steps, W, K, log, label):
    runs = []
    for i, sd in enumerate(seeds):
        t0 = time.perf_counter()
        steps, snap = run_seed(precision, sd, n_steps, PROBLEM_STEPS, timed=(W, K) if i == 0 else None, log=log)
        runs.append(dict(seed=sd, steps=steps, snap=snap))
        log(f"{label} seed {sd}: {n_steps} steps in {time.perf_counter() - t0:.2f} s, "
            f"{sum(s['segments'] for s in steps[:PROBLEM_STEPS]) / 1e9:.2f} G segments in the first {PROBLEM_STEPS}")
    return runs


def window(steps, lo, hi):
    seg = sum(s["segments"] for s in steps[lo:hi])
    wall = sum(s["wall"] for s in steps[lo:hi])
    return seg, wall


def matched_error(mixed, fp64):
    """Per-cell relative error of one run (seed-ensemble standard deviation
    over the mean) of the time-averaged radiation energy density at step 50,
    the FOM of each form and the agreement of the two ensembles.

    Cells: those the radiation has reached, mean above 1e-3 of the largest
    (in the cold cube a cell's own emission is absorbed in place, so its
    Erad is nearly deterministic and says nothing about Monte Carlo noise).
    The z of the two forms uses the pooled variance plus (1e-6 mean)^2, the
    representation error of FP32 opacities."""
    A = np.array([r["snap"]["erad"] for r in mixed])
    B = np.array([r["snap"]["erad"] for r in fp64]) if fp64 else None
    ref = B if B is not None else A
    mref = ref.mean(0)
    cells = mref > 1e-3 * mref.max()
    out = {}

    def rel(X):
        m = X.mean(0)[cells]
        sd = X.std(0, ddof=1)[cells]
        ok = m > 0
        return float(np.median(sd[ok] / m[ok])), int(ok.sum())

    t_m = float(np.mean([sum(s["wall"] for s in r["steps"][:PROBLEM_STEPS]) for r in mixed]))
    ra, na = rel(A)
    out["mixed"] = dict(rel_err=ra, cells=na, seconds_per_run=t_m, seconds_per_step=t_m / PROBLEM_STEPS,
                        fom=1.0 / (ra * ra * t_m))
    if B is not None:
        t_f = float(np.mean([sum(s["wall"] for s in r["steps"][:PROBLEM_STEPS]) for r in fp64]))
        rb, nb = rel(B)
        out["fp64"] = dict(rel_err=rb, cells=nb, seconds_per_run=t_f, seconds_per_step=t_f / PROBLEM_STEPS,
                           fom=1.0 / (rb * rb * t_f))
        out["fom_ratio"] = out["mixed"]["fom"] / out["fp64"]["fom"]
        from scipy import stats
        ma, mb = A.mean(0)[cells], B.mean(0)[cells]
        se2 = A.var(0, ddof=1)[cells] / len(A) + B.var(0, ddof=1)[cells] / len(B) + (1e-6 * np.maximum(ma, mb)) ** 2
        d = ma - mb
        z = np.where(se2 > 0, d / np.sqrt(np.where(se2 > 0, se2, 1.0)), np.where(d == 0, 0.0, np.inf))
        nu = len(A) + len(B) - 2
        out["cells_within_3sigma"] = float(np.mean(np.abs(z) <= 3.0))
        out["expected_within_3sigma"] = float(1.0 - 2.0 * stats.t.sf(3.0, nu))
        out["mean_z2"] = float(np.mean(z[np.isfinite(z)] ** 2))
        out["expected_mean_z2"] = nu / (nu - 2.0)
        out["max_abs_z"] = float(np.max(np.abs(z)))
    out["tally"] = "time-averaged radiation energy density per cell at step 50 (getEradTimeAvg)"
    out["cells"] = "cells with a mean above 1e-3 of the largest: " + str(int(cells.sum()))
    out["seeds"] = len(A)
    return out


def c3_line(W, K, log):
    """Two-material step slab (C3: random walk and DDMC, 2e7 packets per step)."""
    from storm_amd import capi
    from storm_amd import problems as P
    from storm_amd.engine import GpuEngine

    spec = P.c3_step(seed=1)
    seg = wall = t_tr = 0.0
    rw = ddmc = packets = launches = 0
    by_kernel = KernelTotals()
    with GpuEngine(spec, capi.MIXED) as e:
        for s in range(W + K):
            t0 = time.perf_counter()
            e.step()
            dtw = time.perf_counter() - t0
            c = e.counters()
            by_kernel.add(c["transport_schedule"], c["segments"], c["seconds_transport"], c["transport_launches"])
            log(f"c3 step {s + 1}/{W + K}: {c['segments']} segments, {c['start_particles']} packets, "
                f"{c['random_walk_steps']} walk, {c['ddmc_steps']} DDMC, step {dtw * 1e3:.2f} ms")
            if s >= W:
                seg += c["segments"]
                wall += dtw
                t_tr += c["seconds_transport"]
                rw += c["random_walk_steps"]
                ddmc += c["ddmc_steps"]
                packets += c["start_particles"]
                launches += c["transport_launches"]
    # every transport launch of the run (launches of the C2 cube's kernels,
    # counted in its roofline)
    return dict(launch_totals=by_kernel,
                workload="c3_step: 256x16x16 two-material slab, sigma_a 1 | 1e4 (+ sigma_s 1e4), RW + DDMC, "
                         "~2e7 packets/step, dt=1e-11 s",
                value=seg / wall, unit="segments/s", steps=K, warmup=W, ms_per_step=wall / K * 1e3,
                packets_per_step=packets / K, transport_segments_per_s=seg / t_tr,
                random_walk_steps_per_step=rw / K, ddmc_steps_per_step=ddmc / K,
                kernels=by_kernel.names() + " (random walk + DDMC features)",
                roofline_frac=seg * B_SEG_MIXED / t_tr / 1e9 / HBM_PEAK_GBS)


def _run_steps(spec, precision, n, W, log, label, schedule=None):
    """n steps of one engine (in the engine's default transport schedule, or
    `schedule`); segments, wall, transport seconds and packets of the steps
    after the first W."""
    from storm_amd.engine import GpuEngine

    tot = dict(segments=0, wall=0.0, transport=0.0, packets=0, rw=0, ddmc=0, steps=0)
    with GpuEngine(spec, precision) as e:
        if schedule is not None:
            e.set_tuning(block=256, schedule=schedule)
        for s in range(n):
            t0 = time.perf_counter()
            e.step()
            dtw = time.perf_counter() - t0
            c = e.counters()
            if log:
                log(f"{label} step {s + 1}/{n}: {c['segments']} segments, {c['start_particles']} packets, "
                    f"{c['random_walk_steps']} walk, {c['ddmc_steps']} DDMC, transport "
                    f"{c['seconds_transport'] * 1e3:.2f} ms, step {dtw * 1e3:.2f} ms")
            if s >= W:
                tot["segments"] += c["segments"]
                tot["wall"] += dtw
                tot["transport"] += c["seconds_transport"]
                tot["packets"] += c["start_particles"]
                tot["rw"] += c["random_walk_steps"]
                tot["ddmc"] += c["ddmc_steps"]
                tot["steps"] += 1
    return tot


def c4_line(W, K, log, fp64=True):
    """Multigroup IMC on a Voronoi tessellation (C4: 2e5 cells, 64 groups,
    ~5e7 packets per step), mixed engine and, beside it, the FP64 engine."""
    from storm_amd import capi
    from storm_amd import problems as P

    t0 = time.perf_counter()
    mesh = P.voronoi_mesh(P.voronoi_points(200000, 0xC4))
    log(f"c4 mesh: {mesh.ncells} cells, {mesh.nfaces} faces in {time.perf_counter() - t0:.2f} s")
    spec = P.c4_voronoi(seed=1, mesh=mesh)
    m = _run_steps(spec, capi.MIXED, W + K, W, log, "c4 mixed")
    out = dict(workload="c4_voronoi: Voronoi tessellation of 2e5 uniform points in the unit cube (flattened face "
                        "tables), 64 log-spaced groups 0.01-20 keV, sigma = 10 nu^-3 T^-1/2, rho = 1 + 0.5 sin(2 pi "
                        "x), ~5e7 packets/step, dt=1e-11 s",
               value=m["segments"] / m["wall"], unit="segments/s", steps=K, warmup=W,
               ms_per_step=m["wall"] / K * 1e3, packets_per_step=m["packets"] / K,
               segments_per_step=m["segments"] / K, transport_segments_per_s=m["segments"] / m["transport"],
               faces=int(mesh.nfaces), kernel="k_transport<float, 2u, UMeshView<float>> (multigroup)")
    if fp64:
        f = _run_steps(spec, capi.FP64, W + K, W, log, "c4 fp64")
        out["fp64_gpu"] = dict(value=f["segments"] / f["wall"], ms_per_step=f["wall"] / K * 1e3,
                               transport_segments_per_s=f["segments"] / f["transport"])
        out["speedup_vs_fp64"] = out["value"] / out["fp64_gpu"]["value"]
    return out


def c5_line(W, K, log):
    """Batched independent instances (C5: 256 1-D slabs of 50-2000 cells and
    1e4-1e6 packets per step): the batch as one mixed engine against one FP64
    engine per instance, stepped one after another."""
    from storm_amd import capi
    from storm_amd import problems as P

    inst = P.c5_instances()
    t0 = time.perf_counter()
    batch = P.c5_batch(inst, seed=1, steps=W + K)
    log(f"c5 batch: {len(inst)} instances, {batch.umesh.ncells} cells, built in {time.perf_counter() - t0:.2f} s")
    b = _run_steps(batch, capi.MIXED, W + K, W, log, "c5 batched mixed")
    per = dict(segments=0, wall=0.0, transport=0.0, packets=0, rw=0, ddmc=0)
    t0 = time.perf_counter()
    for i, d in enumerate(inst):
        # the pass schedule: faster than the FP64 engine's default lane
        # refill on these small slabs (16.0 against 14.8 G segments/s)
        r = _run_steps(P.c5_instance_spec(d, seed=1 + i, steps=W + K), capi.FP64, W + K, W, None, "",
                       schedule=capi.SCHEDULE_PASS)
        for k in per:
            per[k] += r[k]
    log(f"c5 per-instance fp64: {len(inst)} engines x {W + K} steps in {time.perf_counter() - t0:.2f} s "
        f"(timed steps {per['wall']:.3f} s)")
    return dict(workload=f"c5_batch: {len(inst)} independent 1-D slab gray IMC instances (cells log-uniform 50-2000, "
                         "packets log-uniform 1e4-1e6 per step, sigma_a log-uniform 0.1-1e3, boundary 0.3-1.5 keV, "
                         "DDMC in thick cells) as one engine",
                value=b["segments"] / b["wall"], unit="segments/s", steps=K, warmup=W,
                ms_per_step=b["wall"] / K * 1e3, cells=int(batch.umesh.ncells), packets_per_step=b["packets"] / K,
                ddmc_steps_per_step=b["ddmc"] / K, transport_segments_per_s=b["segments"] / b["transport"],
                per_instance_fp64=dict(value=per["segments"] / per["wall"], ms_per_step=per["wall"] / K * 1e3,
                                       packets_per_step=per["packets"] / K,
                                       form="one FP64 engine per instance, each stepped on its own"),
                step_time_ratio=per["wall"] / b["wall"],
                speedup_vs_per_instance_fp64=(b["segments"] / b["wall"]) / (per["segments"] / per["wall"]))


def cpu_baseline(W, log):
    """FP64 oracle on the host: steps W+1, W+2 of the same workload at 1/100 of
    the packets, timed from the oracle's own per-step clock; one run on one
    thread, then one independent seed per thread on up to 16 threads."""
    import concurrent.futures as cf
    from oracle import pyoracle
    from storm_amd import problems as P

    n_steps = W + 2

    def spec(sd):
        return P.c2_cube(seed=sd, literal=True, packets_scale=0.01, **C2_KNOBS)

    def timed(out):
        led = out["ledger"]
        return float(led["segments"][W:].sum()), float(led["seconds"][W:].sum())

    pyoracle.lib()  # build / load outside the timed region
    s1, t1 = timed(pyoracle.run(spec(1), steps=n_steps))
    log(f"cpu baseline, 1 thread: {s1:.0f} segments in {t1:.2f} s")
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        runs = list(ex.map(lambda sd: pyoracle.run(spec(sd), steps=n_steps), range(1, threads + 1)))
    wall = time.perf_counter() - t0
    seg = sum(timed(r)[0] for r in runs)
    busy = max(sum(r["ledger"]["seconds"]) for r in runs)
    log(f"cpu baseline, {threads} threads: {seg:.0f} segments of steps {W + 1}-{W + 2}; batch wall {wall:.2f} s")
    # aggregate: timed steps' segments over the slowest thread's timed seconds
    slowest = max(timed(r)[1] for r in runs)
    return dict(value=seg / slowest, unit="segments/s", cores=threads, kind="port",
                sample=f"64^3 cube as stated (sigma0 T^-3, random walk), steps {W + 1}-{W + 2} (the first timed GPU "
                       f"steps) at 1/100 of the packets, FP64 oracle, {threads} independent seeds on {threads} threads, "
                       f"timed by the oracle's per-step clock (slowest thread {busy:.1f} s for all steps)",
                one_core=dict(value=s1 / t1, cores=1, sample="seed 1 of the same sample on one thread"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seeds", type=int, default=16, help="seed ensemble per form for the matched-error comparison")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=256,
                    help="after the run, one plain copy kernel over this many MiB (scales a profiler's HBM byte "
                         "counters; 0: none)")
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
    W, K = args.warmup, args.steps
    n_steps = max(PROBLEM_STEPS, W + K)
    seeds = list(range(1, max(args.seeds, 2) + 1))
    log(f"mixed engine: {len(seeds)} seeds x {n_steps} steps (seed 1 timed at steps {W + 1}-{W + K})")
    mixed = ensemble(capi.MIXED, seeds, n_steps, W, K, log, "mixed")
    seg, wall = window(mixed[0]["steps"], W, W + K)
    value = seg / wall
    by_kernel = KernelTotals()
    for r in mixed:
        for st in r["steps"]:
            by_kernel.add(st["schedule"], st["segments"], st["t_tr"], st["launches"])
    c2_transport_rate = by_kernel.segments() / by_kernel.seconds()
    c3 = None
    if not args.no_c3:
        c3 = c3_line(2, 3, log)
        by_kernel.merge(c3.pop("launch_totals"))
    # the roofline of the kernel that ran the timed window (the schedule
    # watch may switch kernels later in a run, at steps that depend on timing)
    win = {}
    for st in mixed[0]["steps"][W:W + K]:
        win[st["schedule"]] = win.get(st["schedule"], 0.0) + st["t_tr"]
    dom = max(win, key=win.get)
    tr_seg, tr_t, tr_n = by_kernel.tot[dom]
    achieved = tr_seg * B_SEG_MIXED / tr_t / 1e9
    traffic, traffic_src = measured_traffic(KERNEL_NAME[dom])
    packets = sum(s["packets"] for s in mixed[0]["steps"][W:W + K]) / K
    seg50, wall50 = window(mixed[0]["steps"], 0, PROBLEM_STEPS)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "segments/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "c2_cube: 64^3 gray IMC, sigma_a=sigma0*T^-3 (no floor), random walk, >=1e7 packets/step, "
                        "dt=1e-11 s, 50 steps",
            "cells": 64 ** 3,
            "packets_per_step": packets,
            "segments_per_step": seg / K,
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
            "kernel": KERNEL_NAME[dom],
            "bytes_per_segment": B_SEG_MIXED,
            "segments": tr_seg,
            "launches": tr_n,
            "avg_launch_us": tr_t / tr_n * 1e6,
            "over": f"every launch of {KERNEL_NAME[dom]} in this process (the C2 cube's {len(seeds)} seeds x "
                    f"{n_steps} steps" + ("" if c3 is None else " and the c3 line's 5 steps") + "), the kernel that "
                    "ran the timed window",
            "by_kernel": by_kernel.report(),
        },
        "transport_segments_per_s": c2_transport_rate,
        "full_problem": {
            "steps": PROBLEM_STEPS,
            "segments_per_s": seg50 / wall50,
            "ms_per_step": wall50 / PROBLEM_STEPS * 1e3,
            "packets_per_step": sum(s["packets"] for s in mixed[0]["steps"][:PROBLEM_STEPS]) / PROBLEM_STEPS,
            "min_packets_per_step_after_first": min(s["packets"] for s in mixed[0]["steps"][1:PROBLEM_STEPS]),
        },
    }
    fp64 = None
    if not args.no_fp64:
        log(f"plain FP64 engine: {len(seeds)} seeds x {n_steps} steps")
        fp64 = ensemble(capi.FP64, seeds, n_steps, W, K, log, "fp64")
        fseg, fwall = window(fp64[0]["steps"], W, W + K)
        out["fp64_gpu"] = {
            "value": fseg / fwall,
            "ms_per_step": fwall / K * 1e3,
            "transport_segments_per_s": sum(s["segments"] for r in fp64 for s in r["steps"]) /
                                        sum(s["t_tr"] for r in fp64 for s in r["steps"]),
            "form": "plain: FP64 tracking and tallies, device libm",
        }
        out["speedup_vs_fp64"] = value / out["fp64_gpu"]["value"]
    out["matched_error"] = matched_error(mixed, fp64)
    if c3 is not None:
        out["c3"] = c3
    if not args.no_c4:
        out["c4"] = c4_line(1, 2, log, fp64=not args.no_fp64)
    if not args.no_c5:
        out["c5"] = c5_line(1, 3, log)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(W, log)
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
