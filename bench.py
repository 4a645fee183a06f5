#!/usr/bin/env python3
"""Benchmark of the SimpleNavierStokesSolver time-step hot path on one MI355X.

Workload: the large single instance, 4096 x 8192 cells, uniform grid, lx=1,
ly=2, Ra=1e11, Pr=0.7, FP64, Crank-Nicolson diffusion in x and explicit in y,
seeded temperature noise of amplitude 1e-3 (u = 0 initially).  One step is
what the reference's time loop does per iteration (src/main.c:28-35): the
time-step decision (src/param/decide_dt.c, evaluated on the device from the
advective bound that the previous step's last correction reduced) and one
three-substage RK3 integration.  Inputs are resident in HBM before the timed
region starts.

Printed JSON (one line): value = cell-steps/s, plus
  roofline      -- the dominant kernel, timed with HIP events on the
                   instance's stream during the timed steps, against the
                   8 TB/s HBM peak, using that kernel's algorithmic bytes;
  step_roofline -- the whole step against 520 B/cell-step of compulsory
                   traffic (65 FP64 words per cell and step, with the
                   corrections of substages 1 and 2 applied by the next
                   predictor; 600 B with a correction kernel per substage);
  cpu_baseline  -- the reference algorithm restated in C (the oracle) with
                   the reference's y-slab MPI decomposition, one MPICH rank
                   per available CPU, timed on a bounded sample of the same
                   workload: --cpu-steps steps of the same grid from the
                   same initial state.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def bytes_per_cell_step(implicitx, implicity, folded=True):
    """Compulsory HBM bytes per cell and RK3 step: 75 FP64 words with a
    correction kernel after every substage, plus 27 for the implicit-y solve
    (write, read the increment, read the field, per field and substage).
    With implicit x and explicit y (folded) the predictors of substages 2 and
    3 apply the previous substage's correction as they load the fields: the
    two corrections (7 words each) give way to a psi read and a pressure
    write in each of those predictors, 65 words."""
    if implicity:
        return 816.0
    return 520.0 if (implicitx and folded) else 600.0


def mix_roofline(value, jobs_weights):
    """HBM roofline of a mix of jobs: jobs_weights = [(job, cell-steps)];
    the cell-step-weighted bytes per cell-step times the measured rate."""
    tot = sum(w for _, w in jobs_weights)
    if tot <= 0:
        return None
    b = sum(bytes_per_cell_step(j["implicitx"], j["implicity"]) * w
            for j, w in jobs_weights) / tot
    ach = value * b / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "bytes_per_cell_step": b}

# Algorithmic (compulsory) HBM bytes per interior cell for one launch of each
# kernel of the step: each array the kernel must read or write, once.
KERNEL_BYTES_PER_CELL = {
    # fused predictor: reads ux, uy, p, T and (substages 1, 2) the three RK
    # histories, writes T', ux*, uy* and (substages 0, 1) the histories:
    # 10, 13, 10 words over the three substages, 11 on average; with the
    # folded corrections substages 1 and 2 also read psi and write p
    # (10, 15, 12 words: pred_bytes_per_cell)
    "pred": 8 * 11,
    "poisson_fwd": 8 * 3,      # ux*, uy* -> spectrum (jtot/2+1 complex per column)
    "poisson_x": 8 * 2,        # spectrum in place
    "poisson_inv": 8 * 2,      # spectrum -> psi
    "corr": 8 * 7,             # ux*, uy*, psi, p -> ux, uy, p
    "dt_reduce": 8 * 2,        # ux, uy
    "dt": 0,
}
KERNEL_NAMES = list(KERNEL_BYTES_PER_CELL.keys())


def kernel_bytes_per_cell(kernel, folded):
    """KERNEL_BYTES_PER_CELL, the predictor's average over its three
    launches of a step with the folded corrections when folded."""
    if kernel == "pred" and folded:
        return 8 * 37 / 3
    return KERNEL_BYTES_PER_CELL[kernel]

WORKLOAD = {
    "workload": "rayleigh_benard_2d_large_single_instance",
    "itot": 4096, "jtot": 8192, "lx": 1.0, "ly": 2.0, "grid": "uniform",
    "Ra": 1e11, "Pr": 0.7, "implicitx": 1, "implicity": 0,
    "ic": "u=0, T=1e-3*(2U-1) seeded (seed 3)", "seed": 3, "amplitude": 1e-3,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--itot", type=int, default=WORKLOAD["itot"])
    ap.add_argument("--jtot", type=int, default=WORKLOAD["jtot"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--per-kernel", action="store_true", help="print per-kernel timings to stderr")
    ap.add_argument("--ensemble-steps", type=int, default=300,
                    help="steps per job of the mixed ensemble (0: skip it)")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="length of the served request stream (0: skip it)")
    ap.add_argument("--serve-slots", type=int, default=16,
                    help="most jobs on the GPU at once in the served stream (0: memory budget "
                         "alone)")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="instead of the step benchmark: time a plain device copy of this "
                         "many GiB (the known-size kernel that calibrates traffic counters)")
    return ap.parse_args()


def make_instance(sns, itot, jtot):
    w = WORKLOAD
    inst = sns.Instance(itot, jtot, ly=w["ly"], stretch=float(itot), Ra=w["Ra"], Pr=w["Pr"],
                        implicitx=w["implicitx"], implicity=w["implicity"])
    inst.state_alloc()
    # seeded noise and the wall / halo boundary values, on the device
    inst.state_init_temperature(w["seed"], w["amplitude"])
    return inst


def step(sns, inst, n=1):
    """n iterations of the reference's time loop body: device-side time-step
    decision (src/param/decide_dt.c for the fixed treatment) + RK3 step."""
    inst.advance(n)


def per_kernel_times(sns, inst, steps):
    """Average duration of each kernel over `steps` steps (HIP events)."""
    out = {}
    for name in KERNEL_NAMES:
        inst.timing_enable(name)
        step(sns, inst, steps)
        n, ms = inst.timing_read()
        inst.timing_disable()
        if n:
            out[name] = (n, ms)
    return out


def host_cpu():
    """CPU model, logical CPUs (all, and those this process may run on) and
    physical cores of the host, from /proc/cpuinfo."""
    model, cores = None, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count(),
            "cpus_available": len(os.sched_getaffinity(0)),
            "physical_cores": len(cores) or None}


def cpu_baseline(itot, jtot, nsteps):
    """The reference algorithm on the host cores, on `nsteps` steps of the
    workload: the oracle's y-slab MPI restatement (the reference's
    decomposition: rows split over the ranks, halo exchanges, MPI_Alltoallv
    transposes for the y solves and the Poisson transforms; one rank per
    available CPU, one thread each) under MPICH, or, where MPI is missing,
    the single-rank oracle with OpenMP threads.  The reference itself cannot
    be built here: its FFTW3 dependency is absent, so the transforms are the
    oracle's recursive mixed-radix FFT with FFTW's r2c/c2r conventions."""
    from oracle import pyoracle as po
    w = WORKLOAD
    host = host_cpu()
    ranks = host["cpus_available"]
    fft = "the oracle's recursive mixed-radix FFT (FFTW3 is absent from the image)"
    if os.path.exists(po.MPI_EXE) and po.mpiexec():
        summ, _ = po.mpi_run(ranks, itot, jtot, w["Ra"], w["Pr"], w["implicitx"],
                             w["implicity"], w["seed"], w["amplitude"], nsteps,
                             stretch=float(itot), timeout=600)
        res = {"value": summ["cell_steps_per_s"], "unit": "cell-steps/s", "cores": ranks,
               "kind": "port",
               "sample": "%d RK3 step(s) (time-step decision + 3 substages) of the same %dx%d "
                         "workload from the same initial state, y-slab MPI restatement of the "
                         "reference algorithm, %d MPICH ranks x 1 thread, step loop timed by "
                         "MPI_Wtime" % (nsteps, itot, jtot, ranks),
               "ranks": ranks, "threads_per_rank": 1, "steps": nsteps,
               "seconds": summ["seconds"], "fft": fft,
               "decomposition": "y slabs of the reference (src/parallel/others.c), halo "
                                "exchanges, MPI_Alltoallv transposes"}
    else:
        o = po.Oracle(itot, jtot, ly=w["ly"], stretch=float(itot), Ra=w["Ra"], Pr=w["Pr"])
        o.set_treatment(w["implicitx"], w["implicity"])
        o.init_temperature(w["seed"], w["amplitude"], linear=False)
        o.apply_boundaries()
        t = time.perf_counter()
        o.run(nsteps)
        el = time.perf_counter() - t
        nthr = po.num_threads()
        res = {"value": itot * jtot * nsteps / el, "unit": "cell-steps/s", "cores": nthr,
               "kind": "port",
               "sample": "%d RK3 step(s) (time-step decision + 3 substages) of the same %dx%d "
                         "workload from the same initial state, CPU oracle (C, gcc -O2, OpenMP "
                         "%d threads)" % (nsteps, itot, jtot, nthr),
               "threads": nthr, "steps": nsteps, "seconds": el, "fft": fft,
               "decomposition": "one rank (MPI not available), OpenMP over rows and columns"}
    res.update(host)
    return res


COPY_SHAPES = [(0, False)] + [(u, nt) for nt in (False, True) for u in (1, 2, 4, 8)]


def copy_sweep(sns, gib, reps=5):
    """Device copies of `gib` GiB: the plain grid-stride kernel (loads 0,
    the one that calibrates the traffic counters) and the tiled kernel at
    each shape, `reps` timed launches after one warm-up each; read+write
    GB/s per shape and the best, the practical copy ceiling."""
    nbytes = int(gib * (1 << 30)) // 16 * 16
    inst = sns.Instance(64, 64)
    src, dst = inst.malloc(nbytes), inst.malloc(nbytes)
    rows = []
    for loads, nt in COPY_SHAPES:
        inst.copy(dst, src, nbytes, loads=loads, nontemporal=nt)
        inst.sync()
        inst.timing_enable("copy")
        for _ in range(reps):
            inst.copy(dst, src, nbytes, loads=loads, nontemporal=nt)
        inst.sync()
        n, ms = inst.timing_read()
        inst.timing_disable()
        avg = ms / max(n, 1)
        rows.append({"loads_per_thread": loads, "nontemporal": nt, "avg_launch_ms": avg,
                     "GB/s": 2 * nbytes / (avg * 1e-3) / 1e9})
    inst.close()
    best = max(rows, key=lambda r: r["GB/s"])
    return {"bytes_read": nbytes, "bytes_written": nbytes, "launches_per_shape": reps,
            "shapes": rows, "best": best, "GB/s": best["GB/s"]}


def copy_calibration(sns, gib, reps=5):
    """Prints the copy sweep as one JSON line."""
    res = copy_sweep(sns, gib, reps)
    print(json.dumps(dict({"metric": "device copy bandwidth"}, **res)))


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def treatments(sns, ens, jobs):
    """Treatment of every job by the reference's rule from the recorded
    cost table (measured now for sizes the table lacks)."""
    try:
        costs = ens.load_costs()
    except (OSError, ValueError, KeyError):
        costs = {}
    missing = sorted({j["nx"] for j in jobs if str(j["nx"]) not in costs})
    if missing:
        log("measuring treatment costs for nx %s" % missing)
        costs.update(ens.measure_costs(sns, missing))
    for j in jobs:
        j["implicitx"], j["implicity"] = ens.job_treatment(costs, j, sns.grid)
    return costs


def job_instance(sns, j):
    inst = sns.Instance(j["nx"], j["ny"], ly=2.0, stretch=float(j["nx"]), Ra=j["Ra"], Pr=j["Pr"],
                        implicitx=j["implicitx"], implicity=j["implicity"])
    inst.state_alloc()
    inst.state_init_temperature(j["seed"], j["amplitude"])
    return inst


def ensemble_run(sns, steps, warmup=2):
    """The mixed ensemble: its 64 jobs advanced one after another through
    snsg_advance (the plain single-instance path), then the same jobs from
    the same initial state as one batch (one launch per kernel and shape
    class, graph replay).  Both after `warmup` untimed steps."""
    from simplenavierstokessolver_amd import ensemble as ens
    jobs = ens.ensemble_jobs(steps=steps)
    treatments(sns, ens, jobs)
    cells = sum(j["nx"] * j["ny"] for j in jobs)
    log("ensemble: %d jobs, %d cells, plain path" % (len(jobs), cells))
    insts = [job_instance(sns, j) for j in jobs]
    for i in insts:
        i.advance(warmup)
    for i in insts:
        i.sync()
    t = time.perf_counter()
    for i in insts:
        i.advance(steps)
        i.sync()
    t_plain = time.perf_counter() - t
    for i in insts:
        i.close()
    log("ensemble: batched")
    insts = [job_instance(sns, j) for j in jobs]
    batch = sns.Batch(insts)
    batch.advance(warmup)
    batch.sync()
    t = time.perf_counter()
    batch.advance(steps)
    batch.sync()
    t_batch = time.perf_counter() - t
    nclass = batch.classes()
    batch.close()
    for i in insts:
        i.close()
    mix = {}
    for j in jobs:
        k = ens.TREATMENTS[(j["implicitx"], j["implicity"])]
        mix[k] = mix.get(k, 0) + 1
    v_b, v_p = cells * steps / t_batch, cells * steps / t_plain
    weights = [(j, j["nx"] * j["ny"] * steps) for j in jobs]
    return {"config": ens.ENSEMBLE, "value": v_b, "unit": "cell-steps/s",
            "roofline": mix_roofline(v_b, weights),
            "plain_roofline": mix_roofline(v_p, weights),
            "ms_per_ensemble_step": 1e3 * t_batch / steps,
            "plain_value": v_p, "plain_ms_per_ensemble_step": 1e3 * t_plain / steps,
            "speedup_vs_plain": v_b / v_p, "shape_classes": nclass, "treatments": mix,
            "cells": cells, "steps": steps, "warmup": warmup,
            "plain": "each job in turn through snsg_advance (fused single-instance path)",
            "batched": "snsg_batch: one launch per kernel and shape class, a stream and a "
                       "replayed HIP graph per class"}


def served_run(sns, window, slots=0):
    """The served request stream for `window` seconds of arrivals: jobs are
    submitted at their arrival times to the serving engine (200 GB budget);
    throughput and per-step latency over the job-steps completed in the
    window.  Jobs still queued or running at its end are dropped."""
    from simplenavierstokessolver_amd import ensemble as ens
    jobs = ens.served_jobs(window)
    treatments(sns, ens, jobs)
    log("served: %d arrivals over %.0f s" % (len(jobs), window))
    srv = sns.Server(int(ens.SERVED["budget_bytes"]), max_active=slots)
    srv.reset_stats()
    t0 = time.perf_counter()
    k = 0
    next_log = 5.0
    while True:
        now = time.perf_counter() - t0
        if now >= window:
            break
        while k < len(jobs) and jobs[k]["arrival"] <= now:
            j = jobs[k]
            j["id"] = srv.submit(j["nx"], j["ny"], j["steps"], Ra=j["Ra"], Pr=j["Pr"],
                                 implicitx=j["implicitx"], implicity=j["implicity"],
                                 seed=j["seed"], amplitude=j["amplitude"])
            k += 1
        if now >= next_log:
            st = srv.stats()
            log("served: t=%.0f s submitted %d admitted %d completed %d job-steps %d" %
                (now, st["submitted"], st["admitted"], st["completed"], st["steps"]))
            next_log += 5.0
        nxt = jobs[k]["arrival"] if k < len(jobs) else window
        time.sleep(max(0.0, min(0.005, nxt - now)))
    st = srv.stats()
    weights = [(j, srv.status(j["id"])[1] * j["nx"] * j["ny"]) for j in jobs[:k]]
    srv.close()
    mix = {}
    for j in jobs[:k]:
        key = ens.TREATMENTS[(j["implicitx"], j["implicity"])]
        mix[key] = mix.get(key, 0) + 1
    value = st["cell_steps"] / st["busy_s"] if st["busy_s"] > 0 else 0.0
    return {"config": ens.SERVED, "window_s": window, "value": value, "unit": "cell-steps/s",
            "roofline": mix_roofline(value, weights),
            "p50_step_ms": st["p50_ms"], "p99_step_ms": st["p99_ms"], "max_step_ms": st["max_ms"],
            "submitted": st["submitted"], "admitted": st["admitted"],
            "completed": st["completed"], "job_steps": st["steps"], "lanes": st["lanes"],
            "slots": slots, "active_peak": st["active_peak"],
            "mem_peak_GB": st["mem_peak"] / 1e9, "treatments": mix,
            "latency": "device duration of the lane step that advanced the job-step"}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_4096x8192.json")


def measured_traffic(group, cells):
    """HBM bytes per launch of the kernels of `group` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("cells") != cells:
        return None
    rows = [r for r in t["kernels"] if r["group"] == group]
    return sum(r["traffic_bytes"] for r in rows) if rows else None


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: the solver runs on one GPU; --gpus %d is not supported "
              "(a y-slab sharding over GPUs is designed, not implemented)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    import simplenavierstokessolver_amd as sns
    if a.copy_gib > 0:
        copy_calibration(sns, a.copy_gib)
        return
    itot, jtot = a.itot, a.jtot
    cells = itot * jtot
    # launch shapes from the autotuner's table (every instance below, the
    # ensemble's and the served jobs' included, is created after this)
    ntuned = sns.load_tuning() if os.path.exists(sns.TUNING_FILE) else 0
    inst = make_instance(sns, itot, jtot)
    launch = inst.launch()
    w = WORKLOAD
    folded = bool(w["implicitx"] and not w["implicity"] and not launch["split_corr"])
    b_step = bytes_per_cell_step(w["implicitx"], w["implicity"], folded)
    step(sns, inst, a.warmup)
    inst.sync()
    # timed region, with the dominant-kernel events recorded live on the stream
    kt = per_kernel_times(sns, inst, 1)
    dominant = max(kt.items(), key=lambda kv: kv[1][1])[0] if kt else None
    inst.sync()
    if dominant:
        inst.timing_enable(dominant)
    t0 = time.perf_counter()
    step(sns, inst, a.steps)
    inst.sync()
    el = time.perf_counter() - t0
    nlaunch, kms = inst.timing_read() if dominant else (0, 0.0)
    inst.timing_disable()
    value = cells * a.steps / el
    res = {
        "metric": "FP64 cell-steps/s @4096x8192 (%HBM roofline); ensemble cell-steps/s + p99 "
                  "step lat",
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * el / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": dict(WORKLOAD, itot=itot, jtot=jtot, launch=launch,
                       launch_table_entries=ntuned),
    }
    if dominant and nlaunch:
        avg_s = kms / nlaunch * 1e-3
        bytes_launch = kernel_bytes_per_cell(dominant, folded) * cells
        ach = bytes_launch / avg_s / 1e9
        res["roofline"] = {"bound": "hbm", "kernel": dominant, "achieved": ach,
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                           "traffic": measured_traffic(dominant, cells),
                           "avg_launch_ms": kms / nlaunch,
                           "bytes_per_launch": bytes_launch}
    # the practical copy ceiling, measured in the same run
    ceiling = copy_sweep(sns, 1.0, reps=3)
    res["copy_ceiling"] = {"GB/s": ceiling["GB/s"], "best_shape": ceiling["best"],
                           "bytes": ceiling["bytes_read"],
                           "what": "best read+write bandwidth of a device copy over the "
                                   "shapes of the tiled copy kernel and the plain one"}
    if "roofline" in res:
        res["roofline"]["frac_of_copy"] = res["roofline"]["achieved"] / ceiling["GB/s"]
    step_gbs = value * b_step / 1e9
    res["step_roofline"] = {"bytes_per_cell_step": b_step, "achieved": step_gbs,
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS}
    if a.per_kernel:
        kt = per_kernel_times(sns, inst, 2)
        tot = sum(v[1] for v in kt.values())
        for k, (n, ms) in sorted(kt.items(), key=lambda kv: -kv[1][1]):
            gbs = kernel_bytes_per_cell(k, folded) * cells / max(ms / n * 1e-3, 1e-12) / 1e9
            print("%-12s launches %4d  avg %8.3f ms  share %5.1f%%  %7.0f GB/s" %
                  (k, n, ms / n, 100 * ms / tot, gbs), file=sys.stderr)
    inst.close()
    if a.ensemble_steps > 0:
        res["ensemble"] = ensemble_run(sns, a.ensemble_steps)
    if a.serve_seconds > 0:
        res["served"] = served_run(sns, a.serve_seconds, a.serve_slots)
    if not a.no_cpu_baseline:
        log("cpu baseline")
        res["cpu_baseline"] = cpu_baseline(itot, jtot, a.cpu_steps)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
