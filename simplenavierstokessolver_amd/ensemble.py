"""Many Rayleigh-Benard runs on one GPU: the mixed ensemble (batched form)
and the served request stream (serving engine), their job generators, and
the choice of each job's diffusive treatment by the reference's rule.

The treatment rule is src/param/decide_dt.c:68-97: for every combination of
explicit/implicit diffusion in x and y, the time step it allows (the
advective bound, and the diffusive bounds of the explicit directions)
gives the wall time per unit of simulated time, cost[implicity][implicitx]
/ dt; the cheapest wins, the first in the loop order on a tie.  The cost
table (seconds per step of each treatment, per grid size) is measured on
the GPU by `measure_costs`, as src/param/estimate_cost.c measures it for one
run, and recorded in tuning/treatment_costs_gfx950.json:

    python -m simplenavierstokessolver_amd.ensemble tune [--out FILE]

`trace` runs the mixed ensemble as one batch and brackets a few batched
steps with two marker copies, for rocprofv3 --kernel-trace and the
`timeline` summary of simplenavierstokessolver_amd.profiling:

    python -m simplenavierstokessolver_amd.ensemble trace [--steps N]
"""
import argparse
import json
import math
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
COSTS_FILE = os.path.join(HERE, "tuning", "treatment_costs_gfx950.json")

# 64 independent runs: nx drawn from {128, 256, 512, 1024}, ny = 2 nx, uniform
# grid, lx = 1, ly = 2, log10 Ra ~ U[6, 10], Pr ~ U[0.5, 4], temperature noise
# of amplitude 1e-3 (job k: seed 4000 + k), 300 steps each
ENSEMBLE = {"workload": "rayleigh_benard_2d_ensemble", "jobs": 64, "nx": [128, 256, 512, 1024],
            "ny": "2*nx", "log10_Ra": [6, 10], "Pr": [0.5, 4.0], "amplitude": 1e-3,
            "steps": 300, "job_seed": "4000+k", "draw_seed": 4}

# jobs arriving as a Poisson process of 20 jobs/s: nx drawn from
# {128, ..., 2048}, ny = 2 nx, log10 Ra ~ U[6, 11], Pr ~ U[0.7, 7], step counts
# uniform in 100..2000, admitted against a 200 GB device-memory budget
SERVED = {"workload": "rayleigh_benard_2d_served", "rate_per_s": 20.0,
          "nx": [128, 256, 512, 1024, 2048], "ny": "2*nx", "log10_Ra": [6, 11],
          "Pr": [0.7, 7.0], "steps": [100, 2000], "amplitude": 1e-3, "budget_bytes": 200e9,
          "draw_seed": 5, "job_seed": "5000+k"}

TREATMENTS = {(0, 0): "EXP-EXP", (1, 0): "IMP-EXP", (0, 1): "EXP-IMP", (1, 1): "IMP-IMP"}


def ensemble_jobs(n=ENSEMBLE["jobs"], seed=ENSEMBLE["draw_seed"], steps=ENSEMBLE["steps"]):
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n):
        nx = int(rng.choice(ENSEMBLE["nx"]))
        Ra = 10.0 ** rng.uniform(*ENSEMBLE["log10_Ra"])
        Pr = rng.uniform(*ENSEMBLE["Pr"])
        jobs.append({"nx": nx, "ny": 2 * nx, "Ra": Ra, "Pr": Pr, "seed": 4000 + k,
                     "steps": steps, "amplitude": ENSEMBLE["amplitude"]})
    return jobs


def served_jobs(duration, rate=SERVED["rate_per_s"], seed=SERVED["draw_seed"]):
    """Arrivals in [0, duration): exponential gaps of mean 1/rate."""
    rng = np.random.default_rng(seed)
    jobs, t, k = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return jobs
        nx = int(rng.choice(SERVED["nx"]))
        jobs.append({"arrival": t, "nx": nx, "ny": 2 * nx,
                     "Ra": 10.0 ** rng.uniform(*SERVED["log10_Ra"]),
                     "Pr": rng.uniform(*SERVED["Pr"]),
                     "steps": int(rng.integers(SERVED["steps"][0], SERVED["steps"][1] + 1)),
                     "seed": 5000 + k, "amplitude": SERVED["amplitude"]})
        k += 1


def diffusive_bounds(dxc, dy, Ra, Pr, lx=1.0):
    """dt_dif_x, dt_dif_y of src/param/decide_dt.c:48-66 (safety factor 0.9)."""
    diffusivity = max(math.sqrt(Pr) / math.sqrt(Ra), 1. / math.sqrt(Pr) / math.sqrt(Ra))
    dx = lx
    for v in dxc[1:-1]:   # DXC(i), i = 2..itot
        dx = min(dx, v)
    return 0.25 / diffusivity * dx ** 2.0 * 0.9, 0.25 / diffusivity * dy ** 2.0 * 0.9


def select_treatment(cost, dt_adv, dt_dif_x, dt_dif_y):
    """src/param/decide_dt.c:68-97: cost[implicity][implicitx] in seconds per
    step -> (implicitx, implicity, dt)."""
    best, choice = float("inf"), None
    for implicity in (0, 1):
        for implicitx in (0, 1):
            if implicitx:
                d = dt_adv if implicity else min(dt_adv, dt_dif_y)
            else:
                d = min(dt_adv, dt_dif_x) if implicity else min(dt_adv, min(dt_dif_x, dt_dif_y))
            wtime = cost[implicity][implicitx] * (1. / d)
            if wtime < best:
                best, choice = wtime, (implicitx, implicity, d)
    return choice


def job_treatment(costs, job, grid):
    """Treatment of a job from its size's cost row, at the start of the run
    (u = 0: the advective bound is its cap, 0.7 * 5e-2)."""
    g = grid(job["nx"], job["ny"], 2.0, float(job["nx"]))
    dx, dy = diffusive_bounds(g["dxc"], g["dy"], job["Ra"], job["Pr"])
    imx, imy, _ = select_treatment(costs[str(job["nx"])], 0.7 * 5.e-2, dx, dy)
    return imx, imy


def measure_costs(sns, sizes, steps=10, warmup=3):
    """Seconds per step of each treatment for each nx (ny = 2 nx), timed on
    the GPU with the single-instance path: {str(nx): [[EXP-EXP, IMP-EXP],
    [EXP-IMP, IMP-IMP]]} indexed [implicity][implicitx]."""
    out = {}
    for nx in sizes:
        row = [[0.0, 0.0], [0.0, 0.0]]
        for imy in (0, 1):
            for imx in (0, 1):
                inst = sns.Instance(nx, 2 * nx, stretch=float(nx), Ra=1e8, Pr=1.0,
                                    implicitx=imx, implicity=imy)
                inst.state_alloc()
                inst.state_init_temperature(1, 1e-3)
                inst.advance(warmup)
                inst.sync()
                t = time.perf_counter()
                inst.advance(steps)
                inst.sync()
                row[imy][imx] = (time.perf_counter() - t) / steps
                inst.close()
        out[str(nx)] = row
    return out


def load_costs(path=COSTS_FILE):
    with open(path) as f:
        return json.load(f)["seconds_per_step"]


def trace_batch(sns, steps=1, warmup=2):
    """The mixed ensemble as one batch, `steps` batched steps bracketed by
    two 16-byte device copies (marker dispatches of k_copy, each followed by
    a synchronisation) for a kernel trace; returns the host wall time of the
    bracketed steps in seconds."""
    jobs = ensemble_jobs()
    costs = load_costs()
    for j in jobs:
        j["implicitx"], j["implicity"] = job_treatment(costs, j, sns.grid)
    sns.load_tuning()
    insts = []
    for j in jobs:
        inst = sns.Instance(j["nx"], j["ny"], ly=2.0, stretch=float(j["nx"]), Ra=j["Ra"],
                            Pr=j["Pr"], implicitx=j["implicitx"], implicity=j["implicity"])
        inst.state_alloc()
        inst.state_init_temperature(j["seed"], j["amplitude"])
        insts.append(inst)
    batch = sns.Batch(insts)
    batch.advance(warmup)
    batch.sync()
    m = insts[0]
    src, dst = m.malloc(16), m.malloc(16)
    m.copy(dst, src, 16)
    m.sync()
    t = time.perf_counter()
    batch.advance(steps)
    batch.sync()
    el = time.perf_counter() - t
    m.copy(dst, src, 16)
    m.sync()
    print(json.dumps({"jobs": len(jobs), "classes": batch.classes(), "steps": steps,
                      "host_s": el}), flush=True)
    batch.close()
    for i in insts:
        i.close()
    return el


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("tune", help="measure the treatment cost table on the GPU")
    t.add_argument("--out", default=COSTS_FILE)
    t.add_argument("--sizes", type=int, nargs="+", default=[128, 256, 512, 1024, 2048])
    t.add_argument("--steps", type=int, default=10)
    tr = sub.add_parser("trace", help="batched ensemble steps between two marker copies")
    tr.add_argument("--steps", type=int, default=1)
    tr.add_argument("--warmup", type=int, default=2)
    a = ap.parse_args(argv)
    import simplenavierstokessolver_amd as sns
    if a.cmd == "trace":
        trace_batch(sns, a.steps, a.warmup)
        return
    costs = measure_costs(sns, a.sizes, a.steps)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump({"what": "seconds per step of each diffusive treatment, indexed "
                           "[implicity][implicitx], ny = 2 nx, uniform grid, single instance",
                   "steps": a.steps, "seconds_per_step": costs}, f, indent=1)
    print(json.dumps(costs))


if __name__ == "__main__":
    main()
