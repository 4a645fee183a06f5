"""Turbulent statistics against the reference's published results.

The reference's example case 4 (docs/source/examples/case4/main.rst):
Rayleigh-Benard convection at Ra = 1e8, 512 x 1024 cells, clipped-Chebyshev
stretch 3, ly = 2, Pr = 0.1, 1 or 10; 500 time units to forget the initial
condition, then 500 time units of statistics sampled once per time unit.
Its published data, kept as fixtures in tests/golden/case4/ (copied from
docs/source/examples/case4/images/data/):

* nusselt_mean.dat: Pr, the mean wall-flux Nusselt number of the
  reference's run and the standard deviation of its samples, and the value
  of van der Poel et al. (2013, table 2);
* Pr*_temp.dat: the y- and time-averaged temperature and its rms at the
  cell centres of the half domain next to the x = 0 wall (the statistics of
  post-processing/statistics.py).

A chaotic flow is reproduced statistically, not sample by sample: the mean
Nusselt number is read against the standard error of a 500-unit mean and
the profiles against their own scale.  The run uses the fused path
(snsg_advance, the time step decided on the device) with Crank-Nicolson
diffusion in x and y, and the device monitors and statistics.

    python -m simplenavierstokessolver_amd.validate --pr 1 [--t-spin 500 --t-avg 500]

prints progress lines and one JSON line with the comparison.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests",
                      "golden", "case4")
PR_TAGS = {0.1: "Pr1e-1", 1.0: "Pr1e+0", 10.0: "Pr1e+1"}


def published(pr):
    """(Nu mean, Nu standard deviation, van der Poel 2013) and the profile
    rows (x, mean T, rms T) of Prandtl number pr."""
    nu = np.loadtxt(os.path.join(GOLDEN, "nusselt_mean.dat"))
    row = nu[np.argmin(np.abs(np.log10(nu[:, 0]) - np.log10(pr)))]
    prof = np.loadtxt(os.path.join(GOLDEN, PR_TAGS[pr] + "_temp.dat"))
    return (row[1], row[2], row[3]), prof


def run(sns, pr, t_spin=500.0, t_avg=500.0, chunk=200, seed=4, log=sys.stderr):
    """Advance the case to t_spin + t_avg; from t_spin on, sample the wall
    Nusselt number and the statistics each time t passes a whole time unit
    (the steps between samples are sized from the last dt to span about half
    a unit, at most `chunk`)."""
    itot, jtot = 512, 1024
    inst = sns.Instance(itot, jtot, ly=2.0, stretch=3.0, Ra=1e8, Pr=pr, implicitx=1, implicity=1)
    inst.state_alloc()
    inst.state_init_temperature(seed, 1e-2)
    inst.set_time(0.0)
    for n in inst.STAT_NAMES:
        inst.upload(n, np.zeros(inst.shape(n)))
    t0 = time.perf_counter()
    t, steps, nus, next_sample, last_log = 0.0, 0, [], t_spin, -1
    n = chunk
    while t < t_spin + t_avg:
        inst.advance(n)
        steps += n
        t, dt = inst.get_time()
        n = max(1, min(chunk, int(0.5 / dt)))
        if t >= next_sample:
            nus.append(inst.diagnostics(state=True)["nu_wall"])
            inst.statistics_collect(state=True)
            next_sample = np.floor(t) + 1.0
        if int(t // 25) != last_log:
            last_log = int(t // 25)
            print("case4 Pr %g: t %.1f dt %.3g steps %d samples %d (%.0f s)" % (
                pr, t, dt, steps, len(nus), time.perf_counter() - t0), file=log, flush=True)
    num = len(nus)
    t1 = inst.download("temp1")[1:-1, :] / num
    t2 = inst.download("temp2")[1:-1, :] / num
    xc = inst.g["xc"]
    inst.close()
    mean = t1.mean(axis=0)
    rms = np.sqrt(np.maximum(t2.mean(axis=0) - mean ** 2, 0.0))
    return dict(nu=np.array(nus), xc=xc, mean=mean, rms=rms, steps=steps, t=t,
                seconds=time.perf_counter() - t0)


def compare(pr, r):
    (nu_ref, nu_sd, nu_vdp), prof = published(pr)
    nu = r["nu"]
    # standard error of the mean from the autocorrelation of the samples
    x = nu - nu.mean()
    ac = np.array([np.dot(x[:-k or None], x[k:]) / np.dot(x, x) for k in range(len(x) // 4)])
    cut = np.argmax(ac <= 0) if (ac <= 0).any() else len(ac)
    tau = 1.0 + 2.0 * ac[1:cut].sum()
    sem = nu.std() * np.sqrt(max(tau, 1.0) / len(nu))
    n = len(prof)
    dx = np.abs(r["xc"][:n] - prof[:, 0]).max()
    return {
        "Pr": pr, "samples": int(len(nu)), "steps": r["steps"], "t_end": r["t"],
        "seconds": r["seconds"],
        "nu_mean": nu.mean(), "nu_std": nu.std(), "nu_sem": sem,
        "nu_published": nu_ref, "nu_published_std": nu_sd, "nu_van_der_poel_2013": nu_vdp,
        "nu_rel_diff_published": nu.mean() / nu_ref - 1.0,
        "nu_diff_in_sem": (nu.mean() - nu_ref) / sem,
        "profile_points": n, "profile_x_max_diff": dx,
        "mean_T_max_abs_diff": np.abs(r["mean"][:n] - prof[:, 1]).max(),
        "rms_T_max_abs_diff": np.abs(r["rms"][:n] - prof[:, 2]).max(),
        "rms_T_peak": float(prof[:, 2].max()),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--pr", type=float, default=1.0, choices=sorted(PR_TAGS))
    ap.add_argument("--t-spin", type=float, default=500.0)
    ap.add_argument("--t-avg", type=float, default=500.0)
    a = ap.parse_args(argv)
    import simplenavierstokessolver_amd as sns
    r = run(sns, a.pr, a.t_spin, a.t_avg)
    print(json.dumps(compare(a.pr, r)), flush=True)


if __name__ == "__main__":
    main()
