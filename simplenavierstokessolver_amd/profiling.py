"""Summaries of rocprofv3 output (kernel statistics, HBM traffic counters).

  python -m simplenavierstokessolver_amd.profiling stats KERNEL_STATS_CSV [--out FILE]
        [--command TEXT]
  python -m simplenavierstokessolver_amd.profiling traffic --fetch CSV [CSV ...]
        --write CSV [CSV ...] --copy-bytes N --cells N [--out FILE] [--json FILE]
        [--command TEXT ...]

  python -m simplenavierstokessolver_amd.profiling counters CSV [CSV ...]
        [--out FILE] [--command TEXT ...]

  python -m simplenavierstokessolver_amd.profiling timeline KERNEL_TRACE_CSV
        [--marker NAME] [--steps N] [--out FILE] [--command TEXT ...]

  python -m simplenavierstokessolver_amd.profiling steps --itot N --jtot N
        [--implicitx 0|1] [--implicity 0|1] [--stretch S] [--ra RA] [--pr PR]
        [--steps N] [--warmup N]

`stats` condenses rocprofv3's kernel_stats.csv into a table of kernel groups
(the names bench.py and the C-ABI timing use).  `traffic` averages the
FETCH_SIZE and WRITE_SIZE counters (KiB per dispatch) of every kernel over
the given counter_collection.csv files and scales each by the factor that
makes the plain copy kernel (k_copy, which reads and writes --copy-bytes
bytes) read exactly its size; the scaled bytes per dispatch are compared
with the kernel's algorithmic bytes (bench.KERNEL_BYTES_PER_CELL x cells).
`counters` tabulates every counter of the given counter_collection.csv
files as its mean per dispatch of each kernel, with the L2 hit rate
TCC_HIT / (TCC_HIT + TCC_MISS) when both are there.
`steps` is a program to profile: it advances one instance of the given
size and treatment (seeded temperature noise 1e-3, the time step decided on
the device) by --warmup and then --steps fused steps, so that a
rocprofv3 run of it sees (warmup + steps) launches of each step kernel.
`timeline` reads a rocprofv3 kernel_trace.csv and takes the window between
the last two dispatches of the marker kernel (a 16-byte k_copy the traced
program issues before and after the steps it wants measured): the share of
the window in which no kernel runs, and per shape class (the dispatches'
grid z, which the batched kernels set to the class's member count, plus
their stream) the kernel time and the time any of its kernels runs.
"""
import argparse
import collections
import csv
import json
import re
import sys

# kernel function -> group
GROUPS = [
    ("k_pred_rows", "pred"), ("k_pred_exp", "pred"),
    ("k_pfwd_a", "poisson_fwd"), ("k_pfwd_b", "poisson_fwd"),
    ("k_poisson_x3", "poisson_x"), ("k_poisson_factor", "setup"),
    ("k_pinv_b", "poisson_inv"), ("k_pinv_a", "poisson_inv"),
    ("k_corr", "corr"), ("k_dt_reduce", "dt_reduce"), ("k_fill_u64", "dt_reduce"),
    ("k_adv_bound", "dt_reduce"),
    ("k_dt", "dt"), ("k_lu_factor", "lu_factor"), ("k_copy", "copy"),
    ("k_transpose", "ysolve"), ("k_cny_rows", "ysolve"), ("k_add_out", "ysolve"),
    ("k_ytrans", "ysolve"), ("k_yfactor", "ysolve"), ("k_ylocal", "ysolve"),
    ("k_yreduced", "ysolve"), ("k_yapply", "ysolve"), ("k_yapply_rows", "ysolve"),
]


def short_name(full):
    m = re.search(r"sns::(\w+)", full)
    return m.group(1) if m else full.split("(")[0]


def group_of(full):
    """Group of a kernel: its name, or a variant of it (name + "_" + suffix,
    e.g. k_pfwd_a_p2, the power-of-two variant of k_pfwd_a)."""
    s = short_name(full)
    for k, g in GROUPS:
        if s == k or s.startswith(k + "_"):
            return g
    return s


def stats(path):
    """[(kernel, group, calls, avg_us, total_ms, pct)] from kernel_stats.csv."""
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            out.append((short_name(r["Name"]), group_of(r["Name"]), int(r["Calls"]),
                        float(r["AverageNs"]) / 1e3, float(r["TotalDurationNs"]) / 1e6,
                        float(r["Percentage"])))
    out.sort(key=lambda t: -t[4])
    return out


def counter_avgs(paths, counter):
    """kernel short name -> (dispatches, mean counter value per dispatch)."""
    acc = collections.defaultdict(list)
    for p in paths:
        with open(p) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                acc[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: (len(v), sum(v) / len(v)) for k, v in acc.items()}


def traffic(fetch_paths, write_paths, copy_bytes, cells, folded=True):
    fe = counter_avgs(fetch_paths, "FETCH_SIZE")
    wr = counter_avgs(write_paths, "WRITE_SIZE")
    if "k_copy" not in fe or "k_copy" not in wr:
        raise SystemExit("traffic: the counter files hold no k_copy dispatch to calibrate with")
    f_fac = copy_bytes / (fe["k_copy"][1] * 1024.0)
    w_fac = copy_bytes / (wr["k_copy"][1] * 1024.0)
    sys.path.insert(0, __file__.rsplit("/", 2)[0])
    # algorithmic bytes per cell and launch (folded: the corrections of
    # substages 1 and 2 applied by the next predictor)
    from bench import KERNEL_BYTES_PER_CELL, kernel_bytes_per_cell
    rows = []
    for k in sorted(set(fe) | set(wr)):
        if k == "k_copy":
            continue
        g = group_of("sns::" + k + "(")
        fb = fe.get(k, (0, 0.0))[1] * 1024.0 * f_fac
        wb = wr.get(k, (0, 0.0))[1] * 1024.0 * w_fac
        alg = kernel_bytes_per_cell(g, folded) if g in KERNEL_BYTES_PER_CELL else None
        rows.append({"kernel": k, "group": g, "dispatches": fe.get(k, (0, 0))[0],
                     "fetch_bytes": fb, "write_bytes": wb, "traffic_bytes": fb + wb,
                     "group_algorithmic_bytes": None if alg is None else alg * cells})
    return {"fetch_factor": f_fac, "write_factor": w_fac, "copy_bytes": copy_bytes,
            "copy_fetch_kib": fe["k_copy"][1], "copy_write_kib": wr["k_copy"][1],
            "cells": cells, "kernels": rows}


def counter_table(paths):
    """(counter names, {kernel: {counter: (dispatches, mean per dispatch)}})."""
    names = []
    with_kernels = collections.defaultdict(dict)
    for p in paths:
        with open(p) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] not in names:
                    names.append(r["Counter_Name"])
    for n in names:
        for k, v in counter_avgs(paths, n).items():
            with_kernels[k][n] = v
    return names, with_kernels


def union_length(intervals):
    """Total length covered by a set of [start, end) intervals."""
    tot, cur_s, cur_e = 0, None, None
    for s, e in sorted(intervals):
        if cur_e is None or s > cur_e:
            if cur_e is not None:
                tot += cur_e - cur_s
            cur_s, cur_e = s, e
        else:
            cur_e = max(cur_e, e)
    if cur_e is not None:
        tot += cur_e - cur_s
    return tot


def timeline(path, marker="k_copy"):
    """Window between the last two marker dispatches: its length, the time
    covered by kernels, and per (stream, grid z) class the summed kernel
    time, the covered time and the dispatch count (times in ns)."""
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append(r)
    marks = [r for r in rows if short_name(r["Kernel_Name"]) == marker]
    if len(marks) < 2:
        raise SystemExit("timeline: fewer than two %s dispatches in %s" % (marker, path))
    marks.sort(key=lambda r: int(r["Start_Timestamp"]))
    t0, t1 = int(marks[-2]["End_Timestamp"]), int(marks[-1]["Start_Timestamp"])
    inside = [r for r in rows if int(r["Start_Timestamp"]) >= t0 and int(r["End_Timestamp"]) <= t1
              and short_name(r["Kernel_Name"]) != marker]
    if not inside:
        raise SystemExit("timeline: no dispatch between the markers")
    stream_key = "Stream_Id" if "Stream_Id" in rows[0] else "Queue_Id"
    classes = collections.defaultdict(list)
    for r in inside:
        z = int(r.get("Grid_Size_Z", 1)) // max(1, int(r.get("Workgroup_Size_Z", 1)))
        classes[(r.get(stream_key, "0"), z)].append(r)
    span = max(int(r["End_Timestamp"]) for r in inside) - min(int(r["Start_Timestamp"])
                                                               for r in inside)
    iv = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in inside]
    out = {"window_ns": t1 - t0, "span_ns": span, "busy_ns": union_length(iv),
           "kernel_ns": sum(e - s for s, e in iv), "dispatches": len(inside), "classes": []}
    for (stream, z), rs in sorted(classes.items(), key=lambda kv: -sum(
            int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in kv[1])):
        civ = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in rs]
        names = collections.Counter(group_of(r["Kernel_Name"]) for r in rs)
        out["classes"].append({"stream": stream, "members": z, "dispatches": len(rs),
                               "kernel_ns": sum(e - s for s, e in civ),
                               "busy_ns": union_length(civ), "groups": dict(names)})
    return out


def run_steps(a):
    import simplenavierstokessolver_amd as sns
    stretch = float(a.itot) if a.stretch is None else a.stretch
    inst = sns.Instance(a.itot, a.jtot, stretch=stretch, Ra=a.ra, Pr=a.pr,
                        implicitx=a.implicitx, implicity=a.implicity)
    inst.state_alloc()
    inst.state_init_temperature(3, 1e-3)
    inst.advance(a.warmup)
    inst.sync()
    inst.advance(a.steps)
    inst.sync()
    print("steps: %dx%d implicitx %d implicity %d, %d + %d steps" % (
        a.itot, a.jtot, a.implicitx, a.implicity, a.warmup, a.steps), flush=True)
    inst.close()


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s1 = sub.add_parser("stats")
    s1.add_argument("csv")
    s1.add_argument("--out")
    s1.add_argument("--command", action="append", default=[])
    s2 = sub.add_parser("traffic")
    s2.add_argument("--fetch", nargs="+", required=True)
    s2.add_argument("--write", nargs="+", required=True)
    s2.add_argument("--copy-bytes", type=int, required=True)
    s2.add_argument("--cells", type=int, required=True)
    s2.add_argument("--split-corr", action="store_true",
                    help="the step ran a correction kernel after every substage")
    s2.add_argument("--out")
    s2.add_argument("--json")
    s2.add_argument("--command", action="append", default=[])
    s3 = sub.add_parser("counters")
    s3.add_argument("csv", nargs="+")
    s3.add_argument("--out")
    s3.add_argument("--command", action="append", default=[])
    s4 = sub.add_parser("timeline")
    s4.add_argument("csv")
    s4.add_argument("--marker", default="k_copy")
    s4.add_argument("--steps", type=int, default=1, help="steps inside the window")
    s4.add_argument("--out")
    s4.add_argument("--json")
    s4.add_argument("--command", action="append", default=[])
    s5 = sub.add_parser("steps")
    s5.add_argument("--itot", type=int, required=True)
    s5.add_argument("--jtot", type=int, required=True)
    s5.add_argument("--implicitx", type=int, default=1)
    s5.add_argument("--implicity", type=int, default=0)
    s5.add_argument("--stretch", type=float, default=None, help="default: uniform (itot)")
    s5.add_argument("--ra", type=float, default=1e11)
    s5.add_argument("--pr", type=float, default=0.7)
    s5.add_argument("--steps", type=int, default=3)
    s5.add_argument("--warmup", type=int, default=1)
    a = ap.parse_args(argv)
    if a.cmd == "steps":
        run_steps(a)
        return
    lines = []
    if a.command:
        lines += ["Commands:", ""] + ["    " + c for c in a.command] + [""]
    if a.cmd == "stats":
        rows = stats(a.csv)
        lines += ["| kernel | group | calls | avg (us) | total (ms) | % |",
                  "|---|---|---:|---:|---:|---:|"]
        for k, g, n, avg, tot, pct in rows:
            lines.append("| %s | %s | %d | %.1f | %.3f | %.2f |" % (k, g, n, avg, tot, pct))
    elif a.cmd == "counters":
        names, tab = counter_table(a.csv)
        hit = "TCC_HIT_sum" in names and "TCC_MISS_sum" in names
        lines += ["Mean per dispatch of each counter.", "",
                  "| kernel | group | dispatches | " + " | ".join(names) +
                  (" | L2 hit % |" if hit else " |"),
                  "|---|---|---:|" + "---:|" * (len(names) + (1 if hit else 0))]
        for k in sorted(tab, key=lambda k: group_of("sns::" + k + "(")):
            row = tab[k]
            nd = max(v[0] for v in row.values())
            vals = ["%.4g" % row[n][1] if n in row else "-" for n in names]
            line = "| %s | %s | %d | %s |" % (k, group_of("sns::" + k + "("), nd, " | ".join(vals))
            if hit:
                h, m = row.get("TCC_HIT_sum", (0, 0.))[1], row.get("TCC_MISS_sum", (0, 0.))[1]
                line += " %s |" % ("%.1f" % (100. * h / (h + m)) if h + m > 0 else "-")
            lines.append(line)
    elif a.cmd == "timeline":
        t = timeline(a.csv, a.marker)
        w = t["window_ns"]
        lines += ["Window between the marker dispatches: %.3f ms for %d step(s), %.3f ms per "
                  "step; %d dispatches." % (w / 1e6, a.steps, w / 1e6 / a.steps, t["dispatches"]),
                  "",
                  "- time with at least one kernel running: %.3f ms (%.1f%% of the window)"
                  % (t["busy_ns"] / 1e6, 100. * t["busy_ns"] / w),
                  "- time with no kernel running (launch gaps, host work): %.3f ms (%.1f%%)"
                  % ((w - t["busy_ns"]) / 1e6, 100. * (w - t["busy_ns"]) / w),
                  "- summed kernel durations: %.3f ms (%.2f kernels running on average "
                  "while busy)" % (t["kernel_ns"] / 1e6, t["kernel_ns"] / max(t["busy_ns"], 1)),
                  "",
                  "| stream | members (grid z) | dispatches | kernel time (ms) | share of kernel "
                  "time | covered (ms) | kernels |", "|---|---:|---:|---:|---:|---:|---|"]
        for c in t["classes"]:
            lines.append("| %s | %d | %d | %.3f | %.1f%% | %.3f | %s |" % (
                c["stream"], c["members"], c["dispatches"], c["kernel_ns"] / 1e6,
                100. * c["kernel_ns"] / t["kernel_ns"], c["busy_ns"] / 1e6,
                ", ".join("%s %d" % kv for kv in sorted(c["groups"].items()))))
        if a.json:
            with open(a.json, "w") as f:
                json.dump(t, f, indent=1)
    else:
        t = traffic(a.fetch, a.write, a.copy_bytes, a.cells, folded=not a.split_corr)
        lines += ["Calibration: k_copy of %d bytes read %.0f KiB (FETCH_SIZE) and wrote %.0f KiB "
                  "(WRITE_SIZE) per dispatch; factors %.4f (fetch), %.4f (write)."
                  % (t["copy_bytes"], t["copy_fetch_kib"], t["copy_write_kib"],
                     t["fetch_factor"], t["write_factor"]), "",
                  "| kernel | group | dispatches | fetch (MB) | write (MB) | traffic (MB) |"
                  " group algorithmic (MB) |", "|---|---|---:|---:|---:|---:|---:|"]
        for r in t["kernels"]:
            alg = r["group_algorithmic_bytes"]
            lines.append("| %s | %s | %d | %.1f | %.1f | %.1f | %s |" % (
                r["kernel"], r["group"], r["dispatches"], r["fetch_bytes"] / 1e6,
                r["write_bytes"] / 1e6, r["traffic_bytes"] / 1e6,
                "-" if alg is None else "%.1f" % (alg / 1e6)))
        if a.json:
            with open(a.json, "w") as f:
                json.dump(t, f, indent=1)
    text = "\n".join(lines) + "\n"
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    else:
        sys.stdout.write(text)


if __name__ == "__main__":
    main()
