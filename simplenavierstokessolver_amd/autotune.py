"""Launch-shape autotuner.

Per grid size it times, on the GPU, the launch shapes the kernels take
(``snsg_launch`` in include/sns_hip.h: the first factor of the four-step y
FFT, the chunk length of the partitioned Crank-Nicolson y solve or no
partition at all, the width of its interface solve, and whether its chunk
interiors are solved again or stored) and writes the fastest into a table
file.  ``load_tuning`` registers that table (``snsg_tuning_add``);
instances created afterwards take its shapes, alone, batched and served
alike.

Every candidate passes ``snsg_launch_check``, and every shape keeps the
results within the parity tolerance of the reference (the GPU tests run
the shapes against the oracle), so the tuner only chooses among launch
configurations that are equally correct.

The search is coordinate descent from the built-in rules: one shape at a
time, the others at their best so far; a value replaces the current one
only when it is faster by more than `--margin`.  Each size is timed the way
it runs: the sizes of the mixed ensemble and of the served stream as a
batch of instances (implicit x and y, the treatment the ensemble's rule
picks for most of them), the large single instance (4096x8192) alone with
implicit x and explicit y.

    python -m simplenavierstokessolver_amd.autotune [--sizes 1024x2048 ...] [--out FILE]
"""
import argparse
import json
import os
import sys
import time

# (itot, jtot): the ensemble's and served stream's sizes, and the large
# single instance
SIZES = [(128, 256), (256, 512), (512, 1024), (1024, 2048), (2048, 4096), (4096, 8192)]


def divisors(n):
    return [d for d in range(1, n + 1) if n % d == 0]


def treatment(itot, jtot):
    """(implicitx, implicity) the size is tuned with."""
    return (1, 0) if (itot, jtot) == (4096, 8192) else (1, 1)


def batch_size(itot, jtot):
    """Instances timed together: about 8 Mi cells, 1 to 16 instances."""
    if (itot, jtot) == (4096, 8192):
        return 1
    return max(1, min(16, (1 << 23) // (itot * jtot)))


def candidates(sns, itot, jtot, field):
    """Values of one launch shape that the kernels take for this size."""
    if field == "fft_n1":
        vals = [d for d in divisors(jtot) if 2 <= d <= jtot // 2]
    elif field == "ysolve_chunk":
        vals = [1] + [d for d in divisors(jtot) if 4 <= d <= 32]  # 1: no partition
    elif field == "ysolve_cols":
        vals = [16, 32, 64]
    elif field == "ysolve_recompute":
        vals = [1, 2]
    elif field == "split_corr":
        vals = [1]
    else:
        raise ValueError(field)
    out = []
    for v in vals:
        try:
            sns.launch_check(itot, jtot, **{field: v})
            out.append(v)
        except sns.GpuError:
            pass
    return out


def time_shapes(sns, itot, jtot, shapes, steps, reps):
    """Milliseconds per step (best of `reps` timings of `steps` steps) of a
    batch of instances of this size with these launch shapes, and the shapes
    in effect."""
    imx, imy = treatment(itot, jtot)
    insts = []
    for k in range(batch_size(itot, jtot)):
        i = sns.Instance(itot, jtot, ly=2.0, stretch=float(itot), Ra=1e9, Pr=1.0,
                         implicitx=imx, implicity=imy)
        i.set_launch(**shapes)
        i.state_alloc()
        i.state_init_temperature(100 + k, 1e-3)
        insts.append(i)
    batch = sns.Batch(insts) if len(insts) > 1 else None
    adv = batch.advance if batch else insts[0].advance
    sync = batch.sync if batch else insts[0].sync
    adv(3)
    sync()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        adv(steps)
        sync()
        best = min(best, (time.perf_counter() - t0) / steps)
    eff = insts[0].launch()
    if batch:
        batch.close()
    for i in insts:
        i.close()
    return 1e3 * best, eff


def tune_size(sns, itot, jtot, steps, reps, margin, log=print):
    imx, imy = treatment(itot, jtot)
    fields = ["fft_n1"] + (["ysolve_chunk", "ysolve_cols", "ysolve_recompute"] if imy else [])
    cur = {}
    rule_ms, rule_eff = time_shapes(sns, itot, jtot, cur, steps, reps)
    best_ms = rule_ms
    log("%dx%d rules %s: %.4f ms/step" % (itot, jtot, rule_eff, rule_ms))
    for field in fields:
        for v in candidates(sns, itot, jtot, field):
            trial = dict(cur, **{field: v})
            ms, eff = time_shapes(sns, itot, jtot, trial, steps, reps)
            log("%dx%d %s=%d: %.4f ms/step" % (itot, jtot, field, v, ms))
            if ms < best_ms * (1.0 - margin):
                cur, best_ms = trial, ms
    # the choice against the rules once more, timed side by side
    final_ms, eff = time_shapes(sns, itot, jtot, cur, steps, reps)
    rule_ms, rule_eff = time_shapes(sns, itot, jtot, {}, steps, reps)
    if final_ms >= rule_ms * (1.0 - margin):
        cur, final_ms, eff = {}, rule_ms, rule_eff
    log("%dx%d chosen %s: %.4f ms/step (rules %.4f)" % (itot, jtot, cur, final_ms, rule_ms))
    entry = {"itot": itot, "jtot": jtot, "treatment": "%s-%s" % ("IMP" if imx else "EXP",
                                                                 "IMP" if imy else "EXP"),
             "batch": batch_size(itot, jtot)}
    entry.update({k: int(cur.get(k, 0)) for k in sns.LAUNCH_FIELDS})
    entry["in_effect"] = eff
    entry["ms_per_step"] = final_ms
    entry["rule_ms_per_step"] = rule_ms
    return entry


def main(argv=None):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import simplenavierstokessolver_amd as sns
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--sizes", nargs="*", default=None, help="itot x jtot, e.g. 1024x2048")
    ap.add_argument("--out", default=sns.TUNING_FILE)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--margin", type=float, default=0.02)
    a = ap.parse_args(argv)
    sizes = SIZES if not a.sizes else [tuple(int(x) for x in s.split("x")) for s in a.sizes]
    sns.tuning_clear()  # time every size from the built-in rules
    entries = []
    for itot, jtot in sizes:
        entries.append(tune_size(sns, itot, jtot, a.steps, a.reps, a.margin,
                                 log=lambda m: print(m, flush=True)))
    doc = {"device": "gfx950", "command": "python -m simplenavierstokessolver_amd.autotune "
           + " ".join(argv if argv is not None else sys.argv[1:]),
           "fields": list(sns.LAUNCH_FIELDS), "entries": entries}
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    print("wrote %s (%d sizes)" % (a.out, len(entries)), flush=True)


if __name__ == "__main__":
    main()
