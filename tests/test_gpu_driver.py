"""The driver (lib/sns_run: the reference's time loop, src/main.c, through the
C host layer and the C-ABI) end to end on the GPU: its NPY checkpoints
against the oracle's time loop, a restart from a checkpoint, the statistics
output, and the monitor logs written from the device diagnostics."""
import glob
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "simplenavierstokessolver_amd", "lib", "sns_run")

CFG = dict(itot=40, jtot=48, stretch=3.0, ly=2.0, Ra=1e6, Pr=1.0, seed=5, amp=1e-2)
STAT_RATE = 0.1


def run_driver(cwd, timemax, restart=None):
    env = dict(os.environ, itot=str(CFG["itot"]), jtot=str(CFG["jtot"]),
               stretch=repr(CFG["stretch"]), ly=repr(CFG["ly"]), Ra=repr(CFG["Ra"]),
               Pr=repr(CFG["Pr"]), timemax=repr(timemax), wtimemax="100", log_rate="0.1",
               log_after="0", save_rate="1e9", save_after="1e9", implicitx="true",
               implicity="false", ic_seed=str(CFG["seed"]), ic_amplitude=repr(CFG["amp"]),
               ic_linear="false", stat_rate=repr(STAT_RATE), stat_after="0", sns_quiet="1")
    env.pop("dirname_restart", None)
    if restart:
        env["dirname_restart"] = restart
    print("driver: timemax %g restart %s" % (timemax, restart), flush=True)
    r = subprocess.run([EXE], cwd=cwd, env=env, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    saves = sorted(glob.glob(os.path.join(cwd, "output", "save", "step*")))
    assert saves
    return saves[-1]


def load(d, name):
    return np.load(os.path.join(d, name + ".npy"), allow_pickle=False)


class Stats:
    """The statistics schedule of src/main.c:46-49 and the sums of
    src/statistics/collect.c, on the oracle's fields."""

    def __init__(self, o, time):
        self.next = STAT_RATE * np.ceil((time + 1e-8) / STAT_RATE)
        self.num = 0
        self.sums = {k: np.zeros_like(o.f[k]) for k in ("ux", "uy", "temp")}
        self.sq = {k: np.zeros_like(o.f[k]) for k in ("ux", "uy", "temp")}

    def after_step(self, o):
        if self.next < o.param.time:
            for k in self.sums:
                self.sums[k] += o.f[k]
                self.sq[k] += o.f[k] ** 2
            self.num += 1
            self.next += STAT_RATE


def oracle_steps(o, steps, stats):
    for _ in range(steps):
        o.run(1)
        stats.after_step(o)


def new_oracle():
    o = po.Oracle(CFG["itot"], CFG["jtot"], ly=CFG["ly"], stretch=CFG["stretch"], Ra=CFG["Ra"],
                  Pr=CFG["Pr"])
    o.set_treatment(1, 0)
    o.init_temperature(CFG["seed"], CFG["amp"], linear=False)
    o.apply_boundaries()
    return o


def compare_stats(cwd, st):
    d = sorted(glob.glob(os.path.join(cwd, "output", "stat", "step*")))[-1]
    assert int(load(d, "num")) == st.num > 0
    for k, n1, n2 in (("ux", "ux1", "ux2"), ("uy", "uy1", "uy2"), ("temp", "temp1", "temp2")):
        for name, ref in ((n1, st.sums[k]), (n2, st.sq[k])):
            got = load(d, name)
            ref = ref[1:-1, :]
            if k != "ux":
                got, ref = got[:, 1:-1], ref[:, 1:-1]   # the walls hold boundary values
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err <= 1e-12, (name, err)


def compare(d, o):
    step = int(load(d, "step"))
    assert abs(float(load(d, "time")) - o.param.time) <= 1e-13 * o.param.time
    for n in ("ux", "uy", "temp", "p"):
        got, ref = load(d, n), o.f[n][1:-1, :]
        assert got.shape == ref.shape, (n, got.shape, ref.shape)
        if n == "p":
            got, ref = got[:, 1:-1] - got[:, 1:-1].mean(), ref[:, 1:-1] - ref[:, 1:-1].mean()
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err <= 1e-12, (n, step, err)
    return step


def test_driver_checkpoint_and_restart(tmp_path):
    po.set_variant(po.VARIANT_PIN_GAUGE)
    try:
        first = run_driver(str(tmp_path), 0.4)
        n1 = int(load(first, "step"))
        assert n1 > 5
        o = new_oracle()
        st = Stats(o, 0.0)
        oracle_steps(o, n1, st)
        compare(first, o)
        compare_stats(str(tmp_path), st)
        # the run continued from its own checkpoint equals the oracle continued;
        # statistics restart from zero, as in the reference
        restart_dir = str(tmp_path / "restart")
        os.makedirs(restart_dir)
        second = run_driver(restart_dir, 0.8, restart=first)
        n2 = int(load(second, "step"))
        assert n2 > n1
        st2 = Stats(o, o.param.time)
        oracle_steps(o, n2 - n1, st2)
        compare(second, o)
        compare_stats(restart_dir, st2)
    finally:
        po.set_variant(0)
    # monitors from the device: divergence at round-off on every logged line
    lines = open(os.path.join(str(tmp_path), "output", "log", "divergence.dat")).read().split("\n")
    rows = [list(map(float, l.split())) for l in lines if l.strip()]
    assert len(rows) >= 3
    assert max(r[1] for r in rows) < 1e-10
