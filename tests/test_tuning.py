"""Launch shapes on the host (no GPU): snsg_launch_check's rules, the
committed launch-shape table, and the autotuner's candidates and search."""
import json
import os

import pytest

import simplenavierstokessolver_amd as sns
from simplenavierstokessolver_amd import autotune


@pytest.mark.parametrize("itot,jtot,shapes", [
    (256, 512, {}),
    (128, 256, dict(fft_n1=2)),              # second pass of 128
    (256, 512, dict(fft_n1=256)),            # second factor 2
    (48, 240, dict(fft_n1=20)),              # 20 x 12: radices 4, 5 and 3
    (4096, 8192, dict(fft_n1=128)),
    (64, 256, dict(ysolve_chunk=4, ysolve_cols=16, ysolve_recompute=2)),
    (16, 16384, dict(ysolve_chunk=32)),      # 512 chunks
    (256, 512, dict(ysolve_chunk=1)),        # no partition
    (40, 8, dict(ysolve_chunk=1)),
])
def test_launch_check_accepts(itot, jtot, shapes):
    sns.launch_check(itot, jtot, **shapes)


@pytest.mark.parametrize("itot,jtot,shapes", [
    (256, 512, dict(fft_n1=7)),              # not a divisor
    (256, 512, dict(fft_n1=2)),              # second pass of 256 exceeds the LDS tile
    (256, 512, dict(fft_n1=512)),            # second factor 1
    (256, 512, dict(fft_n1=1)),
    (4096, 8192, dict(fft_n1=32)),           # second pass of 256 exceeds the LDS tile
    (4096, 8192, dict(fft_n1=512)),          # first pass of 512 exceeds the LDS tile
    (256, 512, dict(ysolve_chunk=33)),
    (256, 512, dict(ysolve_chunk=3)),
    (256, 48, dict(ysolve_chunk=16)),        # 3 chunks
    (16, 16384, dict(ysolve_chunk=16)),      # 1024 chunks
    (256, 512, dict(ysolve_cols=8)),
    (256, 512, dict(ysolve_recompute=3)),
    (256, 512, dict(fft_n1=-2)),
])
def test_launch_check_rejects(itot, jtot, shapes):
    with pytest.raises(sns.GpuError):
        sns.launch_check(itot, jtot, **shapes)


def test_unknown_shape_name():
    with pytest.raises(ValueError):
        sns.launch_check(64, 128, fft_radix=4)


def test_split_corr_values():
    sns.launch_check(64, 128, split_corr=0)
    sns.launch_check(64, 128, split_corr=1)
    for bad in (-1, 2):
        with pytest.raises(sns.GpuError):
            sns.launch_check(64, 128, split_corr=bad)


def test_committed_table():
    """Every entry of the committed table is a valid shape set for its size,
    the sizes are the autotuner's, and no entry is slower than the rules it
    was timed against."""
    assert os.path.exists(sns.TUNING_FILE)
    with open(sns.TUNING_FILE) as f:
        doc = json.load(f)
    assert tuple(doc["fields"]) == sns.LAUNCH_FIELDS
    sizes = set()
    for itot, jtot, shapes in sns.read_tuning():
        sns.launch_check(itot, jtot, **shapes)
        sizes.add((itot, jtot))
    assert sizes == set(autotune.SIZES)
    for e in doc["entries"]:
        sns.launch_check(e["itot"], e["jtot"], **e["in_effect"])
        assert e["ms_per_step"] <= e["rule_ms_per_step"]


@pytest.mark.parametrize("itot,jtot", autotune.SIZES)
def test_candidates_are_valid(itot, jtot):
    for field in sns.LAUNCH_FIELDS:
        vals = autotune.candidates(sns, itot, jtot, field)
        assert vals, field
        for v in vals:
            sns.launch_check(itot, jtot, **{field: v})
    # the built-in first factor is among the candidates
    n1 = max(d for d in autotune.divisors(jtot) if d * d <= jtot)
    assert n1 in autotune.candidates(sns, itot, jtot, "fft_n1")


def test_search_picks_the_faster_shapes(monkeypatch):
    """Coordinate descent on a synthetic cost: a chunk of 16 saves 10%, a
    width of 32 another 5%, an FFT factor of 8 only 1% (inside the margin,
    so the rule stays)."""
    def cost(sns_, itot, jtot, shapes, steps, reps):
        ms = 1.0
        if shapes.get("ysolve_chunk") == 16:
            ms *= 0.9
        if shapes.get("ysolve_cols") == 32:
            ms *= 0.95
        if shapes.get("fft_n1") == 8:
            ms *= 0.99
        return ms, dict(shapes)
    monkeypatch.setattr(autotune, "time_shapes", cost)
    e = autotune.tune_size(sns, 256, 512, 1, 1, 0.02, log=lambda m: None)
    assert (e["fft_n1"], e["ysolve_chunk"], e["ysolve_cols"], e["ysolve_recompute"]) == \
        (0, 16, 32, 0)
    assert e["ms_per_step"] == pytest.approx(0.855)
    assert e["rule_ms_per_step"] == 1.0


def test_search_keeps_the_rules_within_the_margin(monkeypatch):
    monkeypatch.setattr(autotune, "time_shapes",
                        lambda *a: (1.0 - 0.005 * len(a[3]), dict(a[3])))
    e = autotune.tune_size(sns, 128, 256, 1, 1, 0.02, log=lambda m: None)
    assert all(e[k] == 0 for k in sns.LAUNCH_FIELDS)
