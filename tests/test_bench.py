"""The bench line's byte accounting (host logic, no GPU): the per-kernel
algorithmic bytes add up to the per-step figures the step roofline uses,
split and folded, and the mixed-treatment roofline weights them by
cell-steps."""
import pytest

import bench


def step_words(folded):
    """FP64 words per cell and step from the per-kernel figures and each
    kernel's launches per step (explicit y)."""
    k = bench.kernel_bytes_per_cell
    per_step = (3 * k("pred", folded) + 3 * k("poisson_fwd", folded) + 3 * k("poisson_x", folded)
                + 3 * k("poisson_inv", folded) + (1 if folded else 3) * k("corr", folded))
    return per_step / 8


def test_split_step_is_75_words():
    assert step_words(False) == pytest.approx(75.0)
    assert bench.bytes_per_cell_step(1, 0, folded=False) == 600.0
    assert bench.bytes_per_cell_step(0, 0) == 600.0  # explicit x never folds


def test_folded_step_is_65_words():
    # two corrections (7 words each) give way to a psi read and a p write in
    # each of the two folded predictors
    assert step_words(True) == pytest.approx(65.0)
    assert bench.bytes_per_cell_step(1, 0) == 520.0
    assert bench.kernel_bytes_per_cell("pred", True) == pytest.approx(8 * (10 + 15 + 12) / 3)
    assert bench.kernel_bytes_per_cell("pred", False) == 8 * 11


def test_implicit_y_step():
    assert bench.bytes_per_cell_step(1, 1) == 816.0
    assert bench.bytes_per_cell_step(0, 1) == 816.0


def test_mix_roofline_weights_by_cell_steps():
    jobs = [({"implicitx": 1, "implicity": 1}, 3.0), ({"implicitx": 0, "implicity": 0}, 1.0)]
    r = bench.mix_roofline(1e9, jobs)
    b = (816.0 * 3 + 600.0) / 4
    assert r["bytes_per_cell_step"] == pytest.approx(b)
    assert r["achieved"] == pytest.approx(b)  # 1e9 cell-steps/s x b bytes, in GB/s
    assert r["frac"] == pytest.approx(b / bench.HBM_PEAK_GBS)
    assert bench.mix_roofline(1e9, []) is None
