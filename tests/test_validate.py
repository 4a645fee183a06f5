"""The published statistics of the reference's example case 4 (fixtures in
tests/golden/case4/) and the host side of their comparison
(simplenavierstokessolver_amd/validate.py); the GPU run itself is a
validation command, not a test (about 1.5e5 steps per Prandtl number)."""
import numpy as np
import pytest

import simplenavierstokessolver_amd as sns
from simplenavierstokessolver_amd import validate


@pytest.mark.parametrize("pr", [0.1, 1.0, 10.0])
def test_grid_matches_published_cell_centres(pr):
    """The published profiles sit at the cell centres of the reference's
    512-cell clipped-Chebyshev grid (stretch 3): the host grid reproduces
    them to the 8 digits printed."""
    (nu, sd, vdp), prof = validate.published(pr)
    assert prof.shape == (258, 3)
    xc = sns.grid(512, 1024, 2.0, 3.0)["xc"]
    assert np.abs(xc[:258] - prof[:, 0]).max() <= 1e-7
    # walls at the fixed temperatures, the rms vanishing there
    assert prof[0, 1] == 0.5 and prof[0, 2] == 0.0
    assert 20.0 < nu < 30.0 and 0.0 < sd < 5.0 and abs(nu / vdp - 1.0) < 0.02


def test_compare_on_a_synthetic_run():
    """The comparison of a run with the published values: a run equal to the
    published profile and mean reads as zero differences."""
    (nu, sd, vdp), prof = validate.published(1.0)
    xc = sns.grid(512, 1024, 2.0, 3.0)["xc"]
    rng = np.random.default_rng(0)
    samples = nu + sd * rng.standard_normal(500)
    samples += nu - samples.mean()
    mean = np.concatenate([prof[:, 1], -prof[::-1, 1][2:]])
    rms = np.concatenate([prof[:, 2], prof[::-1, 2][2:]])
    r = dict(nu=samples, xc=xc, mean=mean, rms=rms, steps=1, t=1000.0, seconds=1.0)
    c = validate.compare(1.0, r)
    assert abs(c["nu_rel_diff_published"]) < 1e-12
    assert c["mean_T_max_abs_diff"] == 0.0 and c["rms_T_max_abs_diff"] == 0.0
    assert c["profile_x_max_diff"] <= 1e-7
    assert 0.0 < c["nu_sem"] < sd
