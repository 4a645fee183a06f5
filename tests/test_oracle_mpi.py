"""The y-slab MPI restatement of the oracle (oracle/sns_oracle_mpi, the
timed CPU baseline) against the single-rank oracle (CPU only; skipped where
MPI is not installed).  It splits the rows over the ranks as the reference
does (src/parallel/others.c:6-19), exchanges halo rows and runs every
y-direction operation through global transposes; since each operation is
the single-rank one in the same order and the only reduction is a minimum,
the fields after several steps are equal bit for bit, for even and uneven
slabs, every treatment and both closures of the J = 0 system."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.skipif(not os.path.exists(po.MPI_EXE) or po.mpiexec() is None,
                                reason="MPI oracle not built or mpiexec missing")


def single(itot, jtot, Ra, Pr, imx, imy, seed, amp, n, stretch, pin):
    po.set_variant(po.VARIANT_PIN_GAUGE if pin else 0)
    try:
        o = po.Oracle(itot, jtot, stretch=stretch, Ra=Ra, Pr=Pr)
        o.set_treatment(imx, imy)
        o.init_temperature(seed, amp, linear=False)
        o.apply_boundaries()
        o.run(n)
    finally:
        po.set_variant(0)
    return o


@pytest.mark.parametrize("itot,jtot,imx,imy,ranks,stretch,pin", [
    (24, 20, 1, 1, 3, 3.0, True),
    (24, 20, 0, 0, 2, 3.0, False),
    (32, 48, 1, 0, 4, 32.0, True),
    (17, 23, 0, 1, 5, 3.0, False),     # uneven slabs, odd sizes
    (16, 32, 1, 1, 1, 16.0, True),     # one rank
])
def test_mpi_oracle_equals_single_rank(tmp_path, itot, jtot, imx, imy, ranks, stretch, pin):
    out = str(tmp_path / "fields.bin")
    summary, fields = po.mpi_run(ranks, itot, jtot, 1e6, 1.5, imx, imy, 3, 1e-2, 5,
                                 stretch=stretch, pin_gauge=pin, out=out, timeout=120)
    o = single(itot, jtot, 1e6, 1.5, imx, imy, 3, 1e-2, 5, stretch, pin)
    assert summary["ranks"] == ranks and summary["steps"] == 5
    assert summary["time"] == o.param.time and summary["dt"] == o.param.dt
    for n in ("ux", "uy", "temp", "p"):
        want = o.f[n][1:-1, :]
        assert np.array_equal(fields[n].view(np.uint64), want.view(np.uint64)), n
