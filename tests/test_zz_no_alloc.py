"""snsg_diagnostics_noalloc (the monitors without allocation, partial sums
in the library's static device scratch) against snsg_diagnostics, called
through the C-ABI: the same 11 results bit for bit, on two calls in a row on
one GPU, the second on a smaller grid (fewer partial slots of the shared
scratch in use), and again after the first grid's call, so stale partials of
a larger call cannot leak into a smaller one."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def instance(sns, itot, jtot, seed):
    inst = sns.Instance(itot, jtot, stretch=3.0, Ra=1e8, Pr=2.0)
    rng = np.random.default_rng(seed)
    for n in ("ux", "uy", "temp"):
        a = rng.uniform(-1.0, 1.0, inst.shape(n))
        inst.upload(n, a)
        inst.update_boundaries(n)
    return inst


def noalloc(inst):
    out = inst.malloc(11 * 8)
    b = inst.buf
    print("snsg_diagnostics_noalloc %dx%d" % (inst.itot, inst.jtot), flush=True)
    rc = inst.L.snsg_diagnostics_noalloc(inst.ctx, b["ux"], b["uy"], b["temp"], out)
    assert rc == 0, inst.L.snsg_last_error()
    inst.sync()
    res = np.zeros(11)
    assert inst.L.snsg_download(inst.ctx, res.ctypes.data, out, res.nbytes) == 0
    inst.free_raw(out)
    return res


def test_noalloc_diagnostics_match(sns):
    big = instance(sns, 256, 2400, 1)      # 2048 partial slots (the most)
    small = instance(sns, 40, 96, 2)       # 96 partial slots
    want = {k: np.array(list(i.diagnostics().values())) for k, i in
            (("big", big), ("small", small))}
    for name, inst in (("big", big), ("small", small), ("big", big), ("small", small)):
        got = noalloc(inst)
        assert np.array_equal(got.view(np.uint64), want[name].view(np.uint64)), \
            (name, got, want[name])
    for i in (big, small):
        i.close()


def test_noalloc_rejects_null_output(sns):
    inst = instance(sns, 32, 64, 3)
    b = inst.buf
    assert inst.L.snsg_diagnostics_noalloc(inst.ctx, b["ux"], b["uy"], b["temp"], None) != 0
    inst.close()
