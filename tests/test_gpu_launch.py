"""Launch shapes (snsg_launch: the first factor of the four-step y FFT, the
chunk length of the partitioned Crank-Nicolson y solve or none, the width of its
interface solve, recompute or store) against the oracle, and the launch-shape
table (snsg_tuning_add) through the batch and the serving engine.

Every shape the autotuner may choose reorders floating-point operations
only, so each is held to the bar of the default shapes: fields within 1e-12
of the oracle after whole fused steps (p after removing its mean).  The
cases include the extreme FFT factors (a first pass of length 2, a second
of length 2), chunk lengths from 4 to 32 and interface systems beyond 256
rows.  Instances with the same shapes return the same bits alone, batched
and served."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture
def pinned_gauge():
    po.set_variant(po.VARIANT_PIN_GAUGE)
    yield
    po.set_variant(0)


@pytest.fixture
def table(sns):
    sns.tuning_clear()
    yield sns
    sns.tuning_clear()


def interior(a):
    return a[1:-1, 1:-1]


def rel_err(a, ref, demean=False):
    a, ref = interior(a), interior(ref)
    if demean:
        a, ref = a - a.mean(), ref - ref.mean()
    return np.abs(a - ref).max() / max(np.abs(ref).max(), 1e-300)


def same(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def pair(sns, itot, jtot, stretch, Ra, Pr, imx, imy, seed, amp, shapes):
    o = po.Oracle(itot, jtot, stretch=stretch, Ra=Ra, Pr=Pr)
    o.set_treatment(imx, imy)
    o.init_temperature(seed, amp, linear=False)
    o.apply_boundaries()
    inst = sns.Instance(itot, jtot, stretch=stretch, Ra=Ra, Pr=Pr, implicitx=imx, implicity=imy)
    inst.set_launch(**shapes)
    inst.state_alloc()
    for n in ["ux", "uy", "p", "psi", "temp"]:
        inst.state_upload(n, o.f[n])
    return o, inst


def check_fields(o, inst, what):
    for n in ["ux", "uy", "temp", "p"]:
        e = rel_err(inst.state_download(n), o.f[n], demean=n == "p")
        assert e <= 1e-12, (what, n, e)


# (itot, jtot, stretch, Ra, Pr, implicitx, implicity, steps, shapes)
SHAPE_CASES = [
    (64, 256, 3.0, 1e8, 1.0, 1, 1, 10, dict(fft_n1=2)),            # passes of 2 and 128
    (64, 256, 3.0, 1e8, 1.0, 1, 1, 10, dict(fft_n1=128)),          # passes of 128 and 2
    (64, 256, 3.0, 1e8, 1.0, 1, 1, 10, dict(fft_n1=32)),
    (48, 240, 3.0, 1e9, 0.7, 1, 0, 10, dict(fft_n1=20)),           # radices 4, 5 and 3
    (96, 4096, 64.0, 1e9, 1.0, 1, 0, 4, dict(fft_n1=256)),         # runtime schedule of 256
    (64, 256, 3.0, 1e8, 1.0, 1, 1, 10, dict(ysolve_chunk=4)),      # 64 chunks of 4 rows
    (64, 256, 3.0, 1e8, 1.0, 0, 1, 10, dict(ysolve_chunk=16, ysolve_cols=16)),
    (64, 256, 3.0, 1e8, 1.0, 1, 1, 10, dict(ysolve_chunk=8, ysolve_cols=32, ysolve_recompute=1)),
    (20, 960, 3.0, 1e7, 2.0, 0, 1, 6, dict(ysolve_chunk=24, ysolve_recompute=2)),
    (16, 8192, 3.0, 1e8, 1.0, 1, 1, 4, dict(ysolve_chunk=16)),     # 511-row interface block
    (16, 4096, 3.0, 1e8, 1.0, 1, 1, 4, dict(ysolve_chunk=32, ysolve_cols=64, fft_n1=128)),
    (64, 256, 3.0, 1e8, 1.0, 1, 1, 10, dict(ysolve_chunk=1)),      # no partition
    (48, 96, 3.0, 1e6, 2.0, 0, 1, 10, dict(ysolve_chunk=1, fft_n1=8)),
]


@pytest.mark.parametrize("case", SHAPE_CASES)
def test_launch_shapes_match_oracle(sns, pinned_gauge, case):
    itot, jtot, stretch, Ra, Pr, imx, imy, steps, shapes = case
    print("launch shapes %dx%d %s" % (itot, jtot, shapes), flush=True)
    o, inst = pair(sns, itot, jtot, stretch, Ra, Pr, imx, imy, 3, 1e-2, shapes)
    eff = inst.launch()
    for k, v in shapes.items():
        assert eff[k] == v, (k, eff)
    o.run(steps)
    inst.advance(steps)
    check_fields(o, inst, shapes)
    t_gpu, _ = inst.get_time()
    assert abs(t_gpu - o.param.time) <= 1e-12 * o.param.time
    inst.close()


@pytest.mark.parametrize("shapes", [dict(ysolve_chunk=1), dict(ysolve_chunk=4),
                                    dict(ysolve_chunk=16, ysolve_cols=16, ysolve_recompute=2),
                                    dict(ysolve_chunk=8, ysolve_recompute=1)])
@pytest.mark.parametrize("imx", [0, 1])
def test_reference_functions_follow_launch_shapes(sns, shapes, imx):
    """The per-function entry points solve Crank-Nicolson in y with the
    instance's launch shapes (chunk length or none, interface width,
    recompute or store), as the fused step does: over three steps the
    reference-order path and the fused step agree bit for bit."""
    itot, jtot = 64, 256
    print("reference functions %s imx %d" % (shapes, imx), flush=True)
    o = po.Oracle(itot, jtot, stretch=3.0, Ra=1e9, Pr=0.7)
    o.set_treatment(imx, 1)
    o.init_temperature(13, 1e-2, linear=False)
    o.apply_boundaries()
    inst = sns.Instance(itot, jtot, stretch=3.0, Ra=1e9, Pr=0.7, implicitx=imx, implicity=1)
    inst.set_launch(**shapes)
    inst.alloc()
    inst.state_alloc()
    for n in ["ux", "uy", "p", "psi", "temp", "tempforcex", "srcuxa", "srcuxb", "srcuxg",
              "srcuya", "srcuyb", "srcuyg", "srctempa", "srctempb", "srctempg"]:
        inst.upload(n, o.f[n])
    for n in ["ux", "uy", "p", "psi", "temp"]:
        inst.state_upload(n, o.f[n])
    for _ in range(3):
        inst.set_dt(1e-3)
        inst.integrate_reference_order()
        inst.integrate()
    for n in ["ux", "uy", "p", "temp"]:
        assert same(inst.state_download(n), inst.download(n)), (shapes, n)
    inst.close()


def test_fft_factor_reference_functions(sns):
    """The per-function Poisson solve (snsg_fluid_compute_potential) with a
    non-default FFT factor: psi within 1e-11 of the oracle's (after removing
    the mean; the bound of the default factor's test) from the same seeded
    velocities."""
    itot, jtot = 48, 240
    for n1 in (4, 12, 60):
        o = po.Oracle(itot, jtot, stretch=3.0, Ra=1e8, Pr=2.0)
        o.dt = 1e-3
        rng = np.random.default_rng(n1)
        for n in ("ux", "uy"):
            o.f[n][...] = rng.uniform(-1.0, 1.0, o.f[n].shape)
        o.apply_boundaries()
        inst = sns.Instance(itot, jtot, stretch=3.0, Ra=1e8, Pr=2.0)
        inst.set_launch(fft_n1=n1)
        assert inst.launch()["fft_n1"] == n1
        inst.alloc()
        for n in ("ux", "uy"):
            inst.upload(n, o.f[n])
        inst.set_dt(1e-3)
        inst.compute_potential(0)
        o.compute_potential(0)
        e = rel_err(inst.download("psi"), o.f["psi"], demean=True)
        assert e <= 1e-11, (n1, e)
        inst.close()


def test_chunk_change_between_steps(sns, pinned_gauge):
    """A new chunk length after the first steps: the y solve's scratch is
    sized again and the run stays within 1e-12 of the oracle."""
    o, inst = pair(sns, 64, 256, 3.0, 1e8, 1.0, 1, 1, 5, 1e-2, {})
    o.run(3)
    inst.advance(3)
    inst.set_launch(ysolve_chunk=8, ysolve_cols=16)
    assert inst.launch()["ysolve_chunk"] == 8
    o.run(3)
    inst.advance(3)
    check_fields(o, inst, "chunk 8 after 3 steps")
    inst.close()


def test_set_launch_rejects_bad_shapes(sns):
    inst = sns.Instance(64, 256)
    for bad in (dict(fft_n1=3), dict(fft_n1=256), dict(ysolve_chunk=33), dict(ysolve_chunk=6),
                dict(ysolve_cols=8), dict(ysolve_recompute=3), dict(split_corr=2),
                dict(split_corr=-1)):
        with pytest.raises(sns.GpuError):
            inst.set_launch(**bad)
    inst.close()


# (itot, jtot, Ra, Pr, implicitx, implicity, seed)
MEMBERS = [
    (64, 256, 1e8, 1.0, 1, 1, 1),
    (64, 256, 1e7, 2.0, 1, 1, 2),
    (64, 256, 1e9, 0.7, 1, 1, 3),
    (48, 96, 1e8, 1.0, 1, 0, 4),
    (48, 96, 1e6, 3.0, 1, 0, 5),
]
TABLE = {(64, 256): dict(fft_n1=8, ysolve_chunk=8, ysolve_cols=16, ysolve_recompute=2),
         (48, 96): dict(fft_n1=4)}


def make(sns, m):
    itot, jtot, Ra, Pr, imx, imy, seed = m
    inst = sns.Instance(itot, jtot, stretch=float(itot), Ra=Ra, Pr=Pr, implicitx=imx,
                        implicity=imy)
    inst.state_alloc()
    inst.state_init_temperature(seed, 1e-3)
    return inst


def test_table_batch_matches_alone(table):
    """Instances created after snsg_tuning_add take the table's shapes; a
    batch of them ends bitwise where each ends alone, and one instance with
    other shapes (set on it) forms a class of its own."""
    sns = table
    for (itot, jtot), shapes in TABLE.items():
        sns.tuning_add(itot, jtot, **shapes)
    alone = []
    for m in MEMBERS:
        inst = make(sns, m)
        eff = inst.launch()
        for k, v in TABLE[(m[0], m[1])].items():
            assert eff[k] == v, (m, k, eff)
        inst.advance(6)
        alone.append({n: inst.state_download(n) for n in ("ux", "uy", "p", "temp")})
        inst.close()
    insts = [make(sns, m) for m in MEMBERS]
    other = make(sns, MEMBERS[0])
    other.set_launch()  # the rules
    b = sns.Batch(insts + [other])
    assert b.classes() == 3
    with pytest.raises(sns.GpuError):  # shapes are fixed while in a batch
        other.set_launch(fft_n1=8)
    with pytest.raises(sns.GpuError):  # so is the treatment
        other.set_treatment(0, 1)
    with pytest.raises(sns.GpuError):  # the batch's graphs hold its fields
        other.close()
    assert other.ctx and other.state is not None
    b.advance(6)
    b.sync()
    for m, inst, ref in zip(MEMBERS, insts, alone):
        for n in ref:
            assert same(inst.state_download(n), ref[n]), (m, n)
    b.close()
    other.set_launch(fft_n1=8)  # free again once the batch is gone
    for i in insts + [other]:
        i.close()


def test_table_served_matches_alone(table):
    """Served jobs take the table's shapes as their instances are created;
    each result equals the same job alone with the same table."""
    sns = table
    for (itot, jtot), shapes in TABLE.items():
        sns.tuning_add(itot, jtot, **shapes)
    specs = [(m, 5 + k) for k, m in enumerate(MEMBERS)]
    sizes = [sns.instance_bytes(sns.make_config(m[0], m[1], implicitx=m[4], implicity=m[5]))
             for m, _ in specs]
    srv = sns.Server(2 * sum(sizes))
    ids = [srv.submit(m[0], m[1], n, Ra=m[2], Pr=m[3], implicitx=m[4], implicity=m[5],
                      seed=m[6], amplitude=1e-3, keep=True) for m, n in specs]
    for jid, (m, n) in zip(ids, specs):
        srv.wait(jid, 120.0)
        inst = make(sns, m)
        inst.advance(n)
        for f in ("ux", "uy", "p", "temp"):
            assert same(srv.download(jid, f), inst.state_download(f)), (m, f)
        inst.close()
        srv.release(jid)
    srv.close()


# (itot, jtot, stretch, with_temperature, steps): implicit x, explicit y
FOLD_CASES = [
    (64, 256, 3.0, True, 6),       # stretched grid, 64-thread rows
    (128, 64, 1e6, True, 6),       # uniform grid, power-of-two divisors
    (96, 90, 1e6, True, 6),        # uniform, divisors not powers of two
    (40, 45, 2.0, True, 6),        # odd jtot: the last row pair is shifted
    (48, 4, 3.0, True, 6),         # jtot = 4: psi row ja-2 wraps to jtot-1
    (64, 96, 3.0, False, 6),       # no temperature equation (two fields)
    (1500, 16, 3.0, True, 3),      # 256-thread rows
    (4500, 8, 1e6, True, 2),       # 1024-thread rows
]


@pytest.mark.parametrize("case", FOLD_CASES, ids=lambda c: "x".join(map(str, c[:2])))
def test_folded_correction_bitwise(sns, case):
    """With implicit x and explicit y the predictors of substages 2 and 3
    apply the previous substage's correction as they load the fields; every
    field after whole steps equals, bit for bit, the sequence with the
    correction kernel after each substage (split_corr)."""
    itot, jtot, stretch, with_t, steps = case
    out = []
    for split in (0, 1):
        inst = sns.Instance(itot, jtot, stretch=stretch, Ra=1e8, Pr=0.7, implicitx=1,
                            implicity=0, with_temperature=with_t)
        inst.set_launch(split_corr=split)
        assert inst.launch()["split_corr"] == split
        inst.state_alloc()
        r = np.random.default_rng(itot * 7919 + jtot)
        for n in ("ux", "uy", "p", "temp"):
            shape = inst.shape(sns._state_shape(n))
            inst.state_upload(n, 1e-3 * r.standard_normal(shape))
        inst.advance(steps)
        out.append({n: inst.state_download(n) for n in
                    ("ux", "uy", "p", "psi", "temp", "ux_star", "uy_star")})
        out[-1]["time"] = np.array(inst.get_time())
        inst.close()
    for n in out[0]:
        assert same(out[0][n], out[1][n]), (case, n)


@pytest.mark.parametrize("case", [(64, 256, 1, 1), (48, 96, 0, 1)],
                         ids=lambda c: "%dx%d_%d%d" % c)
def test_ysolve_fields_without_temperature(sns, pinned_gauge, case):
    """Without the temperature equation the y solves of a substage cover ux
    and uy only (two fields on the launches' grid z): whole steps from a
    noisy velocity field stay within 1e-12 of the oracle."""
    itot, jtot, imx, imy = case
    o = po.Oracle(itot, jtot, stretch=3.0, Ra=1e8, Pr=1.0, with_temperature=False)
    o.set_treatment(imx, imy)
    rng = np.random.default_rng(itot + jtot)
    for n in ("ux", "uy"):
        o.f[n][...] = 1e-2 * rng.standard_normal(o.f[n].shape)
    o.apply_boundaries()
    inst = sns.Instance(itot, jtot, stretch=3.0, Ra=1e8, Pr=1.0, implicitx=imx, implicity=imy,
                        with_temperature=False)
    inst.state_alloc()
    for n in ("ux", "uy", "p", "psi", "temp"):
        inst.state_upload(n, o.f[n])
    o.run(5)
    inst.advance(5)
    for n in ("ux", "uy", "p"):
        e = rel_err(inst.state_download(n), o.f[n], demean=n == "p")
        assert e <= 1e-12, (case, n, e)
    inst.close()
