"""Batches (snsg_batch_*): every member of a mixed batch ends each advance
call bitwise where the same instance advanced alone with snsg_advance ends
(fields, dt, time), with graph replays and with eager launches, across
several calls with host synchronisation and other work between them."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (itot, jtot, stretch, Ra, Pr, implicitx, implicity, with_temperature, seed)
MEMBERS = [
    (64, 128, 3.0, 1e8, 1.0, 1, 0, True, 1),
    (64, 128, 3.0, 1e7, 2.0, 1, 0, True, 2),     # same class as the first
    (64, 128, 64.0, 1e9, 0.7, 1, 0, True, 3),    # same class, uniform grid
    (32, 64, 3.0, 1e6, 1.0, 0, 0, True, 4),
    (48, 96, 3.0, 1e8, 1.0, 1, 1, True, 5),
    (48, 96, 3.0, 1e7, 4.0, 1, 1, True, 6),
    (40, 80, 3.0, 1e8, 1.0, 0, 1, False, 7),     # no temperature equation
    (100, 60, 3.0, 1e8, 1.0, 1, 0, True, 8),
]
FIELDS = ("ux", "uy", "p", "psi", "temp")


def make(sns, m):
    itot, jtot, stretch, Ra, Pr, imx, imy, wt, seed = m
    inst = sns.Instance(itot, jtot, stretch=stretch, Ra=Ra, Pr=Pr, implicitx=imx, implicity=imy,
                        with_temperature=wt)
    inst.state_alloc()
    rng = np.random.default_rng(seed)
    if wt:
        inst.state_upload("temp", sns.initial_temperature(itot, jtot, inst.g["xc"], seed, 1e-2))
    else:
        # velocity noise; the first projection removes its divergence
        for n in ("ux", "uy"):
            f = 1e-2 * rng.standard_normal(inst.shape(n))
            f[:, 0] = f[:, -1] = 0.   # walls
            inst.state_upload(n, f)
    return inst


def same(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def check(inst, ref, m):
    for n in FIELDS:
        if n == "temp" and not m[7]:
            continue
        got, want = inst.state_download(n), ref.state_download(n)
        assert same(got, want), (m, n, np.abs(got - want).max())
    assert inst.get_time() == ref.get_time(), m


@pytest.mark.parametrize("eager", [False, True])
def test_batch_members_match_alone(sns, eager):
    insts = [make(sns, m) for m in MEMBERS]
    refs = [make(sns, m) for m in MEMBERS]
    batch = sns.Batch(insts, eager=eager)
    assert batch.classes() == 5
    for nsteps in (5, 1, 4):
        print("batch eager=%s advance %d" % (eager, nsteps), flush=True)
        batch.advance(nsteps)
        for r in refs:
            r.advance(nsteps)
        batch.sync()
        for inst, ref, m in zip(insts, refs, MEMBERS):
            check(inst, ref, m)
    batch.close()
    for i in insts + refs:
        i.close()


def test_batch_member_advanced_alone(sns):
    """A member advanced on its own by an odd number of steps leaves its
    class at mixed parity: the batch refuses to advance until it has gone an
    even number; then every member matches its instance advanced alone."""
    ms = MEMBERS[:2]
    insts = [make(sns, m) for m in ms]
    refs = [make(sns, m) for m in ms]
    batch = sns.Batch(insts)
    batch.advance(3)
    for r in refs:
        r.advance(3)
    insts[1].advance(1)
    with pytest.raises(sns.GpuError):
        batch.advance(2)
    insts[1].advance(1)
    refs[1].advance(2)
    batch.advance(2)
    for r in refs:
        r.advance(2)
    batch.sync()
    for inst, ref, m in zip(insts, refs, ms):
        check(inst, ref, m)
    batch.close()
    for i in insts + refs:
        i.close()
