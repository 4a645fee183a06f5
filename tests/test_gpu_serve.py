"""The serving engine (snsg_server_*) and the instance helpers it relies on:
the device initial condition, the memory estimate used for admission, and
jobs whose results equal (bitwise) the same job run alone with snsg_advance,
whether they ran beside others, waited for memory, or reused the instance of
a finished job."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("ux", "uy", "p", "psi", "temp")


def same(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("itot,jtot,seed,amp", [(40, 48, 5, 1e-2), (300, 45, 4001, 1e-3)])
def test_state_init_temperature_matches_host(sns, itot, jtot, seed, amp):
    """Device noise equals the host layer's seeded field; walls +-0.5 and
    periodic halo rows as src/temperature/boundary_conditions.c sets them."""
    inst = sns.Instance(itot, jtot)
    inst.state_alloc()
    inst.state_init_temperature(seed, amp)
    got = inst.state_download("temp")
    ref = sns.initial_temperature(itot, jtot, inst.g["xc"], seed, amp)
    ref[0, 1:-1] = ref[jtot, 1:-1]
    ref[jtot + 1, 1:-1] = ref[1, 1:-1]
    ref[1:jtot + 1, 0] = 0.5
    ref[1:jtot + 1, -1] = -0.5
    assert same(got, ref)
    inst.close()


@pytest.mark.parametrize("itot,jtot", [(64, 128), (100, 60)])
@pytest.mark.parametrize("imx,imy", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_instance_bytes_matches_allocation(sns, itot, jtot, imx, imy):
    inst = sns.Instance(itot, jtot, implicitx=imx, implicity=imy)
    inst.state_alloc()
    inst.advance(1)
    inst.sync()
    assert inst.memory() == sns.instance_bytes(inst.cfg)
    inst.close()


# (itot, jtot, nsteps, Ra, Pr, implicitx, implicity, seed): jobs of one shape
# follow each other so that finished instances are reused
SPECS = [
    (64, 128, 7, 1e8, 1.0, 1, 0, 11),
    (64, 128, 12, 1e7, 2.0, 1, 0, 12),
    (32, 64, 9, 1e6, 1.0, 0, 0, 13),
    (64, 128, 3, 1e9, 0.7, 1, 0, 15),
    (48, 96, 5, 1e8, 1.0, 1, 1, 14),
    (32, 64, 4, 1e7, 3.0, 0, 0, 16),
    (100, 60, 6, 1e8, 1.0, 1, 0, 17),
    (48, 96, 8, 1e6, 4.0, 0, 1, 18),
    (64, 128, 5, 1e8, 1.5, 1, 0, 19),
]


def alone(sns, spec):
    itot, jtot, nsteps, Ra, Pr, imx, imy, seed = spec
    inst = sns.Instance(itot, jtot, stretch=float(itot), Ra=Ra, Pr=Pr, implicitx=imx,
                        implicity=imy)
    inst.state_alloc()
    inst.state_init_temperature(seed, 1e-3)
    inst.advance(nsteps)
    out = {n: inst.state_download(n) for n in FIELDS}
    out["time"] = inst.get_time()
    inst.close()
    return out


def submit(srv, spec, keep):
    itot, jtot, nsteps, Ra, Pr, imx, imy, seed = spec
    return srv.submit(itot, jtot, nsteps, Ra=Ra, Pr=Pr, implicitx=imx, implicity=imy, seed=seed,
                      amplitude=1e-3, keep=keep)


@pytest.mark.parametrize("tight", [False, True])
def test_server_jobs_match_alone(sns, tight):
    """All jobs submitted together (tight: a budget of about two instances,
    so jobs wait for memory and reuse finished instances); each result is
    read, checked against the job alone and released in submission order."""
    sizes = [sns.instance_bytes(sns.make_config(s[0], s[1], implicitx=s[5], implicity=s[6]))
             for s in SPECS]
    budget = int(2.5 * max(sizes)) if tight else 2 * sum(sizes)
    srv = sns.Server(budget)
    ids = []
    for k, spec in enumerate(SPECS):
        ids.append(submit(srv, spec, keep=True))
        if k % 3 == 2:
            time.sleep(0.02)  # later jobs arrive while earlier ones run
    for jid, spec in zip(ids, SPECS):
        print("server tight=%s job %d %s" % (tight, jid, spec), flush=True)
        srv.wait(jid, 120.0)
        state, done = srv.status(jid)
        assert state == 2 and done == spec[2]
        ref = alone(sns, spec)
        for n in FIELDS:
            assert same(srv.download(jid, n), ref[n]), (spec, n)
        assert srv.job_time(jid) == ref["time"], spec
        srv.release(jid)
    st = srv.stats()
    assert st["completed"] == st["submitted"] == len(SPECS)
    assert st["steps"] == sum(s[2] for s in SPECS)
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["mem_peak"] <= budget
    srv.close()


def test_server_stream_drains(sns):
    """Jobs that are not kept: the engine drains them, counts every job-step
    and its cells, and holds no more memory than the budget."""
    specs = SPECS * 2
    sizes = [sns.instance_bytes(sns.make_config(s[0], s[1], implicitx=s[5], implicity=s[6]))
             for s in specs]
    budget = int(3 * max(sizes))
    srv = sns.Server(budget)
    for spec in specs:
        submit(srv, spec, keep=False)
    srv.drain(120.0)
    st = srv.stats()
    assert st["completed"] == len(specs) and st["queued"] == 0 and st["active"] == 0
    assert st["steps"] == sum(s[2] for s in specs)
    assert st["cell_steps"] == sum(s[0] * s[1] * s[2] for s in specs)
    assert st["mem_peak"] <= budget
    srv.close()


def test_server_slots_bound_the_active_jobs(sns):
    """With two slots, at most two jobs are on the GPU at once; the rest
    wait in arrival order and all complete."""
    sizes = [sns.instance_bytes(sns.make_config(s[0], s[1], implicitx=s[5], implicity=s[6]))
             for s in SPECS]
    srv = sns.Server(2 * sum(sizes), max_active=2)
    for spec in SPECS:
        submit(srv, spec, keep=False)
    srv.drain(120.0)
    st = srv.stats()
    assert st["completed"] == len(SPECS) and st["active_peak"] == 2
    assert st["steps"] == sum(s[2] for s in SPECS)
    srv.close()
