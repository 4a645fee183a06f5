"""Runs pinned by the reference's examples, independent of the oracle.

Steady laminar convection: the reference's example case 3
(docs/source/examples/case3/data/exec1.sh:13-17; 32x64, stretch 3, Ra=1e4,
Pr = 0.1, 1 and 10, run to t = 500).

In a steady state the time derivatives vanish,
so the discrete energy relations of the spatial discretisation
(docs/source/numerical_method/spatial_discretisation/discrete_governing_equations/nusselt_number_relations.rst)
make the four Nusselt numbers of src/logging.c -- wall heat flux, energy
injection, kinetic and thermal dissipation -- equal up to rounding, whatever
the time integration; a wrong weight in any stencil, operator or monitor
breaks the agreement.  The steady state is also that of the spatial
operator alone, so two diffusive treatments reach the same Nusselt number.
Measured on MI355X: spreads of 2e-16 to 7e-16 of Nu after t = 500 with
Crank-Nicolson in x and y (2.5190265247, 2.6447173882, 2.6044064856 for the
three Prandtl numbers); explicit diffusion in x and y reaches the same values
to 12 digits (about 1e5 steps, too slow for this suite)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NU_NAMES = ("nu_wall", "nu_inje", "nu_eps_k", "nu_eps_h")


def run_to(sns, Pr, imx, imy, t_end=500.0, chunk=500):
    inst = sns.Instance(32, 64, stretch=3.0, Ra=1e4, Pr=Pr, implicitx=imx, implicity=imy)
    inst.state_alloc()
    inst.state_init_temperature(1, 1e-2)
    inst.set_time(0.0)
    nu_prev = None
    t = 0.0
    while t < t_end:
        d = inst.diagnostics(state=True)
        nu_prev = d["nu_wall"]
        inst.advance(chunk)
        t, _ = inst.get_time()
    d = inst.diagnostics(state=True)
    inst.close()
    return d, nu_prev


@pytest.mark.timeout(200)
@pytest.mark.parametrize("Pr", [0.1, 1.0, 10.0])
def test_steady_nusselt_relations(sns, Pr):
    print("case 3 Pr %g" % Pr, flush=True)
    d, nu_prev = run_to(sns, Pr, 1, 1)
    nu = np.array([d[k] for k in NU_NAMES])
    # a convective state (Nu well above conduction's 1), settled
    assert 2.3 < nu[0] < 2.8, nu
    assert abs(nu[0] - nu_prev) <= 1e-12 * nu[0], (nu[0], nu_prev)
    # the four definitions agree to rounding
    assert (nu.max() - nu.min()) <= 1e-13 * nu.mean(), nu
    # divergence-free, no net x-momentum
    assert d["divmax"] <= 1e-13, d["divmax"]
    assert abs(d["mom_x"]) <= 1e-14, d["mom_x"]


@pytest.mark.timeout(200)
def test_steady_state_independent_of_treatment(sns):
    """Crank-Nicolson in x only reaches the Nusselt number of Crank-Nicolson
    in x and y."""
    print("case 3 Pr 1, IMP-x EXP-y against IMP-IMP", flush=True)
    a, _ = run_to(sns, 1.0, 1, 1)
    b, _ = run_to(sns, 1.0, 1, 0)
    nb = np.array([b[k] for k in NU_NAMES])
    assert (nb.max() - nb.min()) <= 1e-13 * nb.mean(), nb
    assert abs(a["nu_wall"] - b["nu_wall"]) <= 1e-10 * a["nu_wall"], (a["nu_wall"], b["nu_wall"])


def inviscid_state(sns):
    """The random flow of the reference's example case 2
    (docs/source/examples/case2/data/exec1.sh: 32x32, ly=1, stretch 3,
    Ra=1e100, Pr=1, 50 time units with buoyancy), as host arrays."""
    a = sns.Instance(32, 32, ly=1.0, stretch=3.0, Ra=1e100, Pr=1.0)
    a.state_alloc()
    a.state_init_temperature(1, 1e-2)
    a.set_time(0.0)
    t = 0.0
    while t < 50.0:
        a.advance(200)
        t, _ = a.get_time()
    fields = {k: a.state_download(k) for k in ["ux", "uy", "p", "psi", "temp"]}
    a.close()
    return fields


def unforced(sns, fields):
    b = sns.Instance(32, 32, ly=1.0, stretch=3.0, Ra=1e100, Pr=1.0, with_thermal_forcing=False)
    b.state_alloc()
    for k, v in fields.items():
        b.state_upload(k, v)
    b.set_time(0.0)
    return b


def energies(inst):
    d = inst.diagnostics(state=True)
    return d["energy_ux"] + d["energy_uy"], d["energy_t"], d["divmax"]


@pytest.mark.timeout(200)
def test_inviscid_quadratic_quantities_decay(sns):
    """Example case 2 (docs/source/examples/case2/main.rst): without buoyancy
    and diffusion the advective terms conserve the kinetic energy K and the
    thermal energy H exactly in space, and the explicit RK3 dissipates both,
    so each decreases monotonically (exec2.sh: 50 more time units without
    buoyancy), and the loss over a fixed time is third order in dt (the slope
    of the reference's docs/.../adv_convergence_consistent.dat)."""
    print("case 2: inviscid decay", flush=True)
    fields = inviscid_state(sns)
    b = unforced(sns, fields)
    K, H = [], []
    t = 0.0
    while t < 50.0:
        k, h, div = energies(b)
        K.append(k)
        H.append(h)
        assert div <= 1e-13, div
        b.advance(20)
        t, _ = b.get_time()
    b.close()
    dK, dH = np.diff(K), np.diff(H)
    assert len(K) > 50
    assert (dK < 0).all(), dK.max()
    assert (dH < 0).all(), dH.max()
    # the loss over 10 time units at fixed dt = 0.02 and 0.01
    loss = []
    for dt, n in ((0.02, 500), (0.01, 1000)):
        b = unforced(sns, fields)
        k0, h0, _ = energies(b)
        b.set_dt(dt)
        for _ in range(n):
            b.integrate()
        k1, h1, _ = energies(b)
        b.close()
        loss.append((k0 - k1, h0 - h1))
    rk, rh = loss[0][0] / loss[1][0], loss[0][1] / loss[1][1]
    print("loss ratios for dt halved: K %.3f H %.3f" % (rk, rh), flush=True)
    # measured on MI355X: 7.999 (K) and 7.994 (H)
    assert 7.5 < rk < 8.5, rk
    assert 7.5 < rh < 8.5, rh
