"""Pins of the CPU oracle (CPU only).

Every oracle function is checked against something that does not restate
its formula: textbook / library routines (numpy FFT, dense solves), the
reference's own known-answer test (the transpose permutation) and golden
data file, closed forms of the discrete operators, and invariants of the
energy-consistent scheme.  A plausible mistake (a wrong weight, sign, index
or transposed operand) makes one of these fail.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------------ FFT

@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 16, 30, 64, 97, 128, 192, 256, 1000])
def test_rfft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    y = po.rfft(x)
    ref = np.fft.rfft(x)
    assert np.abs(y - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n", [2, 3, 8, 12, 64, 97, 192, 256])
def test_irfft_is_unnormalised_c2r(n):
    rng = np.random.default_rng(100 + n)
    y = np.fft.rfft(rng.standard_normal(n))
    x = po.irfft(y, n)
    assert np.abs(x - np.fft.irfft(y, n) * n).max() <= 1e-12 * n


def test_rfft_direct_dft_small():
    n = 6
    x = np.arange(n, dtype=float) ** 2
    k = np.arange(n // 2 + 1)[:, None]
    j = np.arange(n)[None, :]
    direct = (x[None, :] * np.exp(-2j * np.pi * j * k / n)).sum(axis=1)
    assert np.abs(po.rfft(x) - direct).max() < 1e-12


def test_rfft_single_mode_is_delta():
    n, k0 = 64, 5
    x = np.cos(2 * np.pi * k0 * np.arange(n) / n)
    y = po.rfft(x)
    expect = np.zeros(n // 2 + 1, complex)
    expect[k0] = n / 2
    assert np.abs(y - expect).max() < 1e-12


# ------------------------------------------------------ tridiagonal solvers

def dense(l, c, u, periodic=False):
    n = len(c)
    A = np.diag(c)
    for i in range(1, n):
        A[i, i - 1] = l[i]
    for i in range(n - 1):
        A[i, i + 1] = u[i]
    if periodic:
        A[0, n - 1] += l[0]
        A[n - 1, 0] += u[n - 1]
    return A


@pytest.mark.parametrize("n", [1, 2, 3, 10, 57])
def test_dgtsv_b_matches_dense(n):
    rng = np.random.default_rng(n)
    l, u = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    c = 2.5 + rng.uniform(0, 1, n)
    q = rng.standard_normal(n)
    x = po.dgtsv_b(l, c, u, q)
    assert np.allclose(dense(l, c, u) @ x, q, rtol=0, atol=1e-13)


@pytest.mark.parametrize("n", [2, 3, 10, 57])
def test_zgtsv_b_matches_dense(n):
    rng = np.random.default_rng(10 + n)
    l, u = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    c = -2.5 - rng.uniform(0, 1, n)
    q = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x = po.zgtsv_b(l, c, u, q)
    assert np.allclose(dense(l, c, u) @ x, q, rtol=0, atol=1e-13)


def test_dgtsv_b_pivot_guard():
    """src/linalg.c:22-25: a pivot below DBL_EPSILON zeroes the row."""
    l, c, u, q = [0.0, 1.0, 0.5], [1.0, 1.0, 2.0], [1.0, 0.3, 0.0], [3.0, 4.0, 5.0]
    x = po.dgtsv_b(l, c, u, q)
    # row 1: val = 1 - 1*1 = 0 -> u1 = q1 = 0; row 2: val = 2 - 0.5*0 = 2, q2 = 5/2
    assert x[2] == 2.5 and x[1] == 0.0 and x[0] == 3.0


@pytest.mark.parametrize("n", [3, 4, 9, 64])
def test_dgtsv_p_matches_dense_cyclic(n):
    rng = np.random.default_rng(20 + n)
    l, u = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    c = 2.5 + rng.uniform(0, 1, n)
    q = rng.standard_normal(n)
    x = po.dgtsv_p(l, c, u, q)
    assert np.allclose(dense(l, c, u, periodic=True) @ x, q, rtol=0, atol=1e-13)


# --------------------------------------------- decomposition and transpose

@pytest.mark.parametrize("num,size", [(7, 3), (192, 4), (5, 5), (1000, 7)])
def test_slab_sizes(num, size):
    sizes = [po.get_size(num, size, r) for r in range(size)]
    offs = [po.get_offset(num, size, r) for r in range(size)]
    assert sum(sizes) == num
    assert max(sizes) - min(sizes) <= 1
    assert offs == list(np.cumsum([0] + sizes[:-1]))


@pytest.mark.parametrize("nranks", [1, 2, 3, 4, 7])
@pytest.mark.parametrize("itot,jtot", [(7, 5), (16, 24), (33, 65)])
@pytest.mark.parametrize("dtype", [np.int32, np.float64, np.complex128])
def test_transpose_kat(nranks, itot, jtot, dtype):
    """The reference's known-answer test (src/parallel/transpose.c:98-154):
    qx[j*itot+i] = (j+joffs)*itot+i on each rank, then qy[i*jtot+j] ==
    j*itot+(i+ioffs)."""
    if nranks > min(itot, jtot):
        pytest.skip("more ranks than rows")
    send, recv = [], []
    for r in range(nranks):
        jsize, joffs = po.get_size(jtot, nranks, r), po.get_offset(jtot, nranks, r)
        isize = po.get_size(itot, nranks, r)
        qx = ((np.arange(jsize)[:, None] + joffs) * itot + np.arange(itot)[None, :]).astype(dtype)
        send.append(np.ascontiguousarray(qx))
        recv.append(np.zeros((isize, jtot), dtype=dtype))
    po.transpose_sim(nranks, itot, jtot, send, recv)
    for r in range(nranks):
        isize, ioffs = po.get_size(itot, nranks, r), po.get_offset(itot, nranks, r)
        expect = (np.arange(jtot)[None, :] * itot + np.arange(isize)[:, None] + ioffs).astype(dtype)
        assert np.array_equal(recv[r], expect)


# ---------------------------------------------------------- grid and RK3

def test_uniform_grid_exact():
    o = po.Oracle(64, 32, stretch=32.0)
    assert np.array_equal(o.xf, np.arange(65) * (1.0 / 64))
    assert np.all(o.dxc[1:-1] == o.dxf[0]) or np.allclose(o.dxc[1:-1], 1 / 64, rtol=1e-15)
    assert o.dxc[0] == o.xc[1] and o.dxc[-1] == 1.0 - o.xc[-2]


@pytest.mark.parametrize("itot,stretch", [(32, 3.0), (33, 2.0), (128, 3.9)])
def test_chebyshev_grid_properties(itot, stretch):
    o = po.Oracle(itot, 8, stretch=stretch)
    xf = o.xf
    assert xf[0] == 0.0 and xf[-1] == 1.0
    assert np.all(np.diff(xf) > 0)
    assert np.allclose(xf + xf[::-1], 1.0, atol=1e-14)          # symmetric about x = 1/2
    assert abs(o.dxf.sum() - 1.0) < 1e-14
    assert np.allclose(o.dxc[0], 0.5 * o.dxf[0], rtol=1e-14)     # half cells at the walls
    # clipped Chebyshev: cells shrink towards the walls, ratio of the wall
    # cell to the central cell follows sin(pi/(2(itot+2s)))-type clustering
    assert o.dxf[0] < o.dxf[itot // 2]
    # stretch is truncated to an integer (src/param/init.c:144)
    o2 = po.Oracle(itot, 8, stretch=float(int(stretch)))
    assert np.array_equal(o.xf, o2.xf)


def test_rk3_is_third_order():
    """Amplification factor of the low-storage scheme on y' = z y must be
    1 + z + z^2/2 + z^3/6 (schemes.rst:96-99 of the reference docs)."""
    o = po.Oracle(8, 8)
    a, b, g = o.rk()
    assert abs(sum(g) - 1.0) < 1e-15
    # polynomial arithmetic on coefficient lists
    def mul_z(p):
        return [0.0] + p
    def add(*ps):
        n = max(len(p) for p in ps)
        return [sum(p[k] if k < len(p) else 0.0 for p in ps) for k in range(n)]
    def scale(p, s):
        return [s * v for v in p]
    f0 = [1.0]
    f1 = add(f0, scale(mul_z(f0), a[0]))
    f2 = add(f1, scale(mul_z(f1), a[1]), scale(mul_z(f0), b[1]))
    f3 = add(f2, scale(mul_z(f2), a[2]), scale(mul_z(f1), b[2]))
    assert np.allclose(f3[:4], [1.0, 1.0, 0.5, 1.0 / 6.0], atol=1e-15)


# ------------------------------------------------------ discrete operators

def divfree_velocity(o, rng, amp=1.0):
    """Velocity from a random streamfunction on cell corners, zero at the
    walls: discretely divergence-free by construction."""
    it, jt = o.itot, o.jtot
    phi = np.zeros((jt + 1, it + 1))  # phi[j-1, i-1] at (xf(i), yf(j))
    phi[:jt, 1:it] = rng.uniform(-amp, amp, (jt, it - 1))
    phi[jt, :] = phi[0, :]
    ux = np.zeros((jt + 2, it + 1))
    uy = np.zeros((jt + 2, it + 2))
    ux[1:jt + 1, :] = (phi[1:, :] - phi[:-1, :]) / o.dy
    uy[1:jt + 1, 1:it + 1] = -(phi[:jt, 1:] - phi[:jt, :-1]) / o.dxf[None, :]
    o.f["ux"][...] = ux
    o.f["uy"][...] = uy


def test_divfree_construction(oracle_mod):
    o = po.Oracle(24, 20, stretch=3.0)
    divfree_velocity(o, np.random.default_rng(0))
    o.apply_boundaries()
    dmax, dsum = o.divergence()
    assert dmax < 1e-11


def test_temperature_diffusion_closed_forms():
    """Uniform grid, u = 0: the x-diffusion of x^2 is exactly 2 kappa and the
    y-diffusion of a Fourier mode is its discrete eigenvalue."""
    it, jt, k = 32, 48, 3
    o = po.Oracle(it, jt, stretch=100.0, Ra=1e6, Pr=2.0)
    kappa = 1.0 / np.sqrt(2.0) / np.sqrt(1e6)
    xc = o.xc[1:-1]
    o.f["temp"][1:-1, 1:-1] = xc[None, :] ** 2
    o.apply_boundaries()
    o.update_temp(0)
    A = o.f["srctempa"]
    assert np.allclose(A[:, 1:-1], 2 * kappa, rtol=1e-10)
    yc = o.yc
    o.f["temp"][1:-1, 1:-1] = np.sin(2 * np.pi * k * yc / o.param.ly)[:, None]
    o.apply_boundaries()
    o.update_temp(0)
    lam = -4.0 / o.dy ** 2 * np.sin(np.pi * k / jt) ** 2
    expect = kappa * lam * np.sin(2 * np.pi * k * yc / o.param.ly)[:, None]
    assert np.allclose(o.f["srctempa"][:, 1:-1], expect * np.ones((1, it - 2)), atol=1e-12)


def test_pressure_gradient_and_buoyancy_closed_forms():
    it, jt = 20, 16
    o = po.Oracle(it, jt, stretch=3.0)
    o.f["p"][:, :] = o.xc[None, :]           # p = x_c
    o.f["temp"][:, :] = np.arange(it + 2)[None, :] * 1.0
    o.apply_boundaries()
    o.f["p"][:, :] = o.xc[None, :]           # keep the linear profile in the ghosts too
    o.compute_force()
    assert np.array_equal(o.f["tempforcex"], np.arange(2, it + 1)[None, :] - 0.5 + np.zeros((jt, 1)))
    o.update_velocity(0)
    assert np.array_equal(o.f["srcuxg"], -np.ones((jt, it - 1)))  # -(dp/dx) = -1 exactly
    assert np.array_equal(o.f["srcuyg"], np.zeros((jt, it)))


def smooth_divfree_velocity(o, rng, amp):
    """Streamfunction made of a few random low modes (zero at the walls)."""
    it, jt = o.itot, o.jtot
    x = o.xf[None, :]
    y = o.yf[:, None]
    phi = np.zeros((jt + 1, it + 1))
    for m in (1, 2, 3):
        for k in (0, 1, 2):
            a, ph = rng.uniform(-1, 1), rng.uniform(0, 2 * np.pi)
            phi += a * np.sin(np.pi * m * x) * np.cos(2 * np.pi * k * y / o.param.ly + ph)
    phi *= amp / np.abs(phi).max()
    phi[:, 0] = 0.0
    phi[:, -1] = 0.0
    phi[jt, :] = phi[0, :]
    o.f["ux"][1:jt + 1, :] = (phi[1:, :] - phi[:-1, :]) / o.dy
    o.f["uy"][1:jt + 1, 1:it + 1] = -(phi[:jt, 1:] - phi[:jt, :-1]) / o.dxf[None, :]


def inviscid(it=24, jt=32, seed=0, smooth=False):
    o = po.Oracle(it, jt, stretch=3.0, Ra=1e100, Pr=1.0, with_thermal_forcing=False)
    rng = np.random.default_rng(seed)
    if smooth:
        smooth_divfree_velocity(o, rng, 0.02)
    else:
        divfree_velocity(o, rng)
    o.f["temp"][1:-1, 1:-1] = rng.uniform(-0.5, 0.5, (jt, it))
    o.apply_boundaries()
    return o


def test_advection_is_energy_neutral():
    """For a discretely divergence-free velocity the advective terms conserve
    the quadratic quantities: sum u.A dV = 0 and sum T.A_T dV = 0
    (momentum_balance_and_quadratic_quantity.rst:169-170 of the reference docs)."""
    o = inviscid()
    dy = o.dy
    ux, uy = o.f["ux"][1:-1, 1:-1].copy(), o.f["uy"][1:-1, 1:-1].copy()
    o.update_velocity(0)
    ex = (ux * o.f["srcuxa"] * o.dxc[None, 1:-1] * dy)
    ey = (uy * o.f["srcuya"] * o.dxf[None, :] * dy)
    total = ex.sum() + ey.sum()
    assert abs(total) < 1e-12 * (np.abs(ex).sum() + np.abs(ey).sum())


def test_advection_energy_neutral_temperature():
    o = inviscid(seed=1)
    T = o.f["temp"][1:-1, 1:-1].copy()
    # walls at +-0.5 carry no advective flux because ux = 0 there
    o.update_temp(0)
    eT = T * o.f["srctempa"] * o.dxf[None, :] * o.dy
    assert abs(eT.sum()) < 1e-12 * np.abs(eT).sum()


def test_poisson_residual_and_divergence_after_correction():
    it, jt = 32, 48
    o = po.Oracle(it, jt, stretch=3.0, Ra=1e8, Pr=1.0)
    rng = np.random.default_rng(3)
    o.f["ux"][...] = rng.uniform(-1, 1, o.f["ux"].shape)
    o.f["uy"][...] = rng.uniform(-1, 1, o.f["uy"].shape)
    o.apply_boundaries()
    o.dt = 1e-2
    gam = o.rk()[2][0]
    o.compute_potential(0)
    psi = o.f["psi"]
    dxf, dxc, dy = o.dxf, o.dxc, o.dy
    lap = ((psi[1:-1, 2:] - psi[1:-1, 1:-1]) / dxc[None, 1:] - (psi[1:-1, 1:-1] - psi[1:-1, :-2]) / dxc[None, :-1]) / dxf[None, :] \
        + (psi[2:, 1:-1] - 2 * psi[1:-1, 1:-1] + psi[:-2, 1:-1]) / dy ** 2
    ux, uy = o.f["ux"], o.f["uy"]
    rhs = ((ux[1:-1, 1:] - ux[1:-1, :-1]) / dxf[None, :] + (uy[2:, 1:-1] - uy[1:-1, 1:-1]) / dy) / (gam * o.dt)
    assert np.abs(lap - rhs).max() < 1e-9 * np.abs(rhs).max()
    o.correct_velocity(0)
    assert o.divergence()[0] < 1e-12 * np.abs(rhs).max() * gam * o.dt


def test_step_mass_and_momentum_balance():
    """After full steps: discrete divergence at round-off and, by continuity
    plus impermeable walls, zero net wall-normal momentum on every face
    (src/logging.c:36-122)."""
    o = po.Oracle(32, 64, stretch=3.0, Ra=1e8, Pr=1.0)
    o.set_treatment(0, 0)
    o.init_temperature(1, 0.05, linear=True)
    o.apply_boundaries()
    o.run(30)
    assert o.divergence()[0] < 1e-12
    face_sums = o.f["ux"][1:-1, :].sum(axis=0)
    assert np.abs(face_sums).max() < 1e-13
    assert abs(o.momentum()[0]) < 1e-15


def test_conduction_state_nusselt_is_one():
    o = po.Oracle(32, 16, stretch=3.0, Ra=1e6, Pr=1.0)
    o.init_temperature(0, 0.0, linear=True)
    o.apply_boundaries()
    nu = o.nusselt()
    assert np.allclose(nu, 1.0, atol=1e-12)


def test_inviscid_energy_error_is_third_order():
    """The reference's inviscid test (inconsistent_results.rst:106-141): with
    the energy-consistent scheme K decays only through the RK3 dissipation,
    so its relative change after a fixed time scales like dt^3; the golden
    file holds the reference's measured errors."""
    ref = np.loadtxt(os.path.join(GOLDEN, "adv_convergence_consistent.dat"))
    ref_order = np.log2(ref[:-1, 1] / ref[1:, 1])
    errs = []
    for dt in [0.04, 0.02, 0.01]:
        o = inviscid(20, 24, seed=5, smooth=True)
        K0 = o.energy()[:2].sum()
        Ks = [K0]
        o.dt = dt
        for _ in range(int(round(2.0 / dt))):
            o.integrate()
            Ks.append(o.energy()[:2].sum())
        assert np.all(np.diff(Ks) <= 1e-15 * K0)   # energy never increases
        errs.append(abs(Ks[-1] - K0) / K0)
    order = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(order > ref_order.min() - 0.6) and np.all(order < ref_order.max() + 0.6), (order, ref_order)


# -------------------------------------------------------------- initial field

def test_initial_temperature_distribution():
    o = po.Oracle(64, 128, stretch=3.0)
    o.init_temperature(7, 1e-2, linear=False)
    T = o.f["temp"][1:-1, 1:-1]
    assert np.abs(T).max() <= 1e-2
    assert abs(T.mean()) < 1e-3 * 1e-2 * 10
    assert abs(T.std() - 1e-2 / np.sqrt(3)) < 1e-4
    o2 = po.Oracle(64, 128, stretch=3.0)
    o2.init_temperature(8, 1e-2, linear=False)
    assert not np.array_equal(o2.f["temp"], o.f["temp"])


def test_splitmix64_known_values():
    # first outputs of the SplitMix64 generator seeded with 0 (Steele, Lea,
    # Flood 2014): state increments by the golden gamma before mixing
    assert po.splitmix64(0) == 0xE220A8397B1DCDAF
    assert po.splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


# --------------------------------------- boundaries, time step, pressure

def test_boundaries_match_padding():
    """Boundary values (src/fluid/boundary_conditions.c, src/temperature/
    boundary_conditions.c) against numpy padding: periodic rows in y, then
    impermeable / no-slip walls (ux, uy), Neumann walls (p) and fixed wall
    temperatures +-0.5 on the interior rows (T)."""
    it, jt = 12, 10
    o = po.Oracle(it, jt, stretch=3.0)
    rng = np.random.default_rng(11)
    for name in ("ux", "uy", "p", "temp"):
        a = rng.standard_normal(o.f[name].shape)
        before = a.copy()
        o.bc(name, a)
        if name == "ux":
            inner = before[1:jt + 1, 1:it]                      # faces i = 2..itot
            e = np.pad(np.pad(inner, ((1, 1), (0, 0)), mode="wrap"), ((0, 0), (1, 1)))
        elif name == "uy":
            inner = before[1:jt + 1, 1:it + 1]
            e = np.pad(np.pad(inner, ((1, 1), (0, 0)), mode="wrap"), ((0, 0), (1, 1)))
        elif name == "p":
            inner = before[1:jt + 1, 1:it + 1]
            e = np.pad(np.pad(inner, ((1, 1), (0, 0)), mode="wrap"), ((0, 0), (1, 1)),
                       mode="edge")
        else:
            inner = before[1:jt + 1, 1:it + 1]
            e = np.pad(inner, ((1, 1), (0, 0)), mode="wrap")
            e = np.concatenate([before[:, :1], e, before[:, -1:]], axis=1)
            e[1:jt + 1, 0] = 0.5
            e[1:jt + 1, -1] = -0.5
        assert np.array_equal(a, e), name


def test_decide_dt_special_cases():
    """Time-step limits (src/param/decide_dt.c:11-101) where the minimum is
    known by hand: at rest the advective bound is 0.7*0.05; one fast interior
    face sets dt_adv = 0.7*dx/|u| while a faster value on the wall face (not
    part of the bound) is ignored; the diffusive bounds are 0.9*0.25*dx^2/
    max(kappa_u, kappa_T); the forced treatment selects which bounds apply."""
    it, jt = 32, 64                      # uniform: dx = 1/32, dy = 2/64
    dx = 1.0 / it
    o = po.Oracle(it, jt, stretch=100.0, Ra=1.0, Pr=4.0)   # kappa_u = 2, kappa_T = 0.5
    o.set_treatment(1, 1)
    assert o.decide_dt() == pytest.approx(0.7 * 0.05, rel=1e-15)
    o.f["ux"][5, 9] = -3.0               # face i = 10, row j = 5
    o.f["ux"][7, 0] = 1e3                # wall face i = 1
    o.f["uy"][3, 4] = 2.0                # dy / 2 = dx / 2 > dx / 3
    dt_adv = 0.7 * dx / 3.0
    assert o.decide_dt() == pytest.approx(dt_adv, rel=1e-15)
    assert (o.param.implicitx, o.param.implicity) == (1, 1)
    dt_dif = 0.9 * 0.25 * dx * dx / 2.0
    for imx, imy in ((0, 0), (1, 0), (0, 1)):
        o.set_treatment(imx, imy)
        assert o.decide_dt() == pytest.approx(dt_dif, rel=1e-15)
        assert (o.param.implicitx, o.param.implicity) == (imx, imy)
    o.f["uy"][3, 4] = -7.0               # now the y-velocity limits
    o.set_treatment(1, 1)
    assert o.decide_dt() == pytest.approx(0.7 * (2.0 / jt) / 7.0, rel=1e-15)


@pytest.mark.parametrize("imx,imy", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_update_pressure_closed_forms(imx, imy):
    """p += psi - gamma dt/(2 Re) Lap(psi) over the implicit directions
    (src/fluid/update_pressure.c:7-60): on a uniform grid the x-Laplacian of
    x^2 is exactly 2 away from the walls, and the y-Laplacian of a Fourier
    mode is its discrete eigenvalue times the mode."""
    it, jt, k = 24, 32, 2
    Ra, Pr, dt, rk = 1e4, 2.0, 0.01, 1
    o = po.Oracle(it, jt, stretch=100.0, Ra=Ra, Pr=Pr)
    o.set_treatment(imx, imy)
    o.dt = dt
    gamma = o.rk()[2][rk]
    Re = np.sqrt(Ra) / np.sqrt(Pr)
    xc = o.xc[None, 1:-1]
    ys = np.sin(2 * np.pi * k * o.yc / o.param.ly)[:, None]
    lam = -4.0 / o.dy ** 2 * np.sin(np.pi * k / jt) ** 2
    p0 = np.random.default_rng(2).standard_normal((jt, it))
    for shape, lap_x, lap_y in ((xc ** 2 + 0 * ys, 2.0, 0.0), (ys + 0 * xc, 0.0, lam * ys)):
        o.f["p"][...] = 0.0
        o.f["p"][1:-1, 1:-1] = p0
        o.f["psi"][...] = 0.0
        o.f["psi"][1:-1, 1:-1] = shape
        o.bc("p", o.f["psi"])
        o.bc("p", o.f["p"])
        o.update_pressure(rk)
        lap = imx * lap_x + imy * lap_y
        expect = p0 + shape - gamma * dt / (2 * Re) * lap
        got = o.f["p"][1:-1, 1:-1]
        cols = slice(1, it - 1)          # Neumann ghosts change the x stencil at the walls
        assert np.allclose(got[:, cols], (expect * np.ones((jt, it)))[:, cols], rtol=0,
                           atol=1e-10), (imx, imy)


# -------------------------------------------- alternative operation orders

@pytest.fixture
def variant():
    def set_(bits):
        po.set_variant(bits)
    yield set_
    po.set_variant(0)


def test_reversed_solve_matches_dense(variant):
    """The reversed elimination order solves the same systems."""
    rng = np.random.default_rng(7)
    n = 40
    l, u = rng.uniform(-1, 0, n), rng.uniform(-1, 0, n)
    c = 2.5 + rng.uniform(0, 1, n)
    q = rng.standard_normal(n)
    A = np.diag(c) + np.diag(l[1:], -1) + np.diag(u[:-1], 1)
    variant(po.VARIANT_REVERSED_SOLVE)
    x = po.dgtsv_b(l, c, u, q)
    assert np.allclose(x, np.linalg.solve(A, q), rtol=0, atol=1e-13)


def potential_state(it=32, jt=24, seed=3):
    o = po.Oracle(it, jt, stretch=3.0)
    smooth_divfree_velocity(o, np.random.default_rng(seed), 0.02)
    o.f["ux"][1:-1, 1:it] += 1e-3 * np.random.default_rng(seed + 1).standard_normal((jt, it - 1))
    o.apply_boundaries()
    o.dt = 1e-2
    return o


def test_direct_dft_variant_same_potential(variant):
    o = potential_state()
    o.compute_potential(0)
    ref = o.f["psi"][1:-1, 1:-1].copy()
    variant(po.VARIANT_DIRECT_DFT)
    o.compute_potential(0)
    got = o.f["psi"][1:-1, 1:-1]
    assert np.abs((got - got.mean()) - (ref - ref.mean())).max() <= 1e-12 * np.abs(ref).max()


def test_pinned_gauge_variant(variant):
    """Pinned J = 0 closure: the y-sum of psi vanishes in the last column and
    psi differs from the reference closure by a constant only (to rounding)."""
    o = potential_state()
    o.compute_potential(0)
    ref = o.f["psi"][1:-1, 1:-1].copy()
    variant(po.VARIANT_PIN_GAUGE)
    o.compute_potential(0)
    got = o.f["psi"][1:-1, 1:-1]
    assert abs(got[:, -1].sum()) <= 1e-13 * np.abs(got).max() * got.shape[0]
    assert np.abs((got - got.mean()) - (ref - ref.mean())).max() <= 1e-12 * np.abs(ref).max()
