"""Pins of the oracle's implicit (Crank-Nicolson) branches and of the
momentum diffusion terms (CPU only).

The operators are assembled here from the grid positions (xf, xc, yf, yc)
as products of difference matrices (gradient to faces, then divergence to
cells), independently of the oracle's per-row coefficient formulas.  The
checks are:

* the diffusion terms that a treatment routes into the implicit source
  array G equal the assembled Laplacian applied to the field, times the
  field's diffusivity (kappa_u = sqrt(Pr)/sqrt(Ra) for ux/uy,
  kappa_T = 1/(sqrt(Pr) sqrt(Ra)) for T); on uniform grids they also meet
  the closed forms (x^2 -> 2, a y Fourier mode -> its eigenvalue);
* on stretched grids, the summation-by-parts identity
  sum f L(f) dV = -sum |grad f|^2 dV (plus the wall flux for T);
* the substage increment d solves (I - a Ly)(I - a Lx) d = q over the
  implicit directions, with a = gamma dt kappa / 2 and
  q = dt (alpha A + beta B + gamma G), by a dense solve of the Kronecker
  system on tiny grids (src/temperature/update.c:89-231,
  src/fluid/update_velocity.c:194-488);
* a y Fourier mode at rest is scaled by exactly
  gamma dt kappa lam / (1 - gamma dt kappa lam / 2).

A wrong prefactor (kappa_T against kappa_u, a missing 1/2, sqrt(Pr) on the
wrong side), a DXC/DXF swap or a wrong neighbour index in any operator makes
one of them fail.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

RA, PR = 50.0, 2.0
KAPPA_U = np.sqrt(PR) / np.sqrt(RA)
KAPPA_T = 1.0 / (np.sqrt(PR) * np.sqrt(RA))


# ------------------------------------------------ operators from positions

def grad_div_cells(xc, xf):
    """x Laplacian on cells 1..n with a zero value on each wall (xc(0),
    xc(n+1) are the wall positions): divergence over cell widths of
    gradients over centre distances."""
    n = len(xf) - 1
    G = np.zeros((n + 1, n))                       # faces 1..n+1
    for k in range(n + 1):
        h = xc[k + 1] - xc[k]
        if k < n:
            G[k, k] += 1.0 / h
        if k > 0:
            G[k, k - 1] -= 1.0 / h
    D = np.zeros((n, n + 1))
    for i in range(n):
        w = xf[i + 1] - xf[i]
        D[i, i + 1] += 1.0 / w
        D[i, i] -= 1.0 / w
    return D @ G


def grad_div_faces(xc, xf):
    """x Laplacian on the interior faces 2..n with zero on the wall faces:
    divergence over centre distances of gradients over cell widths."""
    n = len(xf) - 1
    G = np.zeros((n, n - 1))                       # cells 1..n from faces 2..n
    for i in range(n):
        w = xf[i + 1] - xf[i]
        if i < n - 1:
            G[i, i] += 1.0 / w                     # face i+2 (right of cell i+1)
        if i > 0:
            G[i, i - 1] -= 1.0 / w                 # face i+1 (left)
    D = np.zeros((n - 1, n))
    for k in range(n - 1):                         # face k+2 sits between cells k+1, k+2
        h = xc[k + 2] - xc[k + 1]
        D[k, k + 1] += 1.0 / h
        D[k, k] -= 1.0 / h
    return D @ G


def grad_div_periodic(yc, yf, ly):
    n = len(yc)
    G = np.zeros((n, n))                           # face j sits below cell j
    for j in range(n):
        h = yc[j] - (yc[j - 1] if j > 0 else yc[-1] - ly)
        G[j, j] += 1.0 / h
        G[j, j - 1] -= 1.0 / h
    D = np.zeros((n, n))
    for j in range(n):
        w = yf[j + 1] - yf[j]
        D[j, (j + 1) % n] += 1.0 / w
        D[j, j] -= 1.0 / w
    return D @ G


def x_operator(o, field):
    if field == "ux":
        return grad_div_faces(o.xc, o.xf)
    return grad_div_cells(o.xc, o.xf)


def kappa(field):
    return KAPPA_T if field == "temp" else KAPPA_U


SRC = {"temp": "srctemp", "ux": "srcux", "uy": "srcuy"}


def interior(o, field):
    return o.f[field][1:-1, 1:-1]


def fresh(it, jt, stretch, imx, imy):
    o = po.Oracle(it, jt, stretch=stretch, Ra=RA, Pr=PR)
    o.set_treatment(imx, imy)
    return o


def advance(o, field, rk):
    """Run the oracle function that updates `field` at substage rk."""
    if field == "temp":
        o.update_temp(rk)
    else:
        o.compute_force()
        o.update_velocity(rk)


def diffusion_G(o, field, imx, imy, values):
    """Load `values` as the interior of `field` (everything else at rest,
    p = 0) and return the implicit source array G, which then holds only
    the diffusion routed to the implicit directions."""
    o.set_treatment(imx, imy)
    for n in ("ux", "uy", "p", "temp"):
        o.f[n][...] = 0.0
    interior(o, field)[...] = values
    o.apply_boundaries()
    advance(o, field, 0)
    return o.f[SRC[field] + "g"].copy()


# --------------------------------------------------------------- diffusion

@pytest.mark.parametrize("field", ["temp", "ux", "uy"])
@pytest.mark.parametrize("stretch", [100.0, 3.0])
def test_diffusion_matches_assembled_laplacian(field, stretch):
    """G = kappa Lx f (imx) and kappa Ly f (imy), with the operators
    assembled from the positions, on uniform and stretched grids."""
    it, jt = 14, 12
    o = fresh(it, jt, stretch, 1, 0)
    rng = np.random.default_rng(["temp", "ux", "uy"].index(field) + int(stretch))
    f = rng.uniform(-1, 1, interior(o, field).shape)
    Lx = x_operator(o, field)
    Ly = grad_div_periodic(o.yc, o.yf, o.param.ly)
    gx = diffusion_G(o, field, 1, 0, f)
    gy = diffusion_G(o, field, 0, 1, f)
    if field == "temp":
        # the wall values +-0.5 enter the x stencil of the first and last cells
        wall = np.zeros_like(f)
        wall[:, 0] = 0.5 / (o.xc[1] - o.xc[0]) / (o.xf[1] - o.xf[0])
        wall[:, -1] = -0.5 / (o.xc[-1] - o.xc[-2]) / (o.xf[-1] - o.xf[-2])
    else:
        wall = 0.0
    ex = kappa(field) * (f @ Lx.T + wall)
    ey = kappa(field) * (Ly @ f)
    assert np.abs(gx - ex).max() <= 1e-13 * np.abs(ex).max()
    assert np.abs(gy - ey).max() <= 1e-13 * np.abs(ey).max()


@pytest.mark.parametrize("field", ["temp", "ux", "uy"])
def test_diffusion_closed_forms_uniform(field):
    """Uniform grid: c x(1-x) gives -2 c kappa (exact for a quadratic)
    wherever the stencil spans equal spacings, i.e. away from the half cells
    at the walls for the cell fields; a y Fourier mode gives kappa lam."""
    it, jt, k = 16, 24, 3
    o = fresh(it, jt, 100.0, 1, 0)
    x = (o.xf[1:-1] if field == "ux" else o.xc[1:-1])[None, :]
    y = o.yc[:, None]
    c = 0.75
    gx = diffusion_G(o, field, 1, 0, c * x * (1 - x) + 0 * y)
    cols = slice(None) if field == "ux" else slice(1, it - 1)
    assert np.allclose(gx[:, cols], -2 * c * kappa(field), rtol=1e-11, atol=0)
    mode = np.sin(2 * np.pi * k * y / o.param.ly) + 0 * x
    lam = -4.0 / o.dy ** 2 * np.sin(np.pi * k / jt) ** 2
    gy = diffusion_G(o, field, 0, 1, mode)
    assert np.allclose(gy, kappa(field) * lam * mode, rtol=0, atol=1e-12 * abs(lam))


@pytest.mark.parametrize("field", ["temp", "ux", "uy"])
def test_diffusion_summation_by_parts(field):
    """Stretched grid: sum f D(f) dV = -kappa sum |grad f|^2 dV over the
    faces (plus the wall flux for the non-zero wall temperatures), per
    direction, with the volumes of the field's own control cells."""
    it, jt = 18, 10
    o = fresh(it, jt, 3.0, 1, 0)
    rng = np.random.default_rng(5)
    f = rng.uniform(-1, 1, interior(o, field).shape)
    dy = o.dy
    xf, xc = o.xf, o.xc
    gx = diffusion_G(o, field, 1, 0, f)
    gy = diffusion_G(o, field, 0, 1, f)
    if field == "ux":
        vol = (xc[2:-1] - xc[1:-2])[None, :] * dy           # faces 2..itot
        full = np.pad(f, ((0, 0), (1, 1)))                  # wall faces are 0
        grad = np.diff(full, axis=1) / np.diff(xf)[None, :]
        diss_x = (grad ** 2 * np.diff(xf)[None, :] * dy).sum()
        flux = 0.0
    else:
        vol = np.diff(xf)[None, :] * dy
        lo, hi = (0.5, -0.5) if field == "temp" else (0.0, 0.0)
        full = np.concatenate([np.full((jt, 1), lo), f, np.full((jt, 1), hi)], axis=1)
        h = np.diff(xc)[None, :]
        grad = np.diff(full, axis=1) / h                    # faces 1..itot+1
        diss_x = (grad[:, 1:-1] ** 2 * h[:, 1:-1] * dy).sum()
        # boundary term of the discrete integration by parts
        flux = ((f[:, -1] * grad[:, -1]).sum() - (f[:, 0] * grad[:, 0]).sum()) * dy
    lhs_x = (f * gx * vol).sum() / kappa(field)
    assert abs(lhs_x - (flux - diss_x)) <= 1e-12 * (diss_x + abs(flux))
    gyrad = (f - np.roll(f, 1, axis=0)) / dy
    diss_y = (gyrad ** 2 * vol).sum()
    lhs_y = (f * gy * vol).sum() / kappa(field)
    assert abs(lhs_y + diss_y) <= 1e-12 * diss_y


@pytest.mark.parametrize("field", ["temp", "ux", "uy"])
def test_explicit_split_moves_diffusion_into_A(field):
    """The treatment only moves the diffusion between A and G:
    A(exp, exp) - A(imp, imp) = G(imp, imp) - pressure gradient, for a
    general state (advection, buoyancy and pressure present)."""
    it, jt = 12, 10
    rng = np.random.default_rng(9)
    base = {n: rng.uniform(-1, 1, po.Oracle(it, jt).f[n].shape) for n in ("ux", "uy", "p", "temp")}
    res = {}
    for im in (0, 1):
        o = fresh(it, jt, 3.0, im, im)
        for n, a in base.items():
            o.f[n][...] = a
        o.apply_boundaries()
        advance(o, field, 0)
        res[im] = (o.f[SRC[field] + "a"].copy(), o.f[SRC[field] + "g"].copy(), o)
    o = res[1][2]
    if field == "ux":
        pre = -np.diff(o.f["p"][1:-1, 1:-1], axis=1) / np.diff(o.xc[1:-1])[None, :]
    elif field == "uy":
        pre = -(o.f["p"][1:-1, 1:-1] - o.f["p"][:-2, 1:-1]) / o.dy
    else:
        pre = 0.0
    dA = res[0][0] - res[1][0]
    dG = res[1][1] - pre
    scale = np.abs(res[0][0]).max() + np.abs(res[1][1]).max()
    assert np.abs(dA - dG).max() <= 1e-13 * scale
    # explicit treatment leaves only the pressure gradient in G
    if field == "temp":
        assert np.array_equal(res[0][1], np.zeros_like(res[0][1]))
    else:
        assert np.abs(res[0][1] - pre).max() <= 1e-13 * np.abs(pre).max()


# --------------------------------------------------------- implicit solves

def increment_case(field, imx, imy, rk, seed, stretch=3.0, it=9, jt=8):
    """Random state and RK history; returns (oracle, increment d, q)."""
    o = fresh(it, jt, stretch, imx, imy)
    rng = np.random.default_rng(seed)
    for n in ("ux", "uy", "p", "temp"):
        o.f[n][...] = rng.uniform(-1, 1, o.f[n].shape)
    o.apply_boundaries()
    # the previous substage's A becomes B at the start of the call
    o.f[SRC[field] + "a"][...] = rng.uniform(-1, 1, o.f[SRC[field] + "a"].shape)
    o.dt = 0.05
    old = interior(o, field).copy()
    advance(o, field, rk)
    d = interior(o, field) - old
    a, b, g = (o.f[SRC[field] + s] for s in "abg")
    al, be, ga = (v[rk] for v in o.rk())
    q = o.dt * (al * a + be * b + ga * g)
    return o, d, q


def cn_matrix(o, field, imx, imy, rk):
    ga = o.rk()[2][rk]
    a = ga * o.dt * kappa(field) / 2.0
    Lx = x_operator(o, field)
    Ly = grad_div_periodic(o.yc, o.yf, o.param.ly)
    nx, ny = Lx.shape[0], Ly.shape[0]
    Mx = np.eye(nx) - a * Lx if imx else np.eye(nx)
    My = np.eye(ny) - a * Ly if imy else np.eye(ny)
    return np.kron(My, np.eye(nx)) @ np.kron(np.eye(ny), Mx)


@pytest.mark.parametrize("field", ["temp", "ux", "uy"])
@pytest.mark.parametrize("imx,imy", [(1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("rk", [0, 1, 2])
def test_cn_increment_solves_assembled_system(field, imx, imy, rk):
    """Stretched grid: the oracle's increment solves the Crank-Nicolson
    system assembled as a Kronecker product of (I - a L) factors; it also
    equals numpy's dense solve of that system."""
    o, d, q = increment_case(field, imx, imy, rk, seed=100 * rk + 10 * imx + imy)
    M = cn_matrix(o, field, imx, imy, rk)
    dv, qv = d.ravel(), q.ravel()
    res = np.abs(M @ dv - qv).max() / np.abs(qv).max()
    assert res <= 1e-14 * np.abs(M).max(), res
    ref = np.linalg.solve(M, qv)
    assert np.abs(dv - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("field", ["temp", "ux", "uy"])
@pytest.mark.parametrize("rk", [0, 1, 2])
def test_cn_y_mode_closed_form(field, rk):
    """At rest, a y Fourier mode under implicit y diffusion is scaled by
    gamma dt kappa lam / (1 - gamma dt kappa lam / 2) in every column whose
    x stencil does not touch a wall."""
    it, jt, k = 12, 16, 3
    o = fresh(it, jt, 3.0, 0, 1)
    tiny = 1e-30 if field != "temp" else 1.0   # makes advection negligible for u
    y = o.yc[:, None]
    shape = interior(o, field).shape
    mode = tiny * np.sin(2 * np.pi * k * y / o.param.ly) * np.ones((1, shape[1]))
    for n in ("ux", "uy", "p", "temp"):
        o.f[n][...] = 0.0
    interior(o, field)[...] = mode
    o.apply_boundaries()
    o.dt = 0.02
    old = interior(o, field).copy()
    advance(o, field, rk)
    d = interior(o, field) - old
    lam = -4.0 / o.dy ** 2 * np.sin(np.pi * k / jt) ** 2
    z = o.rk()[2][rk] * o.dt * kappa(field) * lam
    expect = z / (1.0 - z / 2.0) * mode
    cols = slice(1, shape[1] - 1)
    assert np.abs(d[:, cols] - expect[:, cols]).max() <= 1e-13 * np.abs(expect).max()


def test_conduction_profile_is_steady():
    """The linear conduction profile (T = 0.5 - x, u = 0) is a discrete
    steady state on any grid: its diffusion vanishes, so no treatment
    changes it."""
    for imx, imy in ((0, 0), (1, 0), (0, 1), (1, 1)):
        o = fresh(16, 8, 3.0, imx, imy)
        o.init_temperature(0, 0.0, linear=True)
        o.apply_boundaries()
        o.dt = 0.05
        old = o.f["temp"].copy()
        o.update_temp(0)
        assert np.abs(o.f["temp"] - old).max() <= 1e-15
