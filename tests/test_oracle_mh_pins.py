"""Pins of the oracle's MUSCL-Hancock subcell reconstruction (no GPU).

The reference defines `calcflux_mh!` for curvilinear meshes only
(src/solvers/dgsem_structured/dg_2d.jl:313-538); the oracle restates it for the
TreeMesh with the Cartesian metric.  The checks below come from what the
method's own rules fix in closed form, on states where they can be evaluated
by hand:
- with Δt = 0 and element-wise constant data every slope is zero and the
  scheme is the first-order one, bit for bit;
- with Δt = 0, linear data is reconstructed exactly at the interior subcell
  faces, the end subcells (ξ = ±1 nodes, where the one-sided differences are
  0/0) keep their node value, and the residual is the flux difference of those
  face states;
- the half-step on constant data cancels in the interior subcells and leaves
  the one-sided term -Δt/2 J⁻¹ f(u) / w_2 in the first subcell (0.5 Δt in MDRK
  stage 1);
- the positivity scaling of a slope whose doubled extrapolation drops the
  density below 1e-10 brings that density to 0.1 of the cell value;
- conservation, elements on a domain boundary falling back to first order, and
  a shock problem that stays admissible.
"""
import numpy as np
import pytest

from conftest import rel_linf

GAMMA = 1.4


def euler_flux(u, orientation, g=GAMMA):
    rho, m1, m2, E = u[..., 0], u[..., 1], u[..., 2], u[..., 3]
    v1, v2 = m1 / rho, m2 / rho
    p = (g - 1) * (E - 0.5 * (m1 * v1 + m2 * v2))
    if orientation == 1:
        return np.stack([m1, m1 * v1 + p, m1 * v2, (E + p) * v1], -1)
    return np.stack([m2, m2 * v1, m2 * v2 + p, (E + p) * v2], -1)


def rusanov(ul, ur, orientation, g=GAMMA):
    def parts(u):
        rho = u[..., 0]
        v = u[..., orientation] / rho
        p = (g - 1) * (u[..., 3] - 0.5 * (u[..., 1] ** 2 + u[..., 2] ** 2) / rho)
        return np.abs(v), np.sqrt(g * p / rho)

    al, cl = parts(ul)
    ar, cr = parts(ur)
    lam = np.maximum(al, ar) + np.maximum(cl, cr)
    return 0.5 * (euler_flux(ul, orientation, g) + euler_flux(ur, orientation, g)) - 0.5 * lam[..., None] * (ur - ul)


def euler_state(rho, v1, v2, p, g=GAMMA):
    return np.stack([rho, rho * v1, rho * v2, p / (g - 1) + 0.5 * rho * (v1 ** 2 + v2 ** 2)], -1)


def mh_oracle(oracle_mod, N, nx, ny, L, periodic=True, rec="muscl_hancock", **kw):
    return oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, L / nx, periodic=(periodic, periodic), shock_capturing=True,
                             reconstruction=rec, **kw)


def fn_low(o):
    return o.aux("fn_low").reshape(o.E, 4, o.n, 4)  # [e, face, node, v]


def random_element_constants(rng, E):
    rho = 1.0 + 0.3 * rng.random(E)
    v1, v2 = 0.5 * rng.standard_normal(E), 0.5 * rng.standard_normal(E)
    p = 1.0 + 0.3 * rng.random(E)
    return euler_state(rho, v1, v2, p)


@pytest.mark.parametrize("N", [3, 4])
def test_dt0_elementwise_constant_is_first_order(oracle_mod, N):
    nx, ny, L = 5, 4, 2.0
    mh = mh_oracle(oracle_mod, N, nx, ny, L)
    fo = mh_oracle(oracle_mod, N, nx, ny, L, rec="first_order")
    n = N + 1
    rng = np.random.default_rng(N)
    c = random_element_constants(rng, nx * ny)
    u = np.broadcast_to(c[:, None, None, :], (nx * ny, n, n, 4)).copy()
    alpha = 0.2 + 0.8 * rng.random(nx * ny)
    for o in (mh, fo):
        o.set_alpha(alpha)
    assert np.array_equal(mh.rhs(u, 0.0, 0.0), fo.rhs(u, 0.0, 0.0))
    assert np.array_equal(fn_low(mh), fn_low(fo))


@pytest.mark.parametrize("N", [3, 4])
def test_dt0_linear_data_reconstructed_exactly(oracle_mod, N):
    """α = 1 (pure subcell FV, β = 1), Δt = 0, each element linear in x and
    constant in y.  Interior subcell faces get the exact values of the linear
    function on both sides; the end subcells keep their node value."""
    nx, L = 6, 3.0
    h = L / nx
    o = mh_oracle(oracle_mod, N, nx, 1, L)
    n = N + 1
    nodes, w, _, _ = o.basis()
    xs = np.concatenate([[-1.0], -1.0 + np.cumsum(w)])
    rng = np.random.default_rng(10 + N)
    c = random_element_constants(rng, nx)
    s = 0.02 * rng.standard_normal((nx, 4)) * np.array([1.0, 1.0, 1.0, 2.0])
    lin = lambda e, xi: c[e] + s[e] * xi  # noqa: E731
    u = np.zeros((nx, n, n, 4))
    for e in range(nx):
        for i in range(n):
            u[e, :, i] = lin(e, nodes[i])
    o.set_alpha(np.ones(nx))
    du = o.rhs(u, 0.0, 0.0)
    ref = np.zeros_like(u)
    for e in range(nx):
        el, er = (e - 1) % nx, (e + 1) % nx
        F = np.zeros((n + 1, 4))
        F[0] = rusanov(lin(el, nodes[-1]), lin(e, nodes[0]), 1)
        F[n] = rusanov(lin(e, nodes[-1]), lin(er, nodes[0]), 1)
        for k in range(1, n):
            left = lin(e, nodes[0]) if k - 1 == 0 else lin(e, xs[k])
            right = lin(e, nodes[-1]) if k == n - 1 else lin(e, xs[k])
            F[k] = rusanov(left, right, 1)
        for i in range(n):
            ref[e, :, i] = -(2.0 / h) * (F[i + 1] - F[i]) / w[i]
    assert rel_linf(du, ref) < 1e-13
    # first-order FV differs: the reconstruction is active
    fo = mh_oracle(oracle_mod, N, nx, 1, L, rec="first_order")
    fo.set_alpha(np.ones(nx))
    assert rel_linf(fo.rhs(u, 0.0, 0.0), ref) > 1e-4


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("kind,scale", [("lw", 1.0), ("mdrk1", 0.5), ("mdrk2", 1.0)])
def test_half_step_one_sided_at_first_subcell(oracle_mod, N, kind, scale):
    """Element-wise constant data, α = 1: interior subcells' half-step terms
    cancel, the first subcell keeps -Δt/2 J⁻¹ f(u)/w_2 (in x and, on the
    bottom row, in y) and the last gets none.  Seen through fn_low: the flux at
    the first interior subcell face of a middle row is Rusanov(u - c f_x(u), u)
    with c = s Δt/2 · (2/h) / w_2, the one at the last is f_x(u)."""
    nx, ny, L = 4, 3, 2.0
    h = L / nx
    o = mh_oracle(oracle_mod, N, nx, ny, L)
    n = N + 1
    _, w, _, _ = o.basis()
    rng = np.random.default_rng(20 + N)
    c = random_element_constants(rng, nx * ny)
    u = np.broadcast_to(c[:, None, None, :], (nx * ny, n, n, 4)).copy()
    o.set_alpha(np.ones(nx * ny))
    dt = 0.5 * o.max_dt(u)
    if kind == "mdrk2":
        o.rhs(u, 0.0, dt, kind="mdrk1")
        o.set_us(u)
    o.rhs(u, 0.0, dt, kind=kind)
    fl = fn_low(o)
    cc = scale * 0.5 * dt * (2.0 / h) / w[1]
    for e in range(nx * ny):
        fx = euler_flux(c[e], 1)
        for j in range(1, n - 1):
            np.testing.assert_allclose(fl[e, 0, j], rusanov(c[e] - cc * fx, c[e], 1), rtol=1e-13, atol=1e-14)
            np.testing.assert_allclose(fl[e, 1, j], fx, rtol=1e-13, atol=1e-14)
    # a two-sided (textbook) half-step would leave the first subcell at u
    assert np.max(np.abs(fl[:, 0, 1:n - 1] - euler_flux(c, 1)[:, None, :])) > 1e-4


@pytest.mark.parametrize("N", [3, 4])
def test_positivity_scaling_of_steep_density_slopes(oracle_mod, N):
    """Linear data whose doubled extrapolation 2·s·(face - node) drops the
    density below 1e-10 at an interior subcell: the slope is scaled by
    θ = 0.9 ρ_i / max|2 s x| (so ρ* = 0.1 ρ_i at the nearer-to-vacuum face),
    velocity and pressure are uniform, so the pressure pass leaves it."""
    nx, L = 4, 2.0
    h = L / nx
    o = mh_oracle(oracle_mod, N, nx, 1, L)
    n = N + 1
    nodes, w, _, _ = o.basis()
    xs = np.concatenate([[-1.0], -1.0 + np.cumsum(w)])
    v1, v2, p = 0.3, -0.1, 1.0
    rho0, srho = 0.6, 0.55  # ρ(ξ) = 0.6 + 0.55 ξ > 0 at the nodes

    def state(rho):
        return euler_state(rho, v1 + 0 * rho, v2 + 0 * rho, p + 0 * rho)

    u = np.zeros((nx, n, n, 4))
    for i in range(n):
        u[:, :, i] = state(np.array(rho0 + srho * nodes[i]))
    o.set_alpha(np.ones(nx))
    o.rhs(u, 0.0, 0.0)
    fl = fn_low(o)
    # subcell 1 (0-based): faces xs[1] (left) and xs[2] (right)
    i = 1
    rho_i = rho0 + srho * nodes[i]
    xl, xr = xs[i] - nodes[i], xs[i + 1] - nodes[i]
    assert min(rho_i + 2 * srho * xl, rho_i + 2 * srho * xr) < 1e-10  # the limiter acts here
    theta = min(0.9 * rho_i / (abs(2 * srho * xl) + 1e-13), 0.9 * rho_i / (abs(2 * srho * xr) + 1e-13), 1.0)
    left_face_of_1 = state(np.array(rho_i + theta * srho * xl))
    expect = rusanov(state(np.array(rho0 + srho * nodes[0])), left_face_of_1, 1)  # end subcell 0 keeps its value
    for e in range(nx):
        for j in range(n):
            np.testing.assert_allclose(fl[e, 0, j], expect, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("kind", ["lw", "mdrk1", "mdrk2"])
def test_mh_conservation(oracle_mod, N, kind):
    from trixilw_hip.equations import IsentropicVortex

    nx = 5
    o = mh_oracle(oracle_mod, N, nx, nx, 20.0)
    x, y = o.node_coords()
    rng = np.random.default_rng(30 + N)
    u = IsentropicVortex(center=(10.0, 10.0), period=(20.0, 20.0))(x, y, 0.0)
    u = u * (1 + 1e-2 * rng.standard_normal(x.shape + (4,)))
    o.set_alpha(rng.random(o.E))
    dt = 0.4 * o.max_dt(u)
    if kind == "mdrk2":
        o.rhs(u, 0.0, dt, kind="mdrk1")
        o.set_us(u + 1e-3 * rng.standard_normal(u.shape))
    du = o.rhs(u, 0.0, dt, kind=kind)
    _, w, _, _ = o.basis()
    W = np.outer(w, w)
    total = np.einsum("ji,ejiv->v", W, du)
    scale = np.einsum("ji,ejiv->v", W, np.abs(du))
    assert np.all(np.abs(total) <= 1e-13 * scale)


def test_boundary_elements_use_first_order_and_shock_stays_admissible(oracle_mod):
    from trixilw_hip.equations import FourQuadrantRiemann

    ic = FourQuadrantRiemann()
    nx, N = 16, 4
    kw = dict(bc_kind=ic.kind, bc_par=ic.params, alpha_max=0.5)
    mh = mh_oracle(oracle_mod, N, nx, nx, 1.0, periodic=False, **kw)
    fo = mh_oracle(oracle_mod, N, nx, nx, 1.0, periodic=False, rec="first_order", **kw)
    x, y = mh.node_coords()
    u = ic(x, y, 0.0)
    rng = np.random.default_rng(4)
    alpha = 0.3 + 0.7 * rng.random(nx * nx)
    for o in (mh, fo):
        o.set_alpha(alpha)
    dt = 0.3 * mh.max_dt(u)
    dm, df = mh.rhs(u, 0.0, dt), fo.rhs(u, 0.0, dt)
    nb = mh.neighbors()
    edge = np.any(nb < 0, axis=1)
    # the boundary elements' own FV terms agree; their neighbours' fn_low may not, so compare
    # fn_low of the boundary elements, which depends on their own subcell fluxes only
    assert np.array_equal(fn_low(mh)[edge], fn_low(fo)[edge])
    assert np.max(np.abs(fn_low(mh)[~edge] - fn_low(fo)[~edge])) > 1e-6
    assert np.max(np.abs(dm - df)) > 1e-6
    for o in (mh, fo):
        o.set_alpha(None)
    um, t, dts = mh.solve(u, 0.0, 1e9, 0.3, max_steps=150)
    uf, tf_, _ = fo.solve(u, 0.0, 1e9, 0.3, max_steps=150)
    rho = um[..., 0]
    p = 0.4 * (um[..., 3] - 0.5 * (um[..., 1] ** 2 + um[..., 2] ** 2) / rho)
    assert np.isfinite(um).all() and rho.min() > 0 and p.min() > 0
    assert np.max(np.abs(um - uf)) > 1e-3  # the reconstruction changes the solution
