"""Pins of the oracle's two conditions on uⁿ with a normal direction,
boundary_condition_outflow and slip_wall_approximate (src/equations/
equations.jl:185-197, 251-260), on the curvilinear mesh: the boundary
surface fluxes against numpy evaluations of their definitions at the boundary
nodes (the node's own contravariant vector as the normal, the outer state on
the side of the boundary), a uniform flow kept by outflow on every face, a
fluid at rest kept by the wall, and the wall's exact conservation of mass and
energy (its mass and energy fluxes vanish)."""
import numpy as np
import pytest

GAMMA = 1.4


def warped_open(oracle_mod, N, nx, kind, L=20.0, A=0.05):
    return oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, L / nx, mapping=(oracle_mod.MAP_WARP, (L, A)),
                             periodic=(False, False), bc_kind=kind)


def prim2cons(rho, v1, v2, p):
    return np.stack([rho, rho * v1, rho * v2, p / (GAMMA - 1) + 0.5 * rho * (v1**2 + v2**2)], axis=-1)


def state(x, y, seed=0):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.2 * np.sin(0.3 * x) * np.cos(0.25 * y) + 0.01 * rng.standard_normal(x.shape)
    v1 = 0.4 + 0.1 * np.cos(0.2 * y)
    v2 = -0.3 + 0.1 * np.sin(0.15 * x)
    p = 1.0 + 0.1 * np.cos(0.2 * (x + y))
    return prim2cons(rho, v1, v2, p)


def normal_flux(u, n1, n2):
    rho = u[..., 0]
    v1, v2 = u[..., 1] / rho, u[..., 2] / rho
    p = (GAMMA - 1) * (u[..., 3] - 0.5 * rho * (v1**2 + v2**2))
    vn = v1 * n1 + v2 * n2
    return np.stack([rho * vn, rho * vn * v1 + p * n1, rho * vn * v2 + p * n2, (u[..., 3] + p) * vn], axis=-1)


def boundary_nodes(nx, n):
    """(element, direction 1..4, face node l, node i, j) of every boundary face node."""
    out = []
    for e in range(nx * nx):
        ix, iy = e % nx, e // nx
        for d, on in ((1, ix == 0), (2, ix == nx - 1), (3, iy == 0), (4, iy == nx - 1)):
            if not on:
                continue
            for l in range(n):
                i, j = {1: (0, l), 2: (n - 1, l), 3: (l, 0), 4: (l, n - 1)}[d]
                out.append((e, d, l, i, j))
    return out


def expected_flux(kind, u, Ja, nx, n, oracle_mod):
    """numpy evaluation of the two conditions at every boundary face node."""
    got = []
    for e, d, l, i, j in boundary_nodes(nx, n):
        J = Ja[e, j, i]
        n1, n2 = (J[0], J[1]) if d <= 2 else (J[2], J[3])
        ui = u[e, j, i]
        if kind == oracle_mod.BV_OUTFLOW:
            got.append(normal_flux(ui, n1, n2))
            continue
        rho = ui[0]
        v1, v2 = ui[1] / rho, ui[2] / rho
        p = (GAMMA - 1) * (ui[3] - 0.5 * rho * (v1**2 + v2**2))
        vn = (v1 * n1 + v2 * n2) / (n1 * n1 + n2 * n2)
        ur = prim2cons(rho, v1 - 2 * vn * n1, v2 - 2 * vn * n2, p)
        ul, uq = (ui, ur) if d % 2 == 0 else (ur, ui)
        c = np.sqrt(GAMMA * p / rho) * np.hypot(n1, n2)
        lam = abs(v1 * n1 + v2 * n2) + c
        got.append(0.5 * (normal_flux(ul, n1, n2) + normal_flux(uq, n1, n2)) - 0.5 * lam * (uq - ul))
    return np.array(got)


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("which", ["outflow", "slip_wall_approximate"])
def test_boundary_fluxes_follow_their_definitions(oracle_mod, N, which):
    kind = oracle_mod.BV_OUTFLOW if which == "outflow" else oracle_mod.BV_SLIP_WALL_APPROX
    nx, n = 5, N + 1
    o = warped_open(oracle_mod, N, nx, kind)
    x, y = o.node_coords()
    u = state(x, y)
    o.rhs(u, 0.3, 0.5 * o.max_dt(u))
    sfv = o.aux("surface_flux_values").reshape(nx * nx, 4, n, 4)
    Ja = o.aux("contravariant_vectors").reshape(nx * nx, n, n, 4)
    got = np.array([sfv[e, d - 1, l] for e, d, l, _, _ in boundary_nodes(nx, n)])
    want = expected_flux(kind, u, Ja, nx, n, oracle_mod)
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
    if which == "slip_wall_approximate":  # no mass or energy crosses the wall
        assert np.max(np.abs(got[:, [0, 3]])) <= 1e-13 * np.max(np.abs(got))


@pytest.mark.parametrize("N", [3, 4])
def test_outflow_keeps_a_uniform_flow(oracle_mod, N):
    o = warped_open(oracle_mod, N, 6, oracle_mod.BV_OUTFLOW)
    x, _ = o.node_coords()
    u = prim2cons(np.full(x.shape, 1.3), np.full(x.shape, 0.6), np.full(x.shape, -0.4), np.full(x.shape, 0.8))
    du = o.rhs(u, 0.0, 0.5 * o.max_dt(u))
    assert np.max(np.abs(du)) <= 1e-11 * np.max(np.abs(u))


@pytest.mark.parametrize("N", [3, 4])
def test_approximate_wall_keeps_rest_and_conserves(oracle_mod, N):
    o = warped_open(oracle_mod, N, 6, oracle_mod.BV_SLIP_WALL_APPROX)
    x, y = o.node_coords()
    rest = prim2cons(np.full(x.shape, 1.1), np.zeros(x.shape), np.zeros(x.shape), np.full(x.shape, 0.9))
    assert np.max(np.abs(o.rhs(rest, 0.0, 0.5 * o.max_dt(rest)))) <= 1e-11
    u0 = state(x, y, seed=3)
    w = np.outer(*(2 * [o.basis()[1]]))[None] / o.aux("inverse_jacobian").reshape(x.shape)
    u, t, dts = o.solve(u0, 0.0, 1e9, 0.3 * 0.103 * 5, max_steps=10)
    assert len(dts) == 10
    m0 = np.einsum("eji,ejiv->v", w, u0)
    m1 = np.einsum("eji,ejiv->v", w, u)
    scale = np.einsum("eji,ejiv->v", w, np.abs(u0))
    assert abs(m1[0] - m0[0]) <= 1e-13 * scale[0] and abs(m1[3] - m0[3]) <= 1e-13 * scale[3]
    assert np.max(np.abs(m1[1:3] - m0[1:3])) > 1e-8 * scale[0]  # momentum is not conserved at walls


def test_conditions_need_a_curvilinear_mesh(oracle_mod):
    for kind in (oracle_mod.BV_OUTFLOW, oracle_mod.BV_SLIP_WALL_APPROX):
        with pytest.raises(ValueError):
            oracle_mod.Oracle(4, 4, 4, 0.0, 0.0, 0.25, periodic=(False, False), bc_kind=kind)
