"""Pins of the oracle's reflecting wall (boundary_condition_slip_wall_vertical
on x faces, _horizontal on y faces, src/equations/equations.jl:95-139) on the
TreeMesh (no GPU).

A state on the periodic [-1, 1]² that is mirror-symmetric about x = 0 and
y = 0 (ρ, p even, v1 odd in x, v2 odd in y) stays symmetric, and its residual
restricted to [0, 1]² is the residual of the walled quarter: at the symmetry
lines the periodic interface flux sees the mirrored state and the mirrored
time-averaged flux, which is exactly what the wall builds.  This holds for LW
and both MDRK stages, to the round-off of the mirrored element computations.
The wall also conserves mass and energy exactly."""
import numpy as np
import pytest

from conftest import rel_linf


def symmetric_state(x, y, gamma=1.4):
    rho = 1.0 + 0.2 * np.cos(np.pi * x) * np.cos(np.pi * y)
    v1 = 0.3 * np.sin(np.pi * x) * (1.0 + 0.5 * np.cos(np.pi * y))
    v2 = -0.25 * np.sin(np.pi * y) * (1.0 + 0.4 * np.cos(np.pi * x))
    p = 1.0 + 0.1 * np.cos(np.pi * x) * np.cos(2 * np.pi * y)
    return np.stack([rho, rho * v1, rho * v2, p / (gamma - 1) + 0.5 * rho * (v1 ** 2 + v2 ** 2)], axis=-1)


def quarter_rows(nx_full):
    """Elements of the periodic mesh that cover [0, 1]², in the quarter mesh's order."""
    h = nx_full // 2
    return [ix + nx_full * iy for iy in range(h, nx_full) for ix in range(h, nx_full)]


@pytest.mark.parametrize("N", [3, 4])
def test_slip_wall_equals_mirror_symmetric_periodic_solution(oracle_mod, N):
    nx = 8
    full = oracle_mod.Oracle(N, nx, nx, -1.0, -1.0, 2.0 / nx)
    quarter = oracle_mod.Oracle(N, nx // 2, nx // 2, 0.0, 0.0, 2.0 / nx, periodic=(False, False),
                                bc_kind=oracle_mod.BV_SLIP_WALL)
    xf, yf = full.node_coords()
    xq, yq = quarter.node_coords()
    rows = quarter_rows(nx)
    uf = symmetric_state(xf, yf)
    uq = symmetric_state(xq, yq)
    assert np.max(np.abs(uf[rows] - uq)) < 1e-15  # the same nodes
    dt = 0.8 * full.max_dt(uf)
    assert rel_linf(quarter.rhs(uq, 0.0, dt)[...], full.rhs(uf, 0.0, dt)[rows]) < 1e-13
    d1f = full.rhs(uf, 0.0, dt, kind="mdrk1")
    d1q = quarter.rhs(uq, 0.0, dt, kind="mdrk1")
    assert rel_linf(d1q, d1f[rows]) < 1e-13
    full.set_us(uf + 0.5 * dt * d1f)
    quarter.set_us(uq + 0.5 * dt * d1q)
    assert rel_linf(quarter.rhs(uq, 0.0, dt, kind="mdrk2"), full.rhs(uf, 0.0, dt, kind="mdrk2")[rows]) < 1e-13


@pytest.mark.parametrize("N", [3, 4])
def test_slip_wall_conserves_mass_and_energy(oracle_mod, N):
    nx = 6
    o = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, 1.0 / nx, periodic=(False, False), bc_kind=oracle_mod.BV_SLIP_WALL)
    x, y = o.node_coords()
    rng = np.random.default_rng(11 + N)
    u = symmetric_state(x + 0.37, y - 0.21) * (1 + 1e-2 * rng.standard_normal(x.shape + (4,)))
    du = o.rhs(u, 0.0, 0.5 * o.max_dt(u))
    _, w, _, _ = o.basis()
    W = np.outer(w, w)[None]
    total = np.einsum("eji,ejiv->v", W, du)
    scale = np.einsum("eji,ejiv->v", W, np.abs(du))
    assert abs(total[0]) <= 1e-13 * scale[0] and abs(total[3]) <= 1e-13 * scale[3]
    assert abs(total[1]) > 1e-6 * scale[1]  # the walls push: momentum is not conserved
