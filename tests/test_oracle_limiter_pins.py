"""Pins of the oracle's positivity limiter (no GPU).

my_limiter_zhang_shu! (src/callbacks/callbacks_stage/positivity_zhang_shu_dg2d.jl:8-45)
with variables (density, pressure) applied in turn
(positivity_zhang_shu.jl:26-41).  The test checks the limiter's defining
properties on states built to trip each branch, instead of retyping it:
untouched elements stay bit-identical, element means are conserved, a
density-limited element has its minimum density exactly at the threshold,
pressure ends at or above its threshold (Jensen), and every limited element
is a single contraction u - ū -> θ (u - ū) toward its mean with θ in [0, 1].
"""
import numpy as np
import pytest

from conftest import rel_linf

THR = (1e-3, 2e-3)


def state(E, n, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.5, 1.0, (E, n, n))
    v1 = rng.uniform(-0.5, 0.5, (E, n, n))
    v2 = rng.uniform(-0.5, 0.5, (E, n, n))
    p = rng.uniform(0.5, 1.0, (E, n, n))
    return rho, v1, v2, p


def cons(rho, v1, v2, p):
    return np.stack([rho, rho * v1, rho * v2, p / 0.4 + 0.5 * rho * (v1 ** 2 + v2 ** 2)], -1)


def pressure(u):
    return 0.4 * (u[..., 3] - 0.5 * (u[..., 1] ** 2 + u[..., 2] ** 2) / u[..., 0])


@pytest.mark.parametrize("N", [3, 4])
def test_zhang_shu_limiter_properties(oracle_mod, N):
    n, nx = N + 1, 4
    o = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, 1.0 / nx, positivity=THR)
    E = o.E
    rho, v1, v2, p = state(E, n, 100 + N)
    rho[1, 2, 3] = -0.05          # density below threshold at one node
    rho[6, 0, 0] = 0.0005         # density just below threshold
    p[9, 1, 1] = -0.2             # negative pressure, density fine
    p[13, n - 1, 0] = 0.001       # pressure just below threshold
    rho[14, 2, 2], p[14, 3, 1] = -0.01, -0.1  # both
    u = cons(rho, v1, v2, p)
    # nodes with negative density carry momentum 0 so that p stays finite
    u[1, 2, 3, 1:3] = 0.0
    u[14, 2, 2, 1:3] = 0.0
    lim = o.positivity_limiter(u)
    _, w, _, _ = o.basis()
    W = np.outer(w, w)
    mean = np.einsum("ji,ejiv->ev", W, u) / 4
    mean_new = np.einsum("ji,ejiv->ev", W, lim) / 4
    touched = [1, 6, 9, 13, 14]
    for e in range(E):
        if e not in touched:
            assert np.array_equal(lim[e], u[e]), e
            continue
        assert np.max(np.abs(mean_new[e] - mean[e])) <= 1e-14 * np.max(np.abs(mean[e]))
        d0, d1 = u[e] - mean[e], lim[e] - mean[e]
        k = np.unravel_index(np.argmax(np.abs(d0)), d0.shape)
        theta = d1[k] / d0[k]
        assert 0.0 <= theta < 1.0
        assert np.max(np.abs(d1 - theta * d0)) <= 1e-13 * np.max(np.abs(d0))
        assert lim[e][..., 0].min() >= THR[0] * (1 - 1e-12)
        assert pressure(lim[e]).min() >= THR[1] * (1 - 1e-9)
    for e in (1, 6):  # limited by density only: the minimum sits at the threshold
        assert pressure(lim[e]).min() > THR[1]
        assert abs(lim[e][..., 0].min() - THR[0]) <= 1e-14
    # a second application changes nothing (the solve loop applies it twice per step)
    assert np.max(np.abs(o.positivity_limiter(lim) - lim)) <= 1e-15


def test_limited_solve_survives_riemann_at_lw_cfl(oracle_mod):
    """The four-quadrant Riemann problem with shock capturing at 0.98·LW CFL on
    32²: without the limiter the state breaks down within 60 steps, with it the
    state stays admissible."""
    import trixilw_hip as tlw

    ic = tlw.FourQuadrantRiemann()
    nx = 32
    o = oracle_mod.Oracle(4, nx, nx, 0.0, 0.0, 1.0 / nx, periodic=(False, False), bc_kind=ic.kind,
                          bc_par=ic.params, shock_capturing=True, alpha_max=0.5, positivity=(5e-6, 5e-6))
    x, y = o.node_coords()
    plain = oracle_mod.Oracle(4, nx, nx, 0.0, 0.0, 1.0 / nx, periodic=(False, False), bc_kind=ic.kind,
                              bc_par=ic.params, shock_capturing=True, alpha_max=0.5)
    u_plain, _, _ = plain.solve(ic(x, y, 0.0), 0.0, 1e9, 0.98 * 5 * 0.103, max_steps=60)
    assert not np.isfinite(u_plain).all()
    u, t, dts = o.solve(ic(x, y, 0.0), 0.0, 1e9, 0.98 * 5 * 0.103, max_steps=60)
    assert len(dts) == 60 and np.isfinite(u).all()
    assert u[..., 0].min() >= 5e-6 * (1 - 1e-12)
    assert pressure(u).min() >= 5e-6 * (1 - 1e-9)
