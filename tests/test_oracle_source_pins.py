"""Pins of the oracle's source terms (TrixiLW folds `source_terms` into the
Taylor procedure, src/solvers/dgsem_tree/dg_2d.jl:89-176 with the kernels'
calls at 980-987, 1033-1045, 1093-1103, 1148-1160, 1228-1247, and the MDRK
stages of tree/multi_derivative_2d.jl:323-379, 487-522):

- on a uniform periodic state every flux derivative vanishes, so the residual
  is the source part S alone.  For a source that depends on t only, S must be
  the exact time average of s over [t, t + dt] whenever the finite differences
  are exact: cubics for LW N = 3, 4 and for the whole MDRK step, linear
  functions for MDRK stage 1 (the average over [t, t + dt/2]).  A dropped or
  mis-weighted Taylor term, a wrong time offset or sign fails these;
- for a state-dependent source (linear damping s = -k u) the stage states are
  the exact Taylor states of u' = -k u, so every finite difference of s is the
  exact derivative except where the reference passes the stencil states in
  another order; the closed form of that call (evaluated here from the Taylor
  states) is what the oracle must give, and it differs from the exact series;
- the manufactured solution of Trixi's convergence test (the reference's
  elixir_euler_source_terms*.jl) converges at high order with the built-in
  source, periodic and with its own Dirichlet data.
"""
import math

import numpy as np
import pytest

from conftest import rel_l2


def uniform_state(o, rho=1.2, v1=0.3, v2=-0.2, p=0.9, gamma=1.4):
    x, _ = o.node_coords()
    u = np.empty(x.shape + (4,))
    u[..., 0] = rho
    u[..., 1] = rho * v1
    u[..., 2] = rho * v2
    u[..., 3] = p / (gamma - 1) + 0.5 * rho * (v1 * v1 + v2 * v2)
    return u


# a cubic in t per variable: s_v(t) = c0 + c1 t + c2 t² + c3 t³
COEF = np.array([[0.3, -0.7, 0.2, 0.05], [-0.1, 0.4, -0.3, 0.02], [0.25, 0.1, 0.15, -0.04], [0.5, -0.2, 0.1, 0.03]])


def poly_source(deg):
    c = COEF[:, : deg + 1]

    def s(u, x, y, t):
        return c @ (t ** np.arange(deg + 1))
    return s


def poly_average(deg, t0, t1):
    """(1/(t1-t0)) ∫ s dt for the cubic above truncated to `deg`."""
    k = np.arange(deg + 1)
    prim = (t1 ** (k + 1) - t0 ** (k + 1)) / (k + 1)
    return COEF[:, : deg + 1] @ prim / (t1 - t0)


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_lw_uniform_state_gets_exact_time_average(oracle_mod, N, deg):
    o = oracle_mod.Oracle(N, 3, 3, 0.0, 0.0, 1.0, source_terms=poly_source(deg))
    u = uniform_state(o)
    t, dt = 0.37, 0.11
    du = o.rhs(u, t, dt)
    want = poly_average(deg, t, t + dt)
    assert np.max(np.abs(du - want)) <= 1e-13 * np.max(np.abs(want))


def test_lw_quartic_source_shows_the_three_point_second_difference(oracle_mod):
    """s = t⁴: s_t (five points), s_ttt and s_tttt are exact for a quartic,
    the three-point s_tt (calc_source_tt_N23) is off by dt⁴·2, so S exceeds the
    exact average by (2 dt⁴)/6 at N = 3 and 4: the test tells the reference's
    three-point s_tt from a five-point one."""
    def s(u, x, y, t):
        return np.full(4, t ** 4)
    t, dt = 0.2, 0.3
    exact = ((t + dt) ** 5 - t ** 5) / 5 / dt
    for N in (3, 4):
        o = oracle_mod.Oracle(N, 3, 3, 0.0, 0.0, 1.0, source_terms=s)
        du = o.rhs(uniform_state(o), t, dt)
        extra = 2 * dt ** 4 / 6
        if N == 3:  # no s_tttt term: S misses the exact dt⁴/120 · 24 of the quartic
            extra -= dt ** 4 * 24 / 120
        assert np.max(np.abs(du - (exact + extra))) <= 1e-13


@pytest.mark.parametrize("N", [3, 4])
def test_mdrk_stages_on_uniform_state(oracle_mod, N):
    t, dt = 0.37, 0.11
    # stage 1 is the average over [t, t + dt/2] (times 1/2) for a linear source
    o = oracle_mod.Oracle(N, 3, 3, 0.0, 0.0, 1.0, source_terms=poly_source(1))
    u = uniform_state(o)
    r1 = o.rhs(u, t, dt, kind="mdrk1")
    want1 = 0.5 * poly_average(1, t, t + 0.5 * dt)
    assert np.max(np.abs(r1 - want1)) <= 1e-13 * np.max(np.abs(want1))
    # the whole step is the exact average for a cubic
    o = oracle_mod.Oracle(N, 3, 3, 0.0, 0.0, 1.0, source_terms=poly_source(3))
    r1 = o.rhs(u, t, dt, kind="mdrk1")
    o.set_us(u + dt * r1)
    r2 = o.rhs(u, t, dt, kind="mdrk2")
    want = poly_average(3, t, t + dt)
    assert np.max(np.abs(r2 - want)) <= 1e-13 * np.max(np.abs(want))


@pytest.mark.parametrize("N", [3, 4])
def test_state_dependent_source_closed_form(oracle_mod, N):
    """s(u) = -k u on a uniform state: ut = -k dt u, utt = (k dt)² u, ... exactly.
    The reference evaluates s_tttt (N = 4) with the states (u, up, um, upp, umm)
    and s_ttt (N = 3) with (u, up, upp, upp, umm) in the calls' argument order;
    the closed form below follows those calls."""
    k = 1.3
    t, dt = 0.1, 0.07
    o = oracle_mod.Oracle(N, 3, 3, 0.0, 0.0, 1.0, source_terms=lambda u, x, y, tt: -k * u)
    u = uniform_state(o)
    du = o.rhs(u, t, dt)
    u0 = u[0, 0, 0]
    a = -k * dt
    d = [u0 * a ** m for m in range(1, 5)]  # ut, utt, uttt, utttt (dt-scaled derivatives)
    s = lambda w: -k * w
    # s_t, s_tt exact: dt s' = -k ut, dt² s'' = -k utt
    S = s(u0) + 0.5 * (-k * d[0]) + (-k * d[1]) / 6
    if N == 4:
        S = S + (-k * d[2]) / 24  # s_ttt of the stage-3 states is exact
        up = u0 + d[0] + d[1] / 2 + d[2] / 6 + d[3] / 24
        um = u0 - d[0] + d[1] / 2 - d[2] / 6 + d[3] / 24
        upp = u0 + 2 * d[0] + 2 * d[1] + 4 * d[2] / 3 + 2 * d[3] / 3
        umm = u0 - 2 * d[0] + 2 * d[1] - 4 * d[2] / 3 + 2 * d[3] / 3
        # calc_source_tttt_N4(u, up, upp, um, umm) called with (u, up, um, upp, umm)
        s4 = s(um) - 4 * s(up) + 6 * s(u0) - 4 * s(upp) + s(umm)
        S = S + s4 / 120
        exact_series = -k * u0 * sum((-k * dt) ** m / math.factorial(m + 1) for m in range(5))
    else:
        up = u0 + d[0] + d[1] / 2 + d[2] / 6
        upp = u0 + 2 * d[0] + 2 * d[1] + 4 * d[2] / 3
        umm = u0 - 2 * d[0] + 2 * d[1] - 4 * d[2] / 3
        # calc_source_ttt_N34(u, up, upp, um, umm) called with (u, up, upp, upp, umm)
        s3 = 0.5 * (s(upp) - 2 * s(up) + 2 * s(upp) - s(umm))
        S = S + s3 / 24
        exact_series = -k * u0 * sum((-k * dt) ** m / math.factorial(m + 1) for m in range(4))
    assert np.max(np.abs(du - S)) <= 1e-13 * np.max(np.abs(S))
    # the argument order makes the last term O(dt) instead of O(dt⁴)
    assert np.max(np.abs(S - exact_series)) > 1e-4 * np.max(np.abs(exact_series))


def convergence_test_state(x, y, t):
    phi = 2.0 + 0.1 * np.sin(np.pi * (x + y - t))
    return np.stack([phi, phi, phi, phi * phi], axis=-1)


def manufactured_errors(oracle_mod, N, td, levels, periodic=True, tf=0.5):
    import trixilw_hip as tlw

    solver = tlw.DGSEM(N, volume_integral=tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW()))
    cfl = tlw.trixi2lw(0.9, solver) * (0.5 if td == "mdrk" else 1.0)
    errs = []
    for nx in levels:
        kw = {} if periodic else dict(periodic=(False, False), bc_kind=oracle_mod.BV_CONVERGENCE_TEST)
        o = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, 2.0 / nx, source_terms="convergence_test", **kw)
        x, y = o.node_coords()
        u, t, _ = o.solve(convergence_test_state(x, y, 0.0), 0.0, tf, cfl, mdrk=(td == "mdrk"))
        assert abs(t - tf) < 1e-12
        errs.append(rel_l2(u, convergence_test_state(x, y, tf)))
    return np.array(errs)


def fitted_order(errs):
    """Least-squares slope of log2(error) against the refinement level."""
    return -np.polyfit(np.arange(len(errs)), np.log2(errs), 1)[0]


@pytest.mark.parametrize("N,td,eoc_min", [(4, "lw", 4.3), (3, "lw", 3.6), (3, "mdrk", 3.6), (4, "mdrk", 4.3)])
def test_manufactured_solution_converges(oracle_mod, N, td, eoc_min):
    """4² … 32² to t = 0.5: seen 4.56 (LW N=4), 3.86 (LW N=3), 3.82 (MDRK N=3),
    4.57 (MDRK N=4)."""
    errs = manufactured_errors(oracle_mod, N, td, (4, 8, 16, 32))
    assert fitted_order(errs) > eoc_min, errs
    # without the source the same run does not approach the manufactured solution
    o = oracle_mod.Oracle(N, 8, 8, 0.0, 0.0, 0.25)
    x, y = o.node_coords()
    u, _, _ = o.solve(convergence_test_state(x, y, 0.0), 0.0, 0.5, 0.01, mdrk=(td == "mdrk"))
    assert rel_l2(u, convergence_test_state(x, y, 0.5)) > 100 * errs[1]


@pytest.mark.parametrize("N", [3, 4])
def test_manufactured_solution_with_dirichlet_data(oracle_mod, N):
    """elixir_euler_source_terms_nonperiodic.jl: BoundaryConditionDirichlet of
    the convergence test's state on every face."""
    errs = manufactured_errors(oracle_mod, N, "lw", (4, 8, 16, 32), periodic=False)
    assert fitted_order(errs) > (4.3 if N == 4 else 3.6), errs


def test_sources_are_refused_where_not_restated(oracle_mod):
    # (the curvilinear LW kernels take sources: tests/test_oracle_curved_source_pins.py)
    warp = (oracle_mod.MAP_WARP, (2.0, 0.05))
    for kw in (dict(equation="linear_advection"), dict(parabolic="navier_stokes", mu=0.01),
               dict(mapping=warp, shock_capturing=True)):
        with pytest.raises(ValueError):
            oracle_mod.Oracle(3, 4, 4, 0.0, 0.0, 0.5, source_terms="convergence_test", **kw)
