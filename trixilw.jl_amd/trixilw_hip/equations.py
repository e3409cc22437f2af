"""Equations and initial/boundary value functions (host side, numpy).

`CompressibleEulerEquations2D` mirrors Trixi's type of the same name as used
by TrixiLW (src/equations/compressible_euler_2d.jl); the kernels evaluate the
flux on the device, this class only converts variables and evaluates initial
and boundary data on the host.

Boundary-value functions that the device must evaluate (time-averaged
Dirichlet condition, src/equations/equations.jl:6-37) are described by a kind
code and a parameter vector shared with the HIP library.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

BV_NONE, BV_VORTEX, BV_CONSTANT, BV_RIEMANN4, BV_SLIP_WALL, BV_TABLE, BV_CONVERGENCE_TEST = 0, 1, 2, 3, 4, 5, 6
BV_OUTFLOW, BV_SLIP_WALL_APPROXIMATE = 7, 8
SRC_NONE, SRC_CONVERGENCE_TEST, SRC_NS_CONVERGENCE_TEST = 0, 1, 2


class BoundaryConditionDirichlet:
    """Trixi BoundaryConditionDirichlet(boundary_value_function) with the LW
    time average of src/equations/equations.jl:6-37, for any function:
    ``boundary_value_function(x, t, equations)`` takes x = (x, y) as numpy
    arrays and returns the conservative states [..., 4].  The host evaluates it
    at the boundary face nodes and uploads the values before each rhs call
    (TLW_BV_TABLE); ``steady=True`` declares a function that does not depend on
    t: it is evaluated once and the device loop can then run with it."""

    kind = BV_TABLE
    params = ()

    def __init__(self, boundary_value_function, steady=False):
        if not callable(boundary_value_function):
            raise TypeError("boundary_value_function must be callable")
        self.boundary_value_function = boundary_value_function
        self.steady = bool(steady)

    def values(self, x, y, t, equations=None):
        u = np.asarray(self.boundary_value_function((x, y), t, equations), dtype=np.float64)
        if u.shape != np.shape(x) + (4,):
            raise ValueError(f"boundary_value_function must return shape {np.shape(x) + (4,)}, got {u.shape}")
        return u


class _SlipWall:
    """The reference's reflecting wall of the LW TreeMesh path
    (src/equations/equations.jl:95-139): boundary_condition_slip_wall_vertical
    mirrors uⁿ and Ũ in v1 and keeps only the x-momentum component of the
    normal flux, _horizontal does the same in y.  Pass the vertical one for the
    x faces and the horizontal one for the y faces (one kind per mesh)."""

    kind = BV_SLIP_WALL
    params = ()

    def __init__(self, orientation):
        self.orientation = orientation

    def __repr__(self):
        return "boundary_condition_slip_wall_" + ("vertical" if self.orientation == 1 else "horizontal")


boundary_condition_slip_wall_vertical = _SlipWall(1)
boundary_condition_slip_wall_horizontal = _SlipWall(2)


class _Periodic:
    """Trixi's boundary_condition_periodic: allowed in the per-direction forms
    for the directions in which the mesh is periodic (no boundary faces)."""

    def __repr__(self):
        return "boundary_condition_periodic"


boundary_condition_periodic = _Periodic()


class _NormalCondition:
    """A condition on uⁿ alone with the boundary node's normal direction, as
    the reference's P4est boundary call uses it (src/solvers/dgsem_p4est/
    dg_2d.jl:361-393): boundary_condition_outflow (equations.jl:185-197) is
    the normal flux of uⁿ, slip_wall_approximate (equations.jl:251-260) the
    Rusanov flux between uⁿ and its reflection about the normal.  The device
    evaluates them on the curvilinear StructuredMesh (Euler, every boundary
    face), with the element's contravariant vector at the node as normal."""

    params = ()

    def __init__(self, kind, name):
        self.kind = kind
        self._name = name

    def __repr__(self):
        return self._name


boundary_condition_outflow = _NormalCondition(BV_OUTFLOW, "boundary_condition_outflow")
slip_wall_approximate = _NormalCondition(BV_SLIP_WALL_APPROXIMATE, "slip_wall_approximate")


class NoSlip:
    """Trixi NoSlip(boundary_value_function): the wall velocity (x, t, equations)
    -> (v1, v2).  The device takes a constant wall velocity: a pair is taken as
    is; a callable is evaluated by the semidiscretization at every boundary face
    node of its walls (at t = 0 and t = 1) and rejected unless it gives one
    value there."""

    def __init__(self, velocity):
        self.function = velocity if callable(velocity) else None
        self.velocity = None if callable(velocity) else (float(velocity[0]), float(velocity[1]))

    def resolve(self, points):
        """The constant wall velocity on the face-node coordinates `points`
        [(x, y), ...] of the walls it is used on."""
        if self.function is None:
            return self.velocity
        if not points:  # no boundary face in this direction (periodic): the value is unused
            return tuple(float(v) for v in self.function((0.0, 0.0), 0.0, None))
        vals = {tuple(float(v) for v in self.function((x, y), t, None)) for (x, y) in points for t in (0.0, 1.0)}
        if len(vals) != 1:
            raise NotImplementedError("the device path takes a constant wall velocity; this function varies "
                                      "over the wall's face nodes or in time")
        return vals.pop()


class Adiabatic:
    """Trixi Adiabatic(boundary_value_normal_flux_function): the normal heat flux;
    the reference's wall condition for conservative gradient variables takes it as
    given, and the device path supports the insulated wall (zero flux)."""

    def __init__(self, normal_flux=0.0):
        if callable(normal_flux):
            normal_flux = normal_flux((0.0, 0.0), 0.0, None)
        if float(normal_flux) != 0.0:
            raise NotImplementedError("the device path implements the insulated wall (zero normal heat flux)")


class Isothermal:
    """Trixi Isothermal(boundary_value_function): the wall temperature T = p/ρ
    (the reference's wall condition takes R = 1); the device takes a constant
    temperature (a callable is evaluated once and must not depend on x or t)."""

    def __init__(self, temperature):
        self.function = temperature if callable(temperature) else None
        self.temperature = None if callable(temperature) else float(temperature)

    def resolve(self, points):
        """The constant wall temperature on the face-node coordinates `points`
        (evaluated at t = 0 and t = 1; a varying function is rejected)."""
        if self.function is None:
            return self.temperature
        if not points:  # no boundary face in this direction (periodic): the value is unused
            return float(self.function((0.0, 0.0), 0.0, None))
        vals = {float(self.function((x, y), t, None)) for (x, y) in points for t in (0.0, 1.0)}
        if len(vals) != 1:
            raise NotImplementedError("the device path takes a constant wall temperature; this function varies "
                                      "over the wall's face nodes or in time")
        return vals.pop()


class BoundaryConditionNavierStokesWall:
    """Trixi BoundaryConditionNavierStokesWall(velocity, heat) with NoSlip and
    Adiabatic or Isothermal (conservative gradient variables:
    src/equations/compressible_navier_stokes_2d.jl:81-124; the reference's
    lid-driven cavity uses Adiabatic)."""

    def __init__(self, boundary_condition_velocity, boundary_condition_heat_flux):
        if not isinstance(boundary_condition_velocity, NoSlip) or not isinstance(boundary_condition_heat_flux,
                                                                                 (Adiabatic, Isothermal)):
            raise NotImplementedError("the device path implements NoSlip velocity with Adiabatic or Isothermal walls")
        self.boundary_condition_velocity = boundary_condition_velocity
        self.boundary_condition_heat_flux = boundary_condition_heat_flux

    @classmethod
    def periodic(cls):
        """The stand-in of a periodic direction (no wall faces there): its values are unused."""
        return cls(NoSlip((0.0, 0.0)), Adiabatic(0.0))

    def resolve(self, points):
        """(wall velocity, wall temperature or None) on the wall's face nodes."""
        heat = self.boundary_condition_heat_flux
        return (self.boundary_condition_velocity.resolve(points),
                heat.resolve(points) if isinstance(heat, Isothermal) else None)


def _builtin_dirichlet(b):
    """BoundaryConditionDirichlet of a boundary value function the device has
    built in (the reference's elixirs wrap their initial condition, e.g.
    BoundaryConditionDirichlet(initial_condition_convergence_test)) is that
    function itself, which the device loop can evaluate; any other stays a
    host-tabulated condition."""
    if isinstance(b, BoundaryConditionDirichlet) and getattr(b.boundary_value_function, "kind", None) in (
            BV_VORTEX, BV_CONSTANT, BV_RIEMANN4, BV_CONVERGENCE_TEST):
        return b.boundary_value_function
    return b


def boundary_conditions_kind(bcs, periodicity=None):
    """One boundary-value description for every boundary face from the
    reference's forms: a single function object, or the NamedTuple (x_neg,
    x_pos, y_neg, y_pos) as a dict or 4-tuple.  Per-direction conditions
    must agree (the same function, or the vertical slip wall on the x faces
    and the horizontal one on the y faces); boundary_condition_periodic may
    stand for the directions in which the mesh is periodic (``periodicity``)."""
    if bcs is None or not isinstance(bcs, (dict, tuple, list)):
        return _builtin_dirichlet(bcs)
    seq = [_builtin_dirichlet(bcs[k]) for k in ("x_neg", "x_pos", "y_neg", "y_pos")] if isinstance(bcs, dict) \
        else [_builtin_dirichlet(b) for b in bcs]
    if len(seq) != 4:
        raise ValueError("per-direction boundary conditions: (x_neg, x_pos, y_neg, y_pos)")
    per = tuple(periodicity) if periodicity is not None else (False, False)
    for k, b in enumerate(seq):
        if isinstance(b, _Periodic) and not per[k // 2]:
            raise ValueError("boundary_condition_periodic in a direction in which the mesh is not periodic")
    live = [(k, b) for k, b in enumerate(seq) if not isinstance(b, _Periodic)]
    if not live:
        return None
    if all(isinstance(b, _SlipWall) for _, b in live):
        if any(b.orientation != (1 if k < 2 else 2) for k, b in live):
            raise NotImplementedError("the slip wall of x faces is the vertical one, of y faces the horizontal one")
        return live[0][1]
    if all(b is live[0][1] for _, b in live):
        return live[0][1]
    raise NotImplementedError("the device path takes one boundary condition for every boundary face")


@dataclass
class CompressibleEulerEquations2D:
    gamma: float = 1.4
    nvariables: int = 4

    def prim2cons(self, rho, v1, v2, p):
        g = self.gamma
        return np.stack([rho, rho * v1, rho * v2, p / (g - 1) + 0.5 * rho * (v1 * v1 + v2 * v2)], axis=-1)

    def cons2prim(self, u):
        rho = u[..., 0]
        v1 = u[..., 1] / rho
        v2 = u[..., 2] / rho
        p = (self.gamma - 1) * (u[..., 3] - 0.5 * (u[..., 1] * v1 + u[..., 2] * v2))
        return rho, v1, v2, p


class GradientVariablesConservative:
    """Gradients of the conservative variables (TrixiLW's
    GradientVariablesConservative, src/equations/compressible_navier_stokes_2d.jl:14-57)."""


class GradientVariablesPrimitive:
    """Trixi's default gradient variables; not implemented on the device."""


class CompressibleNavierStokesDiffusion2D:
    """CompressibleNavierStokesDiffusion2D(equations; mu, Prandtl, gradient_variables)
    with constant dynamic viscosity μ and κ = γ/((γ-1) Pr) (Trixi's constructor;
    used through TrixiLW's compressible_navier_stokes_2d.jl)."""

    def __init__(self, equations, mu, Prandtl, gradient_variables=None):
        if not isinstance(equations, CompressibleEulerEquations2D):
            raise TypeError("the hyperbolic part must be CompressibleEulerEquations2D")
        gv = gradient_variables if gradient_variables is not None else GradientVariablesPrimitive()
        if not isinstance(gv, GradientVariablesConservative):
            raise NotImplementedError("the device path implements GradientVariablesConservative")
        if not (np.isfinite(mu) and mu >= 0.0 and np.isfinite(Prandtl) and Prandtl > 0.0):
            raise ValueError("mu must be >= 0 and Prandtl > 0")
        self.equations_hyperbolic = equations
        self.gamma = equations.gamma
        self.mu = float(mu)
        self.Pr = float(Prandtl)
        self.kappa = self.gamma / ((self.gamma - 1.0) * self.Pr)
        self.gradient_variables = gv
    nvariables: int = 1


@dataclass
class IsentropicVortex:
    """Shu/Yee isentropic vortex translating with the free stream.

    dv = eps/(2π) exp((1-r²)/2) (-y, x), T = 1 - (γ-1) eps²/(8γπ²) exp(1-r²),
    ρ = T^{1/(γ-1)}, p = ρT, background ρ = p = 1 and v∞ = M √γ (1,1)/√2.
    The exact solution at time t is the periodic translate by v∞ t; the
    images beyond the nearest one contribute below 1e-20 on [-10,10]².
    """

    strength: float = 5.0
    mach: float = 0.5
    center: tuple = (0.0, 0.0)
    period: tuple = (20.0, 20.0)
    gamma: float = 1.4

    kind = BV_VORTEX

    @property
    def params(self):
        return [self.strength, self.mach, self.center[0], self.center[1], self.period[0], self.period[1],
                self.gamma]

    def __call__(self, x, y, t, equations=None):
        g = self.gamma
        vinf = self.mach * np.sqrt(g) / np.sqrt(2.0)
        dx = x - (self.center[0] + vinf * t)
        dy = y - (self.center[1] + vinf * t)
        Lx, Ly = self.period
        dx = dx - Lx * np.floor(dx / Lx + 0.5)
        dy = dy - Ly * np.floor(dy / Ly + 0.5)
        r2 = dx * dx + dy * dy
        eps = self.strength
        du = eps / (2 * np.pi) * np.exp(0.5 * (1 - r2))
        T = 1 - (g - 1) * eps * eps / (8 * g * np.pi * np.pi) * np.exp(1 - r2)
        rho = T ** (1 / (g - 1))
        return CompressibleEulerEquations2D(g).prim2cons(rho, vinf - du * dy, vinf + du * dx, rho * T)


@dataclass
class TaylorGreenVortex:
    """Two-dimensional Taylor-Green vortex on the periodic square [x0, x0 + L]²
    (an initial condition, not a boundary value function): with k = 2π/L,
    v1 = U sin(k(x - x0)) cos(k(y - y0)), v2 = -U cos(k(x - x0)) sin(k(y - y0)),
    ρ = ρ0 and p = p0 + ρ0 U²/4 (cos 2k(x - x0) + cos 2k(y - y0)), the
    incompressible vortex's pressure on a background p0 = ρ0 U² / (γ M²) of
    Mach number M.  Viscosity makes it decay; its Reynolds number is
    ρ0 U (L/2) / μ."""

    U: float = 1.0
    rho0: float = 1.0
    mach: float = 0.1
    L: float = 2.0
    origin: tuple = (0.0, 0.0)
    gamma: float = 1.4

    def __call__(self, x, y, t=0.0, equations=None):
        k = 2.0 * np.pi / self.L
        X = k * (np.asarray(x, dtype=np.float64) - self.origin[0])
        Y = k * (np.asarray(y, dtype=np.float64) - self.origin[1])
        p0 = self.rho0 * self.U ** 2 / (self.gamma * self.mach ** 2)
        v1 = self.U * np.sin(X) * np.cos(Y)
        v2 = -self.U * np.cos(X) * np.sin(Y)
        p = p0 + 0.25 * self.rho0 * self.U ** 2 * (np.cos(2 * X) + np.cos(2 * Y))
        rho = self.rho0 * np.ones_like(X)
        return CompressibleEulerEquations2D(self.gamma).prim2cons(rho, v1, v2, p)


@dataclass
class ConstantState:
    rho: float = 1.0
    v1: float = 0.1
    v2: float = -0.2
    p: float = 1.0
    gamma: float = 1.4

    kind = BV_CONSTANT

    @property
    def params(self):
        return [self.rho, self.v1, self.v2, self.p, self.gamma]

    def __call__(self, x, y, t, equations=None):
        one = np.ones_like(np.asarray(x, dtype=np.float64))
        return CompressibleEulerEquations2D(self.gamma).prim2cons(self.rho * one, self.v1 * one, self.v2 * one,
                                                                  self.p * one)


class ConvergenceTest:
    """Trixi's initial_condition_convergence_test for CompressibleEulerEquations2D
    (the manufactured solution of the reference's elixir_euler_source_terms*.jl):
    ρ = ρv1 = ρv2 = φ and ρe = φ², φ = 2 + 0.1 sin(π(x + y - t)); the exact
    solution with source_terms_convergence_test.  Also a boundary value
    function (BoundaryConditionDirichlet of it)."""

    kind = BV_CONVERGENCE_TEST
    params = ()

    def __call__(self, x, y, t, equations=None):
        phi = 2.0 + 0.1 * np.sin(np.pi * (np.asarray(x, dtype=np.float64) + np.asarray(y, dtype=np.float64) - t))
        return np.stack([phi, phi, phi, phi * phi], axis=-1)


class SourceTermsConvergenceTest:
    """Trixi's source_terms_convergence_test for CompressibleEulerEquations2D:
    the residual ∂t u + ∇·f(u) of ConvergenceTest's state, which depends on
    (x, t) only.  The device evaluates it inside the LW and MDRK procedures;
    calling it here gives the same values on the host."""

    kind = SRC_CONVERGENCE_TEST

    def __call__(self, u, x, y, t, equations=None):
        gamma = equations.gamma if equations is not None else 1.4
        arg = np.pi * (np.asarray(x, dtype=np.float64) + np.asarray(y, dtype=np.float64) - t)
        rho = 2.0 + 0.1 * np.sin(arg)
        rho_x = np.pi * 0.1 * np.cos(arg)
        tmp = (2.0 * rho - 1.0) * (gamma - 1.0)
        return np.stack([rho_x, rho_x * (1.0 + tmp), rho_x * (1.0 + tmp), 2.0 * rho_x * (rho + tmp)], axis=-1)


initial_condition_convergence_test = ConvergenceTest()
source_terms_convergence_test = SourceTermsConvergenceTest()


class NavierStokesConvergenceTest:
    """initial_condition_navier_stokes_convergence_test of the reference's
    examples/tree_2d_dgsem/elixir_navierstokes_convergence.jl:31-45:
    ρ = 2 + ½ sin πx cos πy cos πt, v1 = v2 = sin πx log(y + 2)(1 - e^{-(y-1)/2}) cos πt,
    p = ρ² (zero velocity on the walls y = ±1)."""

    def __call__(self, x, y, t, equations=None):
        gamma = getattr(equations, "gamma", 1.4)
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        rho = 2.0 + 0.5 * np.sin(np.pi * x) * np.cos(np.pi * y) * np.cos(np.pi * t)
        v = np.sin(np.pi * x) * np.log(y + 2.0) * (1.0 - np.exp(-0.5 * (y - 1.0))) * np.cos(np.pi * t)
        p = rho * rho
        return np.stack([rho, rho * v, rho * v, p / (gamma - 1.0) + 0.5 * rho * (v * v + v * v)], axis=-1)


class SourceTermsNavierStokesConvergenceTest:
    """source_terms_navier_stokes_convergence_test of the same example: the
    residual of NavierStokesConvergenceTest's state in the compressible
    Navier-Stokes equations.  The reference's function reads the example's
    mu() = 0.01 and prandtl_number() = 0.72, the values its equations take; the
    device evaluates it with the context's μ and Pr, so a semidiscretization
    accepts it only when those match ``mu`` and ``prandtl`` here.  It runs in the
    Navier-Stokes MDRK stages on a TreeMesh."""

    kind = SRC_NS_CONVERGENCE_TEST

    def __init__(self, mu=0.01, prandtl=0.72):
        self.mu, self.prandtl = float(mu), float(prandtl)


initial_condition_navier_stokes_convergence_test = NavierStokesConvergenceTest()
source_terms_navier_stokes_convergence_test = SourceTermsNavierStokesConvergenceTest()


@dataclass
class FourQuadrantRiemann:
    """Piecewise-constant (ρ, v1, v2, p) per quadrant around (x0, y0).

    Default states: configuration 3 of Lax & Liu (SIAM J. Sci. Comput. 19,
    1998), ρ ∈ [0.138, 1.5] and p ∈ [0.029, 1.5]; quadrant order NE, NW, SW, SE.
    """

    x0: float = 0.5
    y0: float = 0.5
    states: tuple = ((1.5, 0.0, 0.0, 1.5),
                     (0.5323, 1.206, 0.0, 0.3),
                     (0.138, 1.206, 1.206, 0.029),
                     (0.5323, 0.0, 1.206, 0.3))
    gamma: float = 1.4

    kind = BV_RIEMANN4

    @property
    def params(self):
        out = [self.x0, self.y0, self.gamma]
        for s in self.states:
            out.extend(s)
        return out

    def __call__(self, x, y, t, equations=None):
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        q = np.where(x >= self.x0, np.where(y >= self.y0, 0, 3), np.where(y >= self.y0, 1, 2))
        st = np.asarray(self.states, dtype=np.float64)
        rho, v1, v2, p = (st[q, k] for k in range(4))
        return CompressibleEulerEquations2D(self.gamma).prim2cons(rho, v1, v2, p)
