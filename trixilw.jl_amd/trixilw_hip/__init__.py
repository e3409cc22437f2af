"""trixilw_hip -- MI355X (gfx950) implementation of TrixiLW's LW-FR / MDRK right-hand side.

Host-side mirror of the reference's API (TreeMesh, DGSEM, LW/MDRK,
SemidiscretizationHyperbolic, rhs!/rhs_mdrk1!/rhs_mdrk2!, solve_lwfr) over the
C ABI of libtrixilw_hip.so.  Importing this package does not load the
library; the first call that needs the GPU does, and fails if it is missing.
"""
from .adjoint import adjoint_step, adjoint_step_mdrk, gaussian_random_field, gradient_lw, kinetic_energy, objective_lw
from .basis import LobattoLegendreBasis
from .batch import SemidiscretizationBatch, SemidiscretizationEnsemble, draw_vortex_ensemble
from .equations import (Adiabatic, BoundaryConditionDirichlet, BoundaryConditionNavierStokesWall, CompressibleEulerEquations2D,
                        CompressibleNavierStokesDiffusion2D, ConstantState, Isothermal, NoSlip,
                        FourQuadrantRiemann, GradientVariablesConservative, GradientVariablesPrimitive,
                        ConvergenceTest, IsentropicVortex, SourceTermsConvergenceTest, TaylorGreenVortex,
                        boundary_condition_outflow, boundary_condition_slip_wall_horizontal,
                        initial_condition_convergence_test, slip_wall_approximate,
                        source_terms_convergence_test, boundary_condition_periodic,
                        NavierStokesConvergenceTest, SourceTermsNavierStokesConvergenceTest,
                        initial_condition_navier_stokes_convergence_test,
                        source_terms_navier_stokes_convergence_test,
                        boundary_condition_slip_wall_vertical)
from .integrator import (LWIntegrator, LWSolution, compute_dt, isapprox, isfinished, perform_step, set_dt,
                         set_t_and_iter, solve_lwfr)
from .limiters import PositivityPreservingLimiterZhangShu, density, pressure
from .mesh import P4estMesh, StructuredMesh, TreeMesh, coordinates2mapping
from .semidiscretization import (DGSEM, LW, MDRK, Adaptive, CFLBased, FirstOrderReconstruction, IndicatorHennemannGassner,
                                 LWUpdate, MUSCLHancockReconstruction, SemidiscretizationHyperbolic,
                                 SemidiscretizationHyperbolicParabolic, VolumeIntegralFR,
                                 VolumeIntegralFRShockCapturing, density_pressure, flux_lax_friedrichs,
                                 get_time_discretization, lw_fourier_cfl, max_dt_noalloc, rhs, rhs_mdrk1, rhs_mdrk2, semidiscretize,
                                 trixi2lw)

__all__ = [
    "LobattoLegendreBasis", "CompressibleEulerEquations2D", "ConstantState", "FourQuadrantRiemann",
    "boundary_condition_slip_wall_vertical", "boundary_condition_slip_wall_horizontal",
    "boundary_condition_outflow", "slip_wall_approximate", "NoSlip", "Adiabatic", "Isothermal",
    "BoundaryConditionNavierStokesWall", "BoundaryConditionDirichlet",
    "IsentropicVortex", "TaylorGreenVortex", "LWIntegrator", "LWSolution", "compute_dt", "isapprox",
    "isfinished", "perform_step", "set_dt", "set_t_and_iter", "solve_lwfr", "TreeMesh",
    "StructuredMesh", "P4estMesh", "coordinates2mapping", "DGSEM", "LW", "MDRK",
    "CFLBased", "Adaptive", "LWUpdate", "SemidiscretizationHyperbolic", "VolumeIntegralFR", "flux_lax_friedrichs",
    "get_time_discretization", "lw_fourier_cfl", "max_dt_noalloc", "rhs", "rhs_mdrk1", "rhs_mdrk2", "semidiscretize", "trixi2lw",
    "FirstOrderReconstruction", "IndicatorHennemannGassner", "MUSCLHancockReconstruction",
    "VolumeIntegralFRShockCapturing", "density_pressure", "PositivityPreservingLimiterZhangShu", "density",
    "pressure", "adjoint_step", "adjoint_step_mdrk", "gradient_lw", "kinetic_energy", "gaussian_random_field", "objective_lw",
    "SemidiscretizationBatch", "SemidiscretizationEnsemble", "draw_vortex_ensemble", "SemidiscretizationHyperbolicParabolic", "CompressibleNavierStokesDiffusion2D",
    "GradientVariablesConservative", "GradientVariablesPrimitive", "ConvergenceTest", "SourceTermsConvergenceTest",
    "initial_condition_convergence_test", "source_terms_convergence_test", "boundary_condition_periodic",
    "NavierStokesConvergenceTest", "SourceTermsNavierStokesConvergenceTest",
    "initial_condition_navier_stokes_convergence_test", "source_terms_navier_stokes_convergence_test",
]
