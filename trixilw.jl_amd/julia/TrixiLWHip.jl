# TrixiLWHip.jl -- Julia binding of libtrixilw_hip.so (the C ABI of include/trixilw_hip.h).
#
# What a TrixiLW.jl maintainer adds to call the MI355X right-hand side and time loop through ccall.
# Julia is not part of this image, so this file is not executed here; tests/test_julia_binding.py
# checks it against the header: the struct layouts (field order, types, offsets, sizes) and one
# ccall per exported entry point with the header's argument types.
#
# Arrays passed as *_dev are device arrays (ROCArray); *_host are ordinary Julia arrays.  The state
# layout is Trixi's [v, i, j, e] column-major array, as wrap_array gives it.
module TrixiLWHip

const lib = "libtrixilw_hip.so"

# error codes; TLW_EADMISSIBLE maps to the DomainError of src/solvers/adaptive.jl:6-21
const TLW_OK = Cint(0)
const TLW_EINVAL = Cint(1)
const TLW_EHIP = Cint(2)
const TLW_EADMISSIBLE = Cint(3)
const TLW_ENOTSUP = Cint(4)
# time discretizations (src/basic_types.jl:4-6)
const TLW_LW = Int32(0)
const TLW_MDRK = Int32(1)
# Rusanov argument order: TreeMesh (λ from Ũ, jump on uⁿ) or Structured/P4est (λ from uⁿ, jump on Ũ)
const TLW_FLUX_TREE = Int32(0)
const TLW_FLUX_STRUCTURED = Int32(1)
# boundary value functions of the time-averaged Dirichlet condition (src/equations/equations.jl:6-37)
const TLW_BV_NONE = Int32(0)
const TLW_BV_VORTEX = Int32(1)
const TLW_BV_CONSTANT = Int32(2)
const TLW_BV_RIEMANN4 = Int32(3)
const TLW_BV_SLIP_WALL = Int32(4)
const TLW_BV_TABLE = Int32(5)      # BoundaryConditionDirichlet(f) for any f: see set_boundary_data!
const TLW_BV_CONVERGENCE_TEST = Int32(6)  # BoundaryConditionDirichlet(initial_condition_convergence_test)
const TLW_BV_OUTFLOW = Int32(7)               # boundary_condition_outflow (curvilinear StructuredMesh)
const TLW_BV_SLIP_WALL_APPROXIMATE = Int32(8) # slip_wall_approximate (curvilinear StructuredMesh)
const TLW_SRC_NONE = Int32(0)
const TLW_SRC_CONVERGENCE_TEST = Int32(1)  # source_terms = source_terms_convergence_test
const TLW_SRC_NAVIER_STOKES_CONVERGENCE_TEST = Int32(2)  # source_terms_navier_stokes_convergence_test (MDRK)
const TLW_AUX_U = Int32(0)
const TLW_AUX_F = Int32(1)
const TLW_AUX_ALPHA = Int32(2)
const TLW_RECON_FIRST_ORDER = Int32(0)
const TLW_RECON_MUSCL_HANCOCK = Int32(1)
const TLW_FLAG_KEEP_AUX = Int32(1)
const TLW_PAR_NAVIER_STOKES = Int32(1)
const TLW_GRAD_CONSERVATIVE = Int32(0)
const TLW_GRAD_PRIMITIVE = Int32(1)

# ---------------------------------------------------------------------------------------------
# Descriptions (isbits structs with the C layout of include/trixilw_hip.h)

# TreeMesh element data (cache.elements) and face neighbours from cache.interfaces/boundaries
struct MeshDesc
    n_elements::Int32
    neighbors_host::Ptr{Int32}            # [4, E]: -x, +x, -y, +y neighbour, -1 = boundary (0-based)
    inverse_jacobian_host::Ptr{Float64}   # cache.elements.inverse_jacobian
    element_center_host::Ptr{Float64}     # [2, E]
    element_length_host::Ptr{Float64}     # [E]
end

# TreeMesh with refined cells: Trixi's mortar container (cache.mortars), 0-based
# {lower, upper, large, orientation, large_side} per mortar; mortar faces carry neighbour -2
const TLW_NEIGHBOR_MORTAR = Int32(-2)
struct MortarDesc
    n_mortars::Int32
    mortars_host::Ptr{Int32}              # [5, M]
end

# LobattoLegendreBasis fields, column-major as Julia stores them
struct BasisDesc
    polydeg::Int32
    nodes_host::Ptr{Float64}
    weights_host::Ptr{Float64}
    derivative_matrix_host::Ptr{Float64}
    derivative_dhat_host::Ptr{Float64}
end

struct SolverDesc
    time_discretization::Int32            # TLW_LW or TLW_MDRK
    gamma::Float64                        # CompressibleEulerEquations2D
    surface_flux_order::Int32             # TLW_FLUX_TREE or TLW_FLUX_STRUCTURED
    boundary_kind::Int32                  # TLW_BV_*
    boundary_params::NTuple{24,Float64}
    flags::Int32                          # TLW_FLAG_*
    # VolumeIntegralFRShockCapturing(IndicatorHennemannGassner(...; variable=density_pressure), ...)
    shock_capturing::Int32                # 0 for VolumeIntegralFR
    alpha_max::Float64
    alpha_min::Float64
    alpha_smooth::Int32
    inverse_vandermonde_host::Ptr{Float64}  # basis.inverse_vandermonde_legendre, column-major n×n
    reconstruction::Int32                 # TLW_RECON_FIRST_ORDER or TLW_RECON_MUSCL_HANCOCK
    source_terms::Int32                   # TLW_SRC_*
end

# P4estMesh: the metric terms and Trixi's P4est interface container, node_indices reduced to
# {primary element, primary face, secondary element, secondary face, flipped} (0-based)
struct P4estDesc
    contravariant_vectors_host::Ptr{Float64}
    inverse_jacobian_host::Ptr{Float64}
    node_coordinates_host::Ptr{Float64}
    n_interfaces::Int32
    interfaces_host::Ptr{Int32}               # [5, I]
end

# CompressibleNavierStokesDiffusion2D(equations; mu, Prandtl, gradient_variables) and its walls
struct ParabolicDesc
    equations::Int32                      # TLW_PAR_NAVIER_STOKES
    mu::Float64
    prandtl::Float64
    gradient_variables::Int32             # TLW_GRAD_CONSERVATIVE
    walls::Int32                          # 1: NoSlip walls on every boundary face
    wall_velocity::NTuple{8,Float64}      # (v1, v2) of x_neg, x_pos, y_neg, y_pos
    wall_isothermal::NTuple{4,Int32}      # per direction: 0 Adiabatic, 1 Isothermal
    wall_temperature::NTuple{4,Float64}
end

# StructuredMesh metrics (cache.elements of a StructuredMesh), passed as Julia holds them
struct CurvedDesc
    contravariant_vectors_host::Ptr{Float64}  # [2, 2, n, n, E]
    inverse_jacobian_host::Ptr{Float64}       # [n, n, E], all > 0
    node_coordinates_host::Ptr{Float64}       # [2, n, n, E] (boundary data)
end

function SolverDesc(; time_discretization = TLW_LW, gamma = 1.4, surface_flux_order = TLW_FLUX_TREE,
                    boundary_kind = TLW_BV_NONE, boundary_params = ntuple(_ -> 0.0, 24), flags = Int32(0),
                    shock_capturing = Int32(0), alpha_max = 0.5, alpha_min = 0.001, alpha_smooth = Int32(1),
                    inverse_vandermonde = Ptr{Float64}(C_NULL), reconstruction = TLW_RECON_FIRST_ORDER,
                    source_terms = TLW_SRC_NONE)
    return SolverDesc(Int32(time_discretization), Float64(gamma), Int32(surface_flux_order), Int32(boundary_kind),
                      NTuple{24,Float64}(boundary_params), Int32(flags), Int32(shock_capturing), Float64(alpha_max),
                      Float64(alpha_min), Int32(alpha_smooth), inverse_vandermonde, Int32(reconstruction),
                      Int32(source_terms))
end

function ParabolicDesc(eqp; walls = nothing, wall_temperature = nothing)
    wv = walls === nothing ? ntuple(_ -> 0.0, 8) : NTuple{8,Float64}(Iterators.flatten(walls))
    iso = ntuple(k -> Int32(wall_temperature === nothing || wall_temperature[k] === nothing ? 0 : 1), 4)
    wt = ntuple(k -> iso[k] == 1 ? Float64(wall_temperature[k]) : 0.0, 4)
    return ParabolicDesc(TLW_PAR_NAVIER_STOKES, eqp.mu, eqp.Pr, TLW_GRAD_CONSERVATIVE,
                         Int32(walls === nothing ? 0 : 1), wv, iso, wt)
end

# ---------------------------------------------------------------------------------------------
# Errors

last_error(ctx) = unsafe_string(ccall((:tlw_last_error, lib), Cstring, (Ptr{Cvoid},), ctx))

function check(ctx, rc)
    rc == TLW_OK && return nothing
    rc == TLW_EADMISSIBLE && throw(DomainError(last_error(ctx)))
    error("libtrixilw_hip error $rc: " * last_error(ctx))
end

# ---------------------------------------------------------------------------------------------
# Contexts: SemidiscretizationHyperbolic(/Parabolic) + create_cache_serial
# (src/semidiscretization/semidiscretization_hyperbolic.jl:58-84, containers.jl:7-64)

function create(mesh::MeshDesc, basis::BasisDesc, solver::SolverDesc)
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    check(C_NULL, ccall((:tlw_create, lib), Cint,
                        (Ref{MeshDesc}, Ref{BasisDesc}, Ref{SolverDesc}, Ref{Ptr{Cvoid}}), mesh, basis, solver, ctx))
    return ctx[]
end

function create_batch(meshes::Vector{MeshDesc}, basis::BasisDesc, solver::SolverDesc)
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    check(C_NULL, ccall((:tlw_create_batch, lib), Cint,
                        (Int32, Ptr{MeshDesc}, Ref{BasisDesc}, Ref{SolverDesc}, Ref{Ptr{Cvoid}}),
                        length(meshes), meshes, basis, solver, ctx))
    return ctx[]
end

function create_parabolic(mesh::MeshDesc, basis::BasisDesc, solver::SolverDesc, par::ParabolicDesc)
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    check(C_NULL, ccall((:tlw_create_parabolic, lib), Cint,
                        (Ref{MeshDesc}, Ref{BasisDesc}, Ref{SolverDesc}, Ref{ParabolicDesc}, Ref{Ptr{Cvoid}}),
                        mesh, basis, solver, par, ctx))
    return ctx[]
end

function create_curved(mesh::MeshDesc, basis::BasisDesc, solver::SolverDesc, elements)
    cd = CurvedDesc(pointer(elements.contravariant_vectors), pointer(elements.inverse_jacobian),
                    pointer(elements.node_coordinates))
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve elements check(C_NULL, ccall((:tlw_create_curved, lib), Cint,
        (Ref{MeshDesc}, Ref{BasisDesc}, Ref{SolverDesc}, Ref{CurvedDesc}, Ref{Ptr{Cvoid}}),
        mesh, basis, solver, cd, ctx))
    return ctx[]
end

function create_curved_parabolic(mesh::MeshDesc, basis::BasisDesc, solver::SolverDesc, elements,
                                 par::ParabolicDesc)
    cd = CurvedDesc(pointer(elements.contravariant_vectors), pointer(elements.inverse_jacobian),
                    pointer(elements.node_coordinates))
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve elements check(C_NULL, ccall((:tlw_create_curved_parabolic, lib), Cint,
        (Ref{MeshDesc}, Ref{BasisDesc}, Ref{SolverDesc}, Ref{CurvedDesc}, Ref{ParabolicDesc}, Ref{Ptr{Cvoid}}),
        mesh, basis, solver, cd, par, ctx))
    return ctx[]
end

# AMR: after DiffEqBase.resize! (src/callbacks/amr.jl:2-20) re-upload the host-built mesh
rebuild_mesh!(ctx, mesh::MeshDesc) =
    check(ctx, ccall((:tlw_rebuild_mesh, lib), Cint, (Ptr{Cvoid}, Ref{MeshDesc}), ctx, mesh))

# P4estMesh (conforming): the curvilinear kernels with the P4est surface terms
function create_p4est(mesh::MeshDesc, p4est::P4estDesc, basis::BasisDesc, solver::SolverDesc)
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    check(C_NULL, ccall((:tlw_create_p4est, lib), Cint,
                        (Ref{MeshDesc}, Ref{P4estDesc}, Ref{BasisDesc}, Ref{SolverDesc}, Ref{Ptr{Cvoid}}),
                        mesh, p4est, basis, solver, ctx))
    return ctx[]
end

# TreeMesh with refined cells (mortars), and its AMR re-upload after resize!
function create_nonconforming(mesh::MeshDesc, mortars::MortarDesc, basis::BasisDesc, solver::SolverDesc)
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    check(C_NULL, ccall((:tlw_create_nonconforming, lib), Cint,
                        (Ref{MeshDesc}, Ref{MortarDesc}, Ref{BasisDesc}, Ref{SolverDesc}, Ref{Ptr{Cvoid}}),
                        mesh, mortars, basis, solver, ctx))
    return ctx[]
end
rebuild_mesh_nonconforming!(ctx, mesh::MeshDesc, mortars::MortarDesc) =
    check(ctx, ccall((:tlw_rebuild_mesh_nonconforming, lib), Cint, (Ptr{Cvoid}, Ref{MeshDesc}, Ref{MortarDesc}),
                     ctx, mesh, mortars))
# Trixi's refine!/coarsen! of the solution on the device: plan [5, E_new] (see the header)
transfer_solution!(ctx, u_old_dev::Ptr{Float64}, n_old, u_new_dev::Ptr{Float64}, n_new, plan::Matrix{Int32}) =
    check(ctx, ccall((:tlw_transfer_solution, lib), Cint,
                     (Ptr{Cvoid}, Ptr{Float64}, Int32, Ptr{Float64}, Int32, Ptr{Int32}),
                     ctx, u_old_dev, n_old, u_new_dev, n_new, plan))
destroy!(ctx) = ccall((:tlw_destroy, lib), Cvoid, (Ptr{Cvoid},), ctx)
set_stream!(ctx, stream::Ptr{Cvoid}) =
    check(ctx, ccall((:tlw_set_stream, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx, stream))
state_len(ctx) = Int(ccall((:tlw_state_len, lib), Csize_t, (Ptr{Cvoid},), ctx))
n_instances(ctx) = (n = Ref{Int32}(0);
    check(ctx, ccall((:tlw_n_instances, lib), Cint, (Ptr{Cvoid}, Ref{Int32}), ctx, n)); Int(n[]))

# The StructuredMesh's face neighbours as the library's table (faces -x, +x, -y, +y; element
# e = cx + nx*cy with Trixi's linear indices, 0-based for C), periodic in both directions
function structured_neighbors(nx, ny)
    nbr = Vector{Int32}(undef, 4 * nx * ny)
    for cy in 0:ny-1, cx in 0:nx-1
        e = cx + nx * cy
        nbr[4e+1] = mod(cx - 1, nx) + nx * cy
        nbr[4e+2] = mod(cx + 1, nx) + nx * cy
        nbr[4e+3] = cx + nx * mod(cy - 1, ny)
        nbr[4e+4] = cx + nx * mod(cy + 1, ny)
    end
    return nbr
end

# An ensemble of uneven instances: a polydeg-3 and a polydeg-4 batch context advanced by the same
# launches (one volume, surface and finalize launch per step for all instances)
function create_ensemble(ctx_a, ctx_b)
    ens = Ref{Ptr{Cvoid}}(C_NULL)
    check(C_NULL, ccall((:tlw_create_ensemble, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}), ctx_a, ctx_b, ens))
    return ens[]
end
ensemble_error(ens) = unsafe_string(ccall((:tlw_ensemble_last_error, lib), Cstring, (Ptr{Cvoid},), ens))
ensemble_check(ens, rc) = rc == TLW_OK ? nothing : error("libtrixilw_hip error $rc: " * ensemble_error(ens))
ensemble_set_stream!(ens, stream::Ptr{Cvoid}) =
    ensemble_check(ens, ccall((:tlw_ensemble_set_stream, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ens, stream))
ensemble_enable_graphs!(ens, on::Bool) =
    ensemble_check(ens, ccall((:tlw_ensemble_enable_graphs, lib), Cint, (Ptr{Cvoid}, Int32), ens, on))
ensemble_solve_init!(ens, t0, tf, cfl3, cfl4) = ensemble_check(ens, ccall((:tlw_ensemble_solve_init, lib), Cint,
    (Ptr{Cvoid}, Float64, Float64, Float64, Float64), ens, t0, tf, cfl3, cfl4))
ensemble_advance!(ens, max_steps) = (fin = Ref{Int32}(0);
    ensemble_check(ens, ccall((:tlw_ensemble_advance, lib), Cint, (Ptr{Cvoid}, Int32, Ref{Int32}), ens, max_steps, fin));
    fin[] != 0)
ensemble_destroy!(ens) = ccall((:tlw_ensemble_destroy, lib), Cvoid, (Ptr{Cvoid},), ens)

# ---------------------------------------------------------------------------------------------
# rhs!, rhs_mdrk1!, rhs_mdrk2! (semidiscretization_hyperbolic.jl:4-56) on device arrays;
# dt is semi.cache.dt[1] as in the reference

rhs_gpu!(du_ode, u_ode, semi, t, ctx) = check(ctx, ccall((:tlw_rhs_lw, lib), Cint,
    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Float64), ctx, u_ode, du_ode, t, semi.cache.dt[1]))
rhs_mdrk1_gpu!(du_ode, u_ode, semi, t, ctx) = check(ctx, ccall((:tlw_rhs_mdrk1, lib), Cint,
    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Float64), ctx, u_ode, du_ode, t, semi.cache.dt[1]))
rhs_mdrk2_gpu!(du_ode, u_ode, semi, t, ctx) = check(ctx, ccall((:tlw_rhs_mdrk2, lib), Cint,
    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Float64), ctx, u_ode, du_ode, t, semi.cache.dt[1]))
# mdrk.jl:17 writes us = uprev + dt*du into element_cache.mdrk_cache._us; on the device:
us_ptr(ctx) = (p = Ref{Ptr{Float64}}(C_NULL);
    check(ctx, ccall((:tlw_us_ptr, lib), Cint, (Ptr{Cvoid}, Ref{Ptr{Float64}}), ctx, p)); p[])
set_us!(ctx, us_dev) = check(ctx, ccall((:tlw_set_us, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx, us_dev))

# Trixi max_dt inside compute_dt (src/solvers/lwfr.jl:176-210)
max_dt_gpu(u_ode, ctx) = (r = Ref{Float64}(0);
    check(ctx, ccall((:tlw_max_dt, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ref{Float64}), ctx, u_ode, r)); r[])

# Allocation-free max_dt for code that cannot allocate (graph capture, serving loops): no context, the
# result lands in device memory (NaN for a non-admissible state); a fixed 2 MiB device array per GPU
# holds the partials, so n_elements <= 1_310_720 at polydeg 4 (2_097_152 at 3); calls overlapping on one
# GPU race on it.
max_dt_noalloc!(out_dev, u_dev, invjac_dev, nelements, polydeg, gamma; stream = C_NULL) =
    check(C_NULL, ccall((:tlw_max_dt_noalloc, lib), Cint,
        (Ptr{Float64}, Ptr{Float64}, Int32, Int32, Float64, Ptr{Float64}, Ptr{Cvoid}),
        u_dev, invjac_dev, nelements, polydeg, gamma, out_dev, stream))

# ---------------------------------------------------------------------------------------------
# The solve_lwfr loop on the device (lwfr.jl:212-263, mdrk.jl:1-29)

set_state!(ctx, u_host::Vector{Float64}) = check(ctx, ccall((:tlw_set_state, lib), Cint,
    (Ptr{Cvoid}, Ptr{Float64}, Csize_t), ctx, u_host, length(u_host)))
get_state!(u_host::Vector{Float64}, ctx) = check(ctx, ccall((:tlw_get_state, lib), Cint,
    (Ptr{Cvoid}, Ptr{Float64}, Csize_t), ctx, u_host, length(u_host)))
state_ptr(ctx) = (p = Ref{Ptr{Float64}}(C_NULL);
    check(ctx, ccall((:tlw_state_ptr, lib), Cint, (Ptr{Cvoid}, Ref{Ptr{Float64}}), ctx, p)); p[])
set_state_dev!(ctx, u_dev) = check(ctx, ccall((:tlw_set_state_dev, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx, u_dev))
get_state_dev!(u_dev, ctx) = check(ctx, ccall((:tlw_get_state_dev, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx, u_dev))
solve_init!(ctx, t0, tf, cfl) = check(ctx, ccall((:tlw_solve_init, lib), Cint,
    (Ptr{Cvoid}, Float64, Float64, Float64), ctx, t0, tf, cfl))
function advance!(ctx, max_steps)
    t = Ref{Float64}(0); dt = Ref{Float64}(0); it = Ref{Int64}(0); fin = Ref{Int32}(0)
    check(ctx, ccall((:tlw_advance, lib), Cint,
                     (Ptr{Cvoid}, Int32, Ref{Float64}, Ref{Float64}, Ref{Int64}, Ref{Int32}),
                     ctx, max_steps, t, dt, it, fin))
    return (t = t[], dt = dt[], iter = it[], finished = fin[] != 0)
end
function batch_status(ctx, n)
    t = zeros(n); dt = zeros(n); it = zeros(Int64, n); fin = zeros(Int32, n)
    check(ctx, ccall((:tlw_batch_status, lib), Cint,
                     (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Int64}, Ptr{Int32}), ctx, n, t, dt, it, fin))
    return t, dt, it, fin
end
enable_graphs!(ctx, on::Bool) = check(ctx, ccall((:tlw_enable_graphs, lib), Cint, (Ptr{Cvoid}, Int32), ctx, on))

# solve_lwfr with CFLBased stepping and no callbacks
function solve_gpu!(u_ode::Vector{Float64}, ctx, t0, tf, cfl; chunk = 64)
    set_state!(ctx, u_ode)
    solve_init!(ctx, t0, tf, cfl)
    r = advance!(ctx, chunk)
    while !r.finished
        r = advance!(ctx, chunk)
    end
    get_state!(u_ode, ctx)
    return r
end

# solve_lwfr with time_step_computation = Adaptive() (src/solvers/adaptive.jl:95-173)
function solve_adaptive_gpu!(u_ode::Vector{Float64}, ctx, t0, tf, dt_initial, tolerances; chunk = 64)
    set_state!(ctx, u_ode)
    check(ctx, ccall((:tlw_solve_init_adaptive, lib), Cint, (Ptr{Cvoid}, Float64, Float64, Float64, Float64, Float64),
                     ctx, t0, tf, dt_initial, tolerances.abstol, tolerances.reltol))
    r = advance!(ctx, chunk)
    while !r.finished
        r = advance!(ctx, chunk)
    end
    nrej = Ref{Int64}(0)
    check(ctx, ccall((:tlw_rejected_steps, lib), Cint, (Ptr{Cvoid}, Ref{Int64}), ctx, nrej))
    get_state!(u_ode, ctx)
    return r, nrej[]
end

# ---------------------------------------------------------------------------------------------
# BoundaryConditionDirichlet(boundary_value_function) for any function (equations.jl:6-37):
# evaluate it at the boundary face nodes and upload before each rhs (or once, if it does not
# depend on t: the device loop then accepts it)

function boundary_nodes(ctx)
    n = Ref{Int64}(0)
    check(ctx, ccall((:tlw_boundary_nodes, lib), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Float64}, Csize_t),
                     ctx, n, C_NULL, 0))
    xy = zeros(2, n[])
    check(ctx, ccall((:tlw_boundary_nodes, lib), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Float64}, Csize_t),
                     ctx, n, xy, length(xy)))
    return xy
end
set_boundary_data!(ctx, values::Array{Float64}, time_levels) = check(ctx, ccall((:tlw_set_boundary_data, lib),
    Cint, (Ptr{Cvoid}, Ptr{Float64}, Csize_t, Int32), ctx, values, length(values), time_levels))
# values for the next rhs: u[v, q, node], q = 1..n at t + ½ s dt (ξ_q + 1), q = n + 1 at t
function upload_boundary!(ctx, xy, f, equations, t, dt, nodes; scaling = 1.0)
    n = length(nodes)
    vals = zeros(4, n + 1, size(xy, 2))
    for b in axes(xy, 2)
        x = (xy[1, b], xy[2, b])
        for q in 1:n
            vals[:, q, b] .= f(x, t + 0.5 * scaling * dt * (nodes[q] + 1), equations)
        end
        vals[:, n+1, b] .= f(x, t, equations)
    end
    set_boundary_data!(ctx, vals, n + 1)
end

# ---------------------------------------------------------------------------------------------
# Checkpointed gradients: per-step dt record / prescription of the device loop and the adjoint step

record_dts!(ctx, capacity) = check(ctx, ccall((:tlw_record_dts, lib), Cint, (Ptr{Cvoid}, Int64), ctx, capacity))
get_dts(ctx, n) = (d = zeros(n);
    check(ctx, ccall((:tlw_get_dts, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64), ctx, d, n)); d)
prescribe_dts!(ctx, dts::Vector{Float64}) = check(ctx, ccall((:tlw_prescribe_dts, lib), Cint,
    (Ptr{Cvoid}, Ptr{Float64}, Int64), ctx, dts, length(dts)))
adjoint_step!(lam, u, lam_next, ctx, t, dt) = check(ctx, ccall((:tlw_adjoint_lw, lib), Cint,
    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64), ctx, u, lam_next, lam, t, dt))
# MDRK: us = u + dt rhs_mdrk1!(u) is the step's stage state (recomputed by the caller)
adjoint_step_mdrk!(lam, u, us, lam_next, ctx, t, dt) = check(ctx, ccall((:tlw_adjoint_mdrk, lib), Cint,
    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64),
    ctx, u, us, lam_next, lam, t, dt))

# Reverse-mode rule of a fixed-dt LW solve u0 -> u_K with checkpoints every k steps (loaded
# into TrixiLW next to ChainRulesCore):
#
#   function ChainRulesCore.rrule(::typeof(solve_fixed_dt), u0, ctx, dts; k = 20)
#       prescribe_dts!(ctx, dts); set_state_dev!(ctx, u0); solve_init!(ctx, 0.0, Inf, 1.0)
#       checkpoints = Dict{Int,typeof(u0)}()
#       for s0 in 1:k:length(dts)
#           checkpoints[s0] = similar(u0); get_state_dev!(checkpoints[s0], ctx)
#           advance!(ctx, min(k, length(dts) - s0 + 1))
#       end
#       u = similar(u0); get_state_dev!(u, ctx); prescribe_dts!(ctx, Float64[])
#       ts = cumsum([0.0; dts[1:end-1]])
#       function pullback(ū)
#           lam = copy(ū); lam_prev = similar(lam)
#           for s0 in sort(collect(keys(checkpoints)); rev = true)
#               s1 = min(s0 + k - 1, length(dts)); states = [checkpoints[s0]]
#               for s in s0:s1-1; push!(states, lw_step(states[end], ctx, ts[s], dts[s])); end
#               for s in s1:-1:s0
#                   adjoint_step!(lam_prev, states[s - s0 + 1], lam, ctx, ts[s], dts[s])
#                   lam, lam_prev = lam_prev, lam
#               end
#           end
#           return NoTangent(), lam, NoTangent(), NoTangent()
#       end
#       return u, pullback
#   end

# ---------------------------------------------------------------------------------------------
# Stage limiter, adaptive statistics, tiles, measurement and the element cache

# PositivityPreservingLimiterZhangShu as a stage limiter on device arrays
# (src/callbacks/callbacks_stage/positivity_zhang_shu.jl:10-41)
struct GPUPositivityLimiter
    ctx::Ptr{Cvoid}
    thresholds::NTuple{2,Float64}    # (density, pressure)
end
set_positivity_limiter!(ctx, enable::Bool, thresholds) = check(ctx, ccall((:tlw_set_positivity_limiter, lib),
    Cint, (Ptr{Cvoid}, Int32, Float64, Float64), ctx, enable, thresholds[1], thresholds[2]))
function (lim::GPUPositivityLimiter)(u_ode, integrator, semi, t)
    set_positivity_limiter!(lim.ctx, true, lim.thresholds)
    check(lim.ctx, ccall((:tlw_apply_positivity_limiter, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}), lim.ctx, u_ode))
end

rejected_steps(ctx) = (r = Ref{Int64}(0);
    check(ctx, ccall((:tlw_rejected_steps, lib), Cint, (Ptr{Cvoid}, Ref{Int64}), ctx, r)); r[])

set_elements_per_workgroup!(ctx, eb) = check(ctx, ccall((:tlw_set_elements_per_workgroup, lib), Cint,
    (Ptr{Cvoid}, Int32), ctx, eb))
elements_per_workgroup(ctx) = (e = Ref{Int32}(0);
    check(ctx, ccall((:tlw_get_elements_per_workgroup, lib), Cint, (Ptr{Cvoid}, Ref{Int32}), ctx, e)); e[])
# pick the volume tile once per mesh, after set_state!
autotune!(ctx; reps = 3) = (c = Ref{Int32}(0); ms = zeros(3);
    check(ctx, ccall((:tlw_autotune, lib), Cint, (Ptr{Cvoid}, Int32, Ref{Int32}, Ptr{Float64}), ctx, reps, c, ms));
    (c[], ms))

enable_kernel_timing!(ctx, on::Bool) = check(ctx, ccall((:tlw_enable_kernel_timing, lib), Cint,
    (Ptr{Cvoid}, Int32), ctx, on))
function kernel_times(ctx)
    vms = Ref{Float64}(0); vn = Ref{Int64}(0); sms = Ref{Float64}(0); sn = Ref{Int64}(0)
    check(ctx, ccall((:tlw_kernel_times, lib), Cint,
                     (Ptr{Cvoid}, Ref{Float64}, Ref{Int64}, Ref{Float64}, Ref{Int64}), ctx, vms, vn, sms, sn))
    return (volume_ms = vms[], volume_launches = vn[], surface_ms = sms[], surface_launches = sn[])
end

# element_cache.U / F of the last rhs (TLW_FLAG_KEEP_AUX) and the indicator's alpha
get_aux!(out::Vector{Float64}, ctx, which) = check(ctx, ccall((:tlw_get_aux, lib), Cint,
    (Ptr{Cvoid}, Int32, Ptr{Float64}, Csize_t), ctx, which, out, length(out)))

end # module
