/*
 * trixilw_hip.h -- C ABI of libtrixilw_hip.so, the MI355X (gfx950) implementation
 * of TrixiLW's Lax-Wendroff flux-reconstruction and MDRK right-hand side on a
 * 2-D TreeMesh, compressible Euler, FP64.
 *
 * The entry points replace the reference's Julia methods on
 * SemidiscretizationHyperbolic (paths relative to the TrixiLW.jl tree):
 *   rhs!        src/semidiscretization/semidiscretization_hyperbolic.jl:4-20
 *   rhs_mdrk1!  src/semidiscretization/semidiscretization_hyperbolic.jl:22-38
 *   rhs_mdrk2!  src/semidiscretization/semidiscretization_hyperbolic.jl:40-56
 *   compute_dt  src/solvers/lwfr.jl:176-190 (Trixi max_dt)
 *   perform_step!/solve_lwfr loop  src/solvers/lwfr.jl:212-263, src/solvers/mdrk.jl:1-29
 *   SemidiscretizationHyperbolic + create_cache_serial
 *               src/semidiscretization/semidiscretization_hyperbolic.jl:58-84,
 *               src/solvers/dgsem_tree/containers.jl:7-64
 *
 * Conventions
 *   - Every function returns TLW_OK (0) or an error code; tlw_last_error()
 *     gives a message.  TLW_EADMISSIBLE plays the role of the reference's
 *     DomainError (src/solvers/adaptive.jl:6-21).
 *   - State arrays use Trixi's layout [v, i, j, e] (column-major, variable
 *     fastest), 4 variables (ρ, ρv1, ρv2, ρE), (N+1)^2 nodes per element.
 *   - Pointers named *_dev are device pointers; *_host are host pointers.
 *   - All work is ordered on the context's stream.  A context is not
 *     reentrant (one host thread at a time), like the reference's cache.
 */
#ifndef TRIXILW_HIP_H
#define TRIXILW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TLW_OK = 0,
    TLW_EINVAL = 1,      /* bad argument or description */
    TLW_EHIP = 2,        /* HIP runtime error */
    TLW_EADMISSIBLE = 3, /* non-admissible state (ρ <= 0 or p <= 0) */
    TLW_ENOTSUP = 4      /* configuration not supported */
};

/* time discretizations (src/basic_types.jl:4-6) */
enum { TLW_LW = 0, TLW_MDRK = 1 };

/* Argument order of the 8-argument Rusanov flux (src/equations/numerical_fluxes.jl:14-38):
 * TreeMesh passes (F̃,F̃,Ũ,Ũ,uⁿ,uⁿ): λ from Ũ, jump on uⁿ (src/solvers/dgsem_tree/dg_2d.jl:1928).
 * Structured/P4est pass (F̃n,F̃n,uⁿ,uⁿ,Ũ,Ũ): λ from uⁿ, jump on Ũ (src/solvers/dgsem_structured/dg_2d.jl:2271). */
enum { TLW_FLUX_TREE = 0, TLW_FLUX_STRUCTURED = 1 };

/* boundary value functions of the time-averaged Dirichlet condition
 * (src/equations/equations.jl:6-37), or TLW_BV_SLIP_WALL: the reflecting wall of
 * boundary_condition_slip_wall_vertical on x faces and boundary_condition_slip_wall_horizontal on
 * y faces (equations.jl:95-139: mirrored uⁿ and Ũ, F_outer = (-F1, F2, -F3, -F4) resp.
 * (-F1, -F2, F3, -F4), the LW flux in the boundary order of the Dirichlet call); TreeMesh only */
enum { TLW_BV_NONE = 0, TLW_BV_VORTEX = 1, TLW_BV_CONSTANT = 2, TLW_BV_RIEMANN4 = 3, TLW_BV_SLIP_WALL = 4,
       TLW_BV_TABLE = 5 /* caller-defined boundary_value_function: values uploaded by tlw_set_boundary_data */,
       TLW_BV_CONVERGENCE_TEST = 6 /* Trixi's initial_condition_convergence_test (Euler): ρ = ρv1 = ρv2 = φ,
                                      ρe = φ², φ = 2 + 0.1 sin(π(x + y - t)) */,
       /* conditions on uⁿ with the boundary node's contravariant vector as normal (curvilinear
        * StructuredMesh, Euler only; TLW_ENOTSUP elsewhere), in the form of their P4est call
        * (src/solvers/dgsem_p4est/dg_2d.jl:361-393).  The reference's StructuredMesh boundary call
        * passes an extra direction argument that these two methods do not take. */
       TLW_BV_OUTFLOW = 7 /* boundary_condition_outflow (equations.jl:185-197): the normal flux of uⁿ */,
       TLW_BV_SLIP_WALL_APPROXIMATE = 8 /* slip_wall_approximate (equations.jl:251-260): the Rusanov flux
                                           between uⁿ and its reflection about the normal, the reflected
                                           state on the side of the boundary */ };

/* source terms of SemidiscretizationHyperbolic (source_terms = ...), folded into the Taylor procedure
 * as the reference does (src/solvers/dgsem_tree/dg_2d.jl:89-176 and the calls in
 * lw_volume_kernel_3!/4!, 674-1267; tree/multi_derivative_2d.jl:323-379, 487-522):
 * TLW_SRC_CONVERGENCE_TEST is Trixi's source_terms_convergence_test for CompressibleEulerEquations2D
 * (the residual of the state of TLW_BV_CONVERGENCE_TEST).  Euler LW and MDRK on a TreeMesh, plain
 * VolumeIntegralFR, and LW and MDRK on a curvilinear StructuredMesh (structured/dg_2d.jl:1379-2105,
 * structured/multi_derivative_2d.jl:82-445; the reference's elixir_euler_source_terms_parallelogram.jl;
 * tlw_create_curved then needs the cache's node_coordinates), also on a P4estMesh (tlw_create_p4est,
 * whose volume kernels are the same); shock capturing, parabolic terms and adaptive steps with
 * source terms return TLW_ENOTSUP.
 * TLW_SRC_NAVIER_STOKES_CONVERGENCE_TEST is source_terms_navier_stokes_convergence_test of
 * examples/tree_2d_dgsem/elixir_navierstokes_convergence.jl (the residual of its manufactured
 * solution under the context's μ and Pr), in the Navier-Stokes MDRK stages on a TreeMesh
 * (multi_derivative_2d_parabolic.jl:221-711; tlw_create_parabolic with TLW_MDRK, periodic or with
 * walls) and in the LW Navier-Stokes kernels with walls (dg_2d_parabolic.jl:725-1860, as the
 * example's walls need); an LW context without walls returns TLW_ENOTSUP. */
enum { TLW_SRC_NONE = 0, TLW_SRC_CONVERGENCE_TEST = 1, TLW_SRC_NAVIER_STOKES_CONVERGENCE_TEST = 2 };

/* tlw_get_aux selectors (element_cache fields, src/solvers/dgsem_tree/containers.jl:73-98,
 * and the shock-capturing blending factors, indicator.cache.alpha) */
enum { TLW_AUX_U = 0, TLW_AUX_F = 1, TLW_AUX_ALPHA = 2 };

/* shock-capturing subcell reconstructions */
enum { TLW_RECON_FIRST_ORDER = 0, TLW_RECON_MUSCL_HANCOCK = 1 };

/* solver flags */
enum { TLW_FLAG_KEEP_AUX = 1 /* store full Ũ [v,i,j,e] and F̃ [v,d,i,j,e] on every rhs */ };

/* Mesh: Trixi TreeMesh element data (cache.elements) plus face neighbours
 * derived from cache.interfaces / cache.boundaries.  Faces between leaves of different levels are
 * described by tlw_mortar_desc (tlw_create_nonconforming). */
typedef struct {
    int32_t n_elements;
    const int32_t* neighbors_host;        /* [4*E]: neighbour of faces -x,+x,-y,+y; -1 = domain boundary */
    const double* inverse_jacobian_host;  /* [E] */
    const double* element_center_host;    /* [2*E] (x, y); used for boundary node coordinates */
    const double* element_length_host;    /* [E] */
} tlw_mesh_desc;

/* TreeMesh with refined cells (leaves of different levels, 2:1 across faces): a face whose
 * neighbours across it are two leaves of the next level, or that is one of those two leaves' faces,
 * carries neighbour TLW_NEIGHBOR_MORTAR and lies on one L2 mortar.  Each mortar record is Trixi's
 * mortar container entry (src/solvers/dgsem_tree/containers_2d.jl in Trixi.jl, read by the
 * reference's prolong2mortars!/calc_mortar_flux!, src/solvers/dgsem_tree/dg_2d.jl:1456-1646), 0-based:
 * {lower element, upper element, large element, orientation (1: x face, 2: y face), large_side
 * (1: the large element lies at -x/-y of the face, 2: at +x/+y)}; lower/upper is the small element at
 * the lower/upper half of the face coordinate (y on x faces, x on y faces). */
#define TLW_NEIGHBOR_MORTAR (-2)
typedef struct {
    int32_t n_mortars;
    const int32_t* mortars_host;  /* [5*n_mortars] */
} tlw_mortar_desc;

/* Basis: Trixi LobattoLegendreBasis fields, arrays column-major as Julia stores them. */
typedef struct {
    int32_t polydeg;                /* 3 or 4 */
    const double* nodes_host;       /* [n] */
    const double* weights_host;     /* [n] */
    const double* derivative_matrix_host; /* [n*n], D[i,k] at i + n*k */
    const double* derivative_dhat_host;   /* [n*n] */
} tlw_basis_desc;

typedef struct {
    int32_t time_discretization;    /* TLW_LW or TLW_MDRK */
    double gamma;                   /* CompressibleEulerEquations2D */
    int32_t surface_flux_order;     /* TLW_FLUX_TREE or TLW_FLUX_STRUCTURED */
    int32_t boundary_kind;          /* TLW_BV_* applied on every face with neighbour -1 */
    double boundary_params[24];
    int32_t flags;                  /* TLW_FLAG_* */
    /* VolumeIntegralFRShockCapturing(IndicatorHennemannGassner(...; variable=density_pressure),
     * volume_flux_fv=flux_lax_friedrichs, reconstruction=FirstOrderReconstruction())
     * (src/solvers/fr.jl:63-86, src/solvers/dgsem_tree/dg_2d.jl:1648-1991); TreeMesh order only */
    int32_t shock_capturing;        /* 0 = VolumeIntegralFR */
    double alpha_max, alpha_min;
    int32_t alpha_smooth;
    const double* inverse_vandermonde_host; /* [n*n] Trixi basis.inverse_vandermonde_legendre, column-major */
    /* subcell FV reconstruction of the blend: TLW_RECON_FIRST_ORDER (FirstOrderReconstruction,
     * dg_2d.jl:251-282) or TLW_RECON_MUSCL_HANCOCK (MUSCLHancockReconstruction, dg_2d.jl:284-324 with
     * calcflux_mh! of src/solvers/dgsem_structured/dg_2d.jl:313-538 on the TreeMesh's Cartesian
     * metric; elements with a face on the domain boundary use first order, as there) */
    int32_t reconstruction;
    int32_t source_terms;           /* TLW_SRC_* */
} tlw_solver_desc;

/* Parabolic terms of SemidiscretizationHyperbolicParabolic
 * (src/semidiscretization/semidiscretization_hyperbolic_parabolic.jl:8-118):
 * CompressibleNavierStokesDiffusion2D(equations; mu, Prandtl, gradient_variables)
 * with constant dynamic viscosity (src/equations/compressible_navier_stokes_2d.jl). */
enum { TLW_PAR_NAVIER_STOKES = 1 };
/* gradient variables: GradientVariablesConservative (compressible_navier_stokes_2d.jl:14-57)
 * is the one implemented; Trixi's GradientVariablesPrimitive is rejected (TLW_ENOTSUP). */
enum { TLW_GRAD_CONSERVATIVE = 0, TLW_GRAD_PRIMITIVE = 1 };
typedef struct {
    int32_t equations;           /* TLW_PAR_NAVIER_STOKES */
    double mu;                   /* dynamic viscosity */
    double prandtl;              /* Prandtl number; κ = γ / ((γ-1) Pr) */
    int32_t gradient_variables;  /* TLW_GRAD_CONSERVATIVE */
    /* 0: periodic mesh.  1: BoundaryConditionNavierStokesWall(NoSlip(v_wall), Adiabatic(0)) on every
     * boundary face (compressible_navier_stokes_2d.jl:81-124: the gradient's boundary value
     * prim2cons(ρ, v_wall, p) of the node, the viscous wall flux (F̃v1, F̃v2, F̃v3, v_wall·(F̃v2, F̃v3))
     * subtracted from the hyperbolic wall flux, dg_2d_parabolic.jl:1951-2085), with the wall velocity
     * (v1, v2) of x_neg, x_pos, y_neg, y_pos in wall_velocity; LW or MDRK (each stage's own
     * viscous time average at the wall, multi_derivative_2d_parabolic.jl:59-218) on a TreeMesh
     * with solver->boundary_kind = TLW_BV_SLIP_WALL (the reference's lid-driven cavity),
     * TLW_ENOTSUP otherwise */
    int32_t walls;
    double wall_velocity[8];
    /* per direction: 0 Adiabatic(0), 1 Isothermal(wall_temperature): the reference's isothermal wall
     * (compressible_navier_stokes_2d.jl:93-111): the gradient's boundary value prim2cons(ρ, v_wall,
     * ρ T_wall) (the reference's R = 1) and the inner viscous flux (F̃v1..F̃v4) as the wall's viscous flux;
     * a temperature that is not finite and > 0 is TLW_EINVAL */
    int32_t wall_isothermal[4];
    double wall_temperature[4];
} tlw_parabolic_desc;
/* Curvilinear StructuredMesh metrics (Trixi cache.elements of a StructuredMesh,
 * src/solvers/dgsem_structured/containers_2d.jl; TrixiLW's kernels for it in
 * src/solvers/dgsem_structured/dg_2d.jl).  Layouts are Julia's column-major arrays:
 * contravariant_vectors[m, l, i, j, element] = component m of Ja^l at node (i, j), so the four
 * entries of a node (Ja11, Ja12, Ja21, Ja22) are contiguous; inverse_jacobian[i, j, element]. */
typedef struct {
    const double* contravariant_vectors_host; /* [2*2*n*n*E] */
    const double* inverse_jacobian_host;      /* [n*n*E], every entry > 0 */
    const double* node_coordinates_host;      /* [2*n*n*E] node_coordinates[d, i, j, element]; needed (for the
                                                 boundary value function) when the mesh has boundary faces */
} tlw_curved_desc;

/* P4estMesh{2} (conforming: leaves of one level): the metric terms as in tlw_curved_desc plus Trixi's
 * P4est interface container (src/solvers/dgsem_p4est/containers_2d.jl in Trixi.jl, read by the
 * reference's prolong2interfaces!/calc_interface_flux!, src/solvers/dgsem_p4est/dg_2d.jl:8-192), its
 * node_indices reduced to faces and a flip: per interface {primary element, primary face, secondary
 * element, secondary face, flipped}, 0-based, faces 0..3 = -ξ, +ξ, -η, +η of the element's own frame
 * (any relative orientation), flipped = 1 when the secondary's face nodes run against the primary's.
 * mesh->neighbors_host holds each interface face's partner element and -1 on boundary faces. */
typedef struct {
    const double* contravariant_vectors_host; /* as in tlw_curved_desc */
    const double* inverse_jacobian_host;
    const double* node_coordinates_host;
    int32_t n_interfaces;
    const int32_t* interfaces_host; /* [5*n_interfaces] */
} tlw_p4est_desc;
typedef struct tlw_ctx tlw_ctx;

/* SemidiscretizationHyperbolic(mesh, time_disc, equations, ic, solver) + cache
 * (semidiscretization_hyperbolic.jl:58-84, containers.jl:7-64): uploads mesh and
 * basis, allocates the device state, traces and MDRK stage buffer. */
int tlw_create(const tlw_mesh_desc* mesh, const tlw_basis_desc* basis, const tlw_solver_desc* solver,
               tlw_ctx** out);
/* A batch: n_instances independent meshes (meshes[k], element ids local to each) with the same
 * basis and solver, advanced together by one launch per kernel and step of tlw_advance; each
 * instance keeps its own t, dt (its own CFL reduction), iteration count and finished flag.
 * No counterpart in the reference, which runs one semidiscretization per solve_lwfr.  States
 * are the instances' states concatenated in order; tlw_advance reports instance 0 and
 * `finished` when all are; tlw_batch_status reports every instance.  tlw_max_dt returns the
 * smallest of the instances' values. */
int tlw_create_batch(int32_t n_instances, const tlw_mesh_desc* meshes, const tlw_basis_desc* basis,
                     const tlw_solver_desc* solver, tlw_ctx** out);
/* SemidiscretizationHyperbolicParabolic(mesh, time_disc, (equations, equations_parabolic), ic, solver)
 * (semidiscretization_hyperbolic_parabolic.jl:8-118) on a periodic mesh, or with the no-slip walls of
 * tlw_parabolic_desc.walls on a non-periodic TreeMesh (LW): every rhs and step of the
 * context then follows rhs! / rhs_mdrk1! / rhs_mdrk2! of the parabolic path
 * (src/solvers/dgsem_tree/dg_2d_parabolic.jl:15-98, multi_derivative_2d_parabolic.jl:59-219):
 * BR1 gradient of the conservative variables, LW procedure on f_a - f_v with Taylor states of
 * the gradient, interface flux LW-Rusanov - ½(F̃v_L + F̃v_R); tlw_max_dt and the device loop use
 * compute_dt's min of the advective and viscous bounds (src/solvers/lwfr.jl:192-210,
 * compressible_navier_stokes_2d.jl:204-261).  Plain VolumeIntegralFR only; tlw_adjoint_lw returns
 * TLW_ENOTSUP on such a context; the positivity limiter applies as on any mesh (quadrature weights,
 * positivity_zhang_shu_dg2d.jl:8-45), the next dt then from the limited state with the viscous bound.  With TLW_FLAG_KEEP_AUX,
 * TLW_AUX_F holds the total F̃a - F̃v. */
int tlw_create_parabolic(const tlw_mesh_desc* mesh, const tlw_basis_desc* basis, const tlw_solver_desc* solver,
                         const tlw_parabolic_desc* parabolic, tlw_ctx** out);
/* SemidiscretizationHyperbolic on a curvilinear StructuredMesh: rhs! of
 * src/solvers/dgsem_structured/dg_2d.jl:17-53 with lw_volume_kernel_3!/4! (dg_2d.jl:1379-2105:
 * the Taylor procedure on the contravariant fluxes Ja^l · (f, g) and the node's inverse Jacobian)
 * and calc_interface_flux! (dg_2d.jl:2194-2290: the normal of the right element's face node,
 * F* = ½(F̃_L·n + F̃_R·n) - ½ λ(uⁿ, n)(Ũ_R - Ũ_L) with λ = max|v·n| + max c |n|); tlw_max_dt and
 * the device loop use Trixi's curvilinear max_dt (per node |Ja^l · (λ1, λ2)| |J⁻¹|).
 * mesh->neighbors_host gives the face neighbours, -1 on a domain boundary, where the time-averaged
 * Dirichlet condition of solver->boundary_kind applies as in calc_boundary_flux! for StructuredMesh
 * (dg_2d.jl:2297-2391: the element's own face normal, TreeMesh argument order, the average over the
 * full step in both MDRK stages); mesh->inverse_jacobian_host, element_center_host and
 * element_length_host are not read and may be NULL.  The structured flux order applies whatever
 * solver->surface_flux_order says.
 * With TLW_MDRK, tlw_rhs_mdrk1/2 are rhs_mdrk1!/2! of src/solvers/dgsem_structured/
 * multi_derivative_2d.jl:1-80 (mdrk_kernel_1!/2!, 82-445: F̃ = f/2 + f_t/8, then F2 + f_t(us)/3,
 * in contravariant form for the volume term).  VolumeIntegralFRShockCapturing with
 * TLW_RECON_FIRST_ORDER blends as on the TreeMesh (IndicatorHennemannGassner on uⁿ) with Trixi's
 * calcflux_fv! for curvilinear meshes (structured/dg_2d.jl:111-173: Rusanov fluxes with the
 * subcell normals Ja^l(node 0) + Σ_{q<k} w_q (D Ja^l)(q); a pure-DG element's fn_low at subcell
 * face n-1 takes the reference's calc_fn_low! normal, the q = 0 and q = n-2 terms only) and
 * compute_alp with each face node's own inverse Jacobian (dg_2d.jl:2169-2192); boundary faces are
 * not blended.  TLW_RECON_MUSCL_HANCOCK, and alpha smoothing on a mesh with boundaries (Trixi
 * asserts a periodic StructuredMesh there), return TLW_ENOTSUP, and so do several instances,
 * adaptive steps, tlw_adjoint_lw and tlw_rebuild_mesh on such a context.  The positivity limiter
 * (my_limiter_zhang_shu!, quadrature weights only, positivity_zhang_shu_dg2d.jl:8-45, for any
 * AbstractMesh{2}) applies, the next dt then from the limited state with the curvilinear speeds.
 * A non-positive inverse Jacobian returns TLW_EINVAL. */
int tlw_create_curved(const tlw_mesh_desc* mesh, const tlw_basis_desc* basis, const tlw_solver_desc* solver,
                      const tlw_curved_desc* metrics, tlw_ctx** out);
/* SemidiscretizationHyperbolicParabolic (compressible Navier-Stokes, as tlw_create_parabolic) on a
 * periodic curvilinear mesh: the LW kernels of src/solvers/dgsem_p4est/dg_2d_parabolic.jl
 * (641-1787: contravariant fluxes Ja·(fa - fv), u_t scaled by the node's J⁻¹, element-local
 * gradients of the time derivatives by compute_local_grad_u!, 51-104, i.e. D in ξ and η turned
 * with the transposed Ja and scaled by J⁻¹) with Trixi's curvilinear BR1 gradient (reference-space
 * D̂ and lifted ½(u_L + u_R), the element's own outward normal, transposed Ja, J⁻¹), and the
 * interfaces of tlw_create_curved (one normal per face node; F* = LW-Rusanov(F̃a) -
 * ½(F̃v_L + F̃v_R)·n, p4est/dg_2d_parabolic.jl:1878-1960).  MDRK stages are the TreeMesh parabolic
 * ones with these operators (the reference has no curvilinear MDRK parabolic kernel).  dt =
 * cfl · min(curvilinear max_dt, 1/((λa + λv)·2(2N+1))), λv = 2(2N+1) max J⁻¹ max(4μ/(3ρ),
 * μκ/(ρ Pr)) with the node's J⁻¹ (the reference's P4est max_dt reads inverse_jacobian[element] of
 * a per-node array; this is the formula that reduces to its TreeMesh one on an affine mesh).
 * A mesh with boundary faces, shock capturing and several instances return TLW_ENOTSUP. */
int tlw_create_curved_parabolic(const tlw_mesh_desc* mesh, const tlw_basis_desc* basis,
                                const tlw_solver_desc* solver, const tlw_curved_desc* metrics,
                                const tlw_parabolic_desc* parabolic, tlw_ctx** out);
int tlw_n_instances(const tlw_ctx* ctx, int32_t* n_instances);
int tlw_batch_status(tlw_ctx* ctx, int32_t n_instances, double* t_host, double* dt_host, int64_t* iter_host,
                     int32_t* finished_host);
/* An ensemble of uneven instances: one batch context of polydeg 3 and one of polydeg 4 (plain LW on
 * periodic TreeMeshes, tiles (16, 10) or (8, 5) elements per volume workgroup) advanced by the same
 * launches: each step is one volume, one surface and one finalize launch for every instance of both,
 * optionally replayed from a captured HIP graph.  No counterpart in the reference (one
 * semidiscretization per solve_lwfr).  The member contexts keep their states (tlw_set_state,
 * tlw_get_state, tlw_batch_status work on them as before); tlw_ensemble_solve_init is tlw_solve_init
 * on both (CFL number per context); tlw_ensemble_advance runs up to max_steps steps and reports whether
 * every instance is finished.  The ensemble does not own the contexts: destroy it first. */
typedef struct tlw_ensemble tlw_ensemble;
int tlw_create_ensemble(tlw_ctx* ctx_a, tlw_ctx* ctx_b, tlw_ensemble** out);
int tlw_ensemble_set_stream(tlw_ensemble* ens, void* stream);
int tlw_ensemble_enable_graphs(tlw_ensemble* ens, int32_t enable);
int tlw_ensemble_solve_init(tlw_ensemble* ens, double t0, double t_final, double cfl_polydeg3, double cfl_polydeg4);
int tlw_ensemble_advance(tlw_ensemble* ens, int32_t max_steps, int32_t* finished_host);
const char* tlw_ensemble_last_error(const tlw_ensemble* ens);
void tlw_ensemble_destroy(tlw_ensemble* ens);
/* Replay tlw_advance's steps from a captured HIP graph (one graph per step count, captured on
 * first use; needs an explicit stream from tlw_set_stream).  No counterpart in the reference. */
int tlw_enable_graphs(tlw_ctx* ctx, int32_t enable);

/* AMR or any host-side mesh change: rebuild the device mesh (src/callbacks/amr.jl:2-20 resize!) */
int tlw_rebuild_mesh(tlw_ctx* ctx, const tlw_mesh_desc* mesh);

/* A TreeMesh with refined cells (mortars, see tlw_mortar_desc): the rhs adds prolong2mortars! and
 * calc_mortar_flux! of src/solvers/dgsem_tree/dg_2d.jl:1456-1646 (also in MDRK,
 * tree/multi_derivative_2d.jl:37-46, 98-107) with Trixi's LobattoLegendreMortarL2 and
 * mortar_fluxes_to_elements!: on each half of a mortar the large element's uⁿ, Ũ and F̃·n
 * interpolated to the small face's nodes, the LW Rusanov flux in the interface order, the small
 * element taking it and the large element its L2 projection.  Euler, LW or MDRK, VolumeIntegralFR,
 * one instance; shock capturing, parabolic terms and curvilinear meshes return TLW_ENOTSUP, and so
 * do the adjoints.  Source terms, the positivity limiter, adaptive steps (polydeg 4), the device
 * loop and graphs work as on a uniform TreeMesh. */
int tlw_create_nonconforming(const tlw_mesh_desc* mesh, const tlw_mortar_desc* mortars, const tlw_basis_desc* basis,
                             const tlw_solver_desc* solver, tlw_ctx** out);
/* SemidiscretizationHyperbolic on a P4estMesh (tlw_p4est_desc): the curvilinear volume kernels of
 * tlw_create_curved with the P4est surface terms of src/solvers/dgsem_p4est/dg_2d.jl:8-393: at an
 * interface each side's F̃ in the primary's outward direction from its own contravariant vector, the
 * LW Rusanov flux with λ from uⁿ and the jump on Ũ, the primary keeping F*, the secondary -F* at its own
 * (possibly reversed) node; boundaries in the conditions' P4est methods with the outward normal:
 * the time-averaged Dirichlet value (equations.jl:66-93: surface_flux(F̃·n, F_outer, uⁿ, U_outer, Ũ,
 * U_outer, n)), TLW_BV_OUTFLOW and TLW_BV_SLIP_WALL_APPROXIMATE; MDRK stage 1 takes the boundary's half
 * time average (scaling_factor 0.5, tree/multi_derivative_2d.jl:34).  LW or MDRK, VolumeIntegralFR,
 * Euler; shock capturing, parabolic terms, source terms, TLW_BV_SLIP_WALL, mortars and the adjoints
 * return TLW_ENOTSUP. */
int tlw_create_p4est(const tlw_mesh_desc* mesh, const tlw_p4est_desc* p4est, const tlw_basis_desc* basis,
                     const tlw_solver_desc* solver, tlw_ctx** out);
/* AMR re-upload of a TreeMesh with refined cells (src/callbacks/amr.jl:2-208 resize! of the element,
 * interface, boundary and mortar caches). */
int tlw_rebuild_mesh_nonconforming(tlw_ctx* ctx, const tlw_mesh_desc* mesh, const tlw_mortar_desc* mortars);
/* AMR solution transfer (Trixi's refine!/coarsen! of the DG solution, which the reference's AMR
 * callback runs before src/callbacks/amr.jl:2-20 resizes the caches): u_new_dev [n_new elements] from
 * u_old_dev [n_old elements], both in the context's state layout, with the mortar operators of the
 * context's basis.  plan_host [5*n_new], per new element: {0, e} keeps old element e; {1, p, child}
 * is a child of old element p (child 0 lower left, 1 lower right, 2 upper left, 3 upper right,
 * forward_lower/upper in x then y); {2, c0, c1, c2, c3} is the parent of old elements c0..c3 (same
 * child order, reverse_lower/upper, summed).  Independent of the context's mesh, so it runs between
 * the host-side mesh change and tlw_rebuild_mesh_nonconforming.  Uploads the plan to a temporary
 * device buffer and waits for the context's stream. */
int tlw_transfer_solution(tlw_ctx* ctx, const double* u_old_dev, int32_t n_old, double* u_new_dev, int32_t n_new,
                          const int32_t* plan_host);
void tlw_destroy(tlw_ctx* ctx);
const char* tlw_last_error(const tlw_ctx* ctx);

/* Stream on which all work of the context is ordered (hipStream_t passed as void*; NULL = default). */
int tlw_set_stream(tlw_ctx* ctx, void* stream);

/* Number of doubles in a state vector: 4 * (N+1)^2 * E (Trixi ndofs * nvariables). */
size_t tlw_state_len(const tlw_ctx* ctx);

/* rhs!(du_ode, u_ode, semi, t) with cache.dt[1] = dt (semidiscretization_hyperbolic.jl:4-20,
 * dg_2d.jl:28-83).  du is overwritten. */
int tlw_rhs_lw(tlw_ctx* ctx, const double* u_dev, double* du_dev, double t, double dt);
/* rhs_mdrk1!/rhs_mdrk2! (semidiscretization_hyperbolic.jl:22-56, tree/multi_derivative_2d.jl:1-116).
 * Stage 2 reads us from the context's stage buffer (tlw_us_ptr), and u must be the
 * same uⁿ that stage 1 received, as in mdrk.jl:13-25. */
int tlw_rhs_mdrk1(tlw_ctx* ctx, const double* u_dev, double* du_dev, double t, double dt);
int tlw_rhs_mdrk2(tlw_ctx* ctx, const double* u_dev, double* du_dev, double t, double dt);
/* Device pointer of the MDRK stage state `us` (element_cache.mdrk_cache._us), and a
 * stream-ordered device-to-device copy into it (update_soln!(integrator, _us, uprev, du),
 * src/solvers/mdrk.jl:17). */
int tlw_us_ptr(tlw_ctx* ctx, double** us_dev);
int tlw_set_us(tlw_ctx* ctx, const double* us_dev);

/* Trixi max_dt(u, t, mesh::TreeMesh{2}, ...) = 2 / (n * max_e J⁻¹ (max λ1 + max λ2));
 * compute_dt multiplies it by the CFL number (lwfr.jl:176-190). */
int tlw_max_dt(tlw_ctx* ctx, const double* u_dev, double* max_dt_host);
/* Allocation-free variant of tlw_max_dt, with no counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Needs no context: u_dev is a TreeMesh Euler state [v,i,j,e] of n_elements
 * elements at polydeg 3 or 4, inverse_jacobian_dev its [n_elements] inverse Jacobians; the result
 * 2 / (n max_e J⁻¹ (max λ1 + max λ2)), the same bits tlw_max_dt gives, is written to max_dt_dev
 * (device memory) in stream order on `stream` (hipStream_t, NULL = default), NaN if a node has
 * ρ <= 0 or p <= 0.  No allocation, no host synchronisation.  Partial maxima use a fixed 2 MiB
 * device array per GPU (held until the process ends): n_elements above 1 310 720 (polydeg 4) or
 * 2 097 152 (polydeg 3) returns TLW_EINVAL.  Calls that overlap on one GPU share that array and
 * race.  Errors: tlw_last_error(NULL) gives the message. */
int tlw_max_dt_noalloc(const double* u_dev, const double* inverse_jacobian_dev, int32_t n_elements, int32_t polydeg,
                       double gamma, double* max_dt_dev, void* stream);

/* Device-resident time loop: the state lives in the context (tlw_set_state /
 * tlw_get_state / tlw_state_ptr).  tlw_solve_init sets t, the final time and the
 * CFL number and computes the first dt on the device; tlw_advance runs up to
 * max_steps steps of solve_lwfr's loop (compute_dt, set_dt! truncation at t_final,
 * rhs!/rhs_mdrk*, u = uprev + dt du, set_t_and_iter!, isfinished = t ≈ t_final)
 * without host synchronisation between steps, then reports the time and the
 * number of steps actually taken. */
int tlw_set_state(tlw_ctx* ctx, const double* u_host, size_t len);
int tlw_get_state(tlw_ctx* ctx, double* u_host, size_t len);
int tlw_state_ptr(tlw_ctx* ctx, double** u_dev);
/* Stream-ordered device-to-device copies into / out of the loop's state (integrator.u held in
 * device memory by the caller), the device counterparts of tlw_set_state / tlw_get_state. */
int tlw_set_state_dev(tlw_ctx* ctx, const double* u_dev);
int tlw_get_state_dev(tlw_ctx* ctx, double* u_dev);
int tlw_solve_init(tlw_ctx* ctx, double t0, double t_final, double cfl);
int tlw_advance(tlw_ctx* ctx, int32_t max_steps, double* t_host, double* dt_host, int64_t* iter_host,
                int32_t* finished_host);
/* BoundaryConditionDirichlet(boundary_value_function) for any function (src/equations/equations.jl:6-37;
 * TreeMesh and curvilinear StructuredMesh): with solver->boundary_kind = TLW_BV_TABLE the caller
 * evaluates its function on the host and uploads the states; the device forms the time averages of
 * the state and its flux (the GLL rule over [t, t + s dt]) and the Rusanov flux as for the built-in
 * functions.  tlw_boundary_nodes gives the number of boundary face nodes and (xy_host != NULL, len =
 * 2 n_nodes) their coordinates [node][2] in table order: element, face (-x, +x, -y, +y), face node.
 * tlw_set_boundary_data uploads the conservative states [node][level][4] (stream-ordered):
 * time_levels = 1, a steady state used at every time (the device loop accepts it), or polydeg + 2:
 * levels q = 0..polydeg at ts_q = t + ½ s dt (ξ_q + 1) with the GLL nodes ξ_q, then level polydeg+1 at
 * t, for the next rhs call's t and dt (s = 0.5 for tlw_rhs_mdrk1 on a TreeMesh, the reference's
 * scaling_factor, else 1); the device loop refuses such per-call data (TLW_ENOTSUP). */
int tlw_boundary_nodes(tlw_ctx* ctx, int64_t* n_nodes, double* xy_host, size_t len);
int tlw_set_boundary_data(tlw_ctx* ctx, const double* values_host, size_t len, int32_t time_levels);
/* Per-step dt record and prescribed dts of the device loop on a single-instance context, for the
 * checkpointed gradient driver (no counterpart in the reference, whose LWIntegrator keeps dt on the
 * host).  tlw_record_dts(capacity): the dt of step k (k = iterations since the last tlw_solve_init,
 * 0-based) is stored on the device while k < capacity (0 turns it off); tlw_get_dts copies the first
 * n.  tlw_prescribe_dts(dts, n): step k takes dts[k] instead of cfl · max_dt while k < n (set_dt!'s
 * truncation at the final time still applies; n = 0 turns it off); CFL steps only. */
int tlw_record_dts(tlw_ctx* ctx, int64_t capacity);
int tlw_get_dts(tlw_ctx* ctx, double* dts_host, int64_t n);
int tlw_prescribe_dts(tlw_ctx* ctx, const double* dts_host, int64_t n);

/* Adaptive (error-controlled) time stepping, solve_lwfr with time_step_computation = Adaptive()
 * (src/solvers/adaptive.jl:95-173, lwfr.jl:141-171): replaces tlw_solve_init.  dt starts at
 * dt_initial (set_dt!); every step of tlw_advance is one attempt: the LW step with the temporal
 * error estimate of lw_volume_kernel_4! (dg_2d.jl:1183-1199, 1251-1265; abstol/reltol are the
 * reference's `tolerances`), the admissibility test of the updated state, then the controller
 * (β = 0.6, -0.2, 0, k = nnodes): a rejected attempt (non-admissible state or factor <= 0.81)
 * restarts from uⁿ with min(factor, 0.95)·dt, an accepted one advances t and proposes
 * min(factor, 1.5)·dt.  tlw_advance's iter counts accepted steps.  Euler LW on a TreeMesh at
 * polydeg 4 (the reference's TreeMesh estimate is in that kernel), or on a curvilinear StructuredMesh
 * at polydeg 3 or 4 (lw_volume_kernel_3!/4! of src/solvers/dgsem_structured/dg_2d.jl: u_np1 with the
 * contravariant F̃ and the node's J⁻¹, u_np1_low without the last stage's term, 1595-1681 and
 * 1983-2077), or compressible Navier-Stokes LW at polydeg 4 on a periodic TreeMesh (the total flux's
 * estimate of lw_volume_kernel_4! in src/solvers/dgsem_tree/dg_2d_parabolic.jl:1749-1775, 1841-1855);
 * not combined with the positivity limiter or source terms (TLW_ENOTSUP).  A rejected step whose factor is not finite (NaN state)
 * shrinks dt by 0.95. */
int tlw_solve_init_adaptive(tlw_ctx* ctx, double t0, double t_final, double dt_initial, double abstol,
                            double reltol);
/* integrator.stats.nreject: rejected attempts of the adaptive loop, summed over instances. */
int tlw_rejected_steps(tlw_ctx* ctx, int64_t* nreject);
/* Measurement hook (no counterpart in the reference; its TimerOutputs/performance_counter
 * timing, semidiscretization_hyperbolic.jl:11-17, is per rhs! call): when enabled,
 * tlw_advance records HIP events around every volume and surface kernel on the context's
 * stream; tlw_kernel_times returns the accumulated milliseconds and launch counts and
 * resets them. */
int tlw_enable_kernel_timing(tlw_ctx* ctx, int32_t enable);
int tlw_kernel_times(tlw_ctx* ctx, double* volume_ms, int64_t* volume_launches, double* surface_ms,
                     int64_t* surface_launches);

/* Discrete adjoint of one LW step u' = u + dt rhs!(u) (periodic mesh, TreeMesh flux order):
 * lam = lam_next + dt (∂rhs/∂u)ᵀ lam_next at the state u.  No counterpart in the reference, whose
 * gradients go through ForwardDiff; it differentiates the same rhs! (dg_2d.jl:28-83 with
 * lw_volume_kernel_4!/3!, 674-1267, and the TreeMesh Rusanov flux, 1903-1946), with dt frozen.
 * The max in λ passes its derivative to the side the forward max picked.  With
 * VolumeIntegralFRShockCapturing and first-order subcell FV, the blend (1 - α) LW + α FV of the
 * volume term (dg_2d.jl:1648-1804, fv_kernel! 251-282) and the blended interface flux
 * alp fn + (1 - alp) F* (1873-1946) are differentiated with the blending factors α of the state u
 * and the interface's admissibility decision frozen.  Uses the context's trace buffer (a buffer of
 * its own with shock capturing) as scratch; lam must not alias u or lam_next. */
int tlw_adjoint_lw(tlw_ctx* ctx, const double* u_dev, const double* lam_next_dev, double* lam_dev, double t,
                   double dt);

/* Discrete adjoint of one MDRK step (plain MDRK, periodic mesh, TreeMesh flux order): with
 * us = u + dt rhs_mdrk1!(u) and u' = u + dt rhs_mdrk2!(u, us) (src/solvers/mdrk.jl:13-25),
 * lam = lam_next + dt (∂R2/∂u)ᵀ lam_next + λ_us + dt (∂R1/∂u)ᵀ λ_us, λ_us = dt (∂R2/∂us)ᵀ lam_next.
 * No counterpart in the reference; it differentiates the stage residuals of
 * tree/multi_derivative_2d.jl:263-555 (F̃ = f/2 + f_t/8, then f + f_t(uⁿ)/6 + f_t(us)/3) and the
 * TreeMesh Rusanov flux, with dt frozen.  us is the stage state of the step (device).  Uses the
 * context's trace buffer and a state-sized buffer of its own as scratch; lam must not alias u, us or
 * lam_next. */
int tlw_adjoint_mdrk(tlw_ctx* ctx, const double* u_dev, const double* us_dev, const double* lam_next_dev,
                     double* lam_dev, double t, double dt);

/* PositivityPreservingLimiterZhangShu(thresholds = (threshold_density, threshold_pressure),
 * variables = (density, pressure)) as a stage limiter of solve_lwfr
 * (src/callbacks/callbacks_stage/positivity_zhang_shu.jl:10-41, positivity_zhang_shu_dg2d.jl:8-45;
 * applied after each update, src/solvers/lwfr.jl:226, 251, 257).  tlw_set_positivity_limiter
 * turns it on for tlw_advance (the next step's dt is then taken from the limited state);
 * tlw_apply_positivity_limiter is limiter!(u_ode, integrator, semi, t) on a device state. */
int tlw_set_positivity_limiter(tlw_ctx* ctx, int32_t enable, double threshold_density, double threshold_pressure);
int tlw_apply_positivity_limiter(tlw_ctx* ctx, double* u_dev);

/* Tile of the fused volume kernel (no counterpart in the reference, whose @threaded loop
 * has no tile): elements per workgroup, one lane per GLL node.  polydeg 4: 5, 10 or 15
 * (default 15); polydeg 3: 4, 8 or 16 (default 16).  Results do not depend on the tile.
 * tlw_autotune times `reps` rhs evaluations per tile on the current state buffer (after
 * tlw_set_state), keeps the fastest and writes it to *chosen and the per-rhs milliseconds
 * of the three tiles, in the order above, to ms[3] (may be NULL).  The other state buffer
 * (and, for MDRK, the stage buffer) is used as scratch. */
int tlw_set_elements_per_workgroup(tlw_ctx* ctx, int32_t elements);
int tlw_get_elements_per_workgroup(const tlw_ctx* ctx, int32_t* elements);
int tlw_autotune(tlw_ctx* ctx, int32_t reps, int32_t* chosen, double* ms);

/* Copy an element-cache array to the host.  TLW_AUX_U: Ũ [v,i,j,e], len 4 n^2 E, and
 * TLW_AUX_F: F̃ [v,d,i,j,e], len 8 n^2 E, of the last rhs call (need TLW_FLAG_KEEP_AUX);
 * TLW_AUX_ALPHA: the smoothed blending factors [E] of the last rhs call or device-loop
 * stage (shock capturing; Trixi's indicator cache alpha). */
int tlw_get_aux(tlw_ctx* ctx, int32_t which, double* out_host, size_t len);

#ifdef __cplusplus
}
#endif
#endif /* TRIXILW_HIP_H */
