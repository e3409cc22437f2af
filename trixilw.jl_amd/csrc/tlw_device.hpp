// tlw_device.hpp -- gfx950 kernels of the TreeMesh Lax-Wendroff / MDRK step.
//
// Work decomposition: one thread per GLL node; a volume workgroup owns EB
// consecutive elements (with Z-ordered meshes, 16 consecutive elements form
// a 4x4 block of the tree).  Fluxes and flux time derivatives are exchanged
// through LDS for the derivative-matrix contractions; Taylor states, Ũ and F̃
// stay in registers.
//
// One step is three launches:
//   volume    uⁿ -> approximate LW procedure, volume term D̂F̃, then the
//             Rusanov flux of every interface whose two elements belong to
//             the workgroup (both face records are in LDS), the lifted
//             surface terms, the Jacobian and the update of every node.
//             Faces leaving the workgroup write a 10-double face record.
//   surface   the remaining faces: Rusanov flux from the two face records
//             (or the time-averaged Dirichlet state on a domain boundary),
//             lifted correction of those face nodes.
//   finalize  CFL reduction and compute_dt / set_dt! / set_t_and_iter! /
//             isfinished on the device state.
// The reference computes the same quantities in
//   lw_volume_kernel_3!/4!         src/solvers/dgsem_tree/dg_2d.jl:674-1267
//   prolong2interfaces!            src/solvers/dgsem_tree/dg_2d.jl:1269-1326
//   calc_interface_flux!           src/solvers/dgsem_tree/dg_2d.jl:1903-1946
//   calc_boundary_flux!            src/solvers/dgsem_tree/dg_2d.jl:1379-1454
//   Trixi calc_surface_integral!, apply_jacobian!, max_dt
//   update_soln!/set_dt!/set_t_and_iter!  src/solvers/lwfr.jl:70-103
//   mdrk_kernel_1!/2!              src/solvers/dgsem_tree/multi_derivative_2d.jl:263-555
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tlw_sincos.hpp"

namespace tlw {

constexpr int kBcSlipWall = 4;  // boundary kind of the reflecting wall (TLW_BV_SLIP_WALL)
constexpr int kBcTable = 5;     // caller-defined boundary data (TLW_BV_TABLE)
constexpr int kBcConvergenceTest = 6;  // the convergence test's state (TLW_BV_CONVERGENCE_TEST)
constexpr int kBcOutflow = 7;          // the normal flux of uⁿ (TLW_BV_OUTFLOW, curvilinear mesh)
constexpr int kBcSlipWallApprox = 8;   // Rusanov with uⁿ's reflection (TLW_BV_SLIP_WALL_APPROXIMATE, curvilinear)
constexpr int kTraceQ = 10;  // face record: jump state (4), normal flux (4), |v_n|, c
// shock-capturing face record: + fn_low (4), blending factor, padding
constexpr int kTraceQSC = 16;
template <bool SC> constexpr int kRQ = SC ? kTraceQSC : kTraceQ;
constexpr int kSurfBlock = 128;

// Volume workgroup geometry: EB elements, NT threads (a multiple of 64).
template <int N, int EB> struct VGeo {
    static constexpr int n = N + 1;
    static constexpr int NN = n * n;
    static constexpr int S = EB * NN;  // LDS stride of one exchanged variable
    static constexpr int NT = ((EB * NN + 63) / 64) * 64;
    static constexpr int LDS = 3 * 8 * S;  // three exchange buffers
};
// Surface / reduction workgroup geometry: 128 threads.
template <int N> struct SGeo {
    static constexpr int n = N + 1;
    static constexpr int NN = n * n;
    static constexpr int EB = kSurfBlock / NN;
    static constexpr int S = EB * NN;
};

// Basis and boundary data, uploaded once per context.
struct DevConst {
    double D[8][8];      // derivative_matrix, row-major D[i][k]
    double Dhat[8][8];   // derivative_dhat
    double nodes[8], weights[8];
    double bnd_lo, bnd_hi;  // boundary_interpolation[1,1], [n,2]
    double gamma;
    double bc_par[24];
    int bc_kind;
    int flux_order;
    // shock capturing (IndicatorHennemannGassner on density_pressure)
    double Vinv[8][8];  // inverse normalised-Legendre Vandermonde, row-major
    double alpha_max, alpha_min;
    int alpha_smooth;
    // positivity limiter thresholds for density and pressure
    double pos_threshold[2];
    // CompressibleNavierStokesDiffusion2D (conservative gradient variables):
    // dynamic viscosity, Prandtl number, κ = γ/((γ-1) Pr)
    double visc_mu, visc_Pr, visc_kappa;
    // no-slip adiabatic walls of the Navier-Stokes path on every boundary face:
    // wall velocity (v1, v2) per face 0..3 (x_neg, x_pos, y_neg, y_pos)
    int ns_walls;
    double wall_v[8];
    // per direction: 1 for Isothermal(wall_T) (NoSlip with a wall temperature), 0 Adiabatic(0)
    int wall_iso[4];
    double wall_T[4];
    // MUSCL-Hancock subcell geometry per node i (compute_subcells): spacings
    // to the neighbouring nodes, the centred-difference weights, the distances
    // from the node to its subcell faces, and hw[k] = xs[k+2] - xs[k+1], the
    // width the half-step divides by for subcell k
    int recon;  // TLW_RECON_*
    double mh_dx1[8], mh_dx2[8], mh_fa[8], mh_fb[8], mh_fc[8], mh_xl[8], mh_xr[8], mh_hw[8];
    int src_kind;  // SrcKind
    // L2 mortar operators of a TreeMesh with refined cells (LobattoLegendreMortarL2),
    // row-major [output node][input node]: [0] lower half, [1] upper half
    double mfwd[2][8][8];  // large face -> the half's GLL nodes
    double mrev[2][8][8];  // the half's fluxes -> large face (L2 projection)
};

// Device-resident integrator state (LWIntegrator fields t, dt, iter and the
// isfinished test), updated by the finalize kernel.
struct DevState {
    double t, dt, t_final, cfl;
    double invJ0;   // interface inverse Jacobian of compute_alp: the instance's first element
    double query;   // Trixi max_dt result of a query reduction
    int finished;
    int cur;        // state buffer holding uⁿ
    int bad_state;  // set when a non-admissible state was met
    long long iter;
    // Adaptive time stepping (src/solvers/adaptive.jl): the error estimates
    // epsilon[-1], epsilon[0], epsilon[1], tolerances, the instance's Trixi ndofs,
    // the retry count of the current step and the accumulated rejections
    double eps[3];
    double abstol, reltol, ndofs;
    int adaptive, nfail;
    long long nreject;
};

enum BufSel { SEL_EXPLICIT = 0, SEL_CUR = 1, SEL_OTHER = 2 };
// RED_ADAPT: adaptive step (per-workgroup sums of the temporal errors and the
// admissibility test of the updated state instead of CFL maxima)
enum Reduce { RED_NONE = 0, RED_STEP = 1, RED_INIT = 2, RED_QUERY = 3, RED_ADAPT = 4 };
// VOL_LW_ADAPT: the LW kernel with the temporal-error estimate of adaptive steps
// (a separate instantiation, so the plain LW kernel carries none of its registers)
// VOL_LW_WALL, VOL_MDRK1_WALL, VOL_MDRK2_WALL: the Navier-Stokes LW and MDRK
// kernels that also keep the viscous time average of boundary face nodes
// (no-slip walls)
// VOL_*_SRC: the Euler LW and MDRK kernels with source terms (separate
// instantiations, VOL_LW_SRC + the plain kind)
enum VolKind {
    VOL_LW = 0, VOL_MDRK1 = 1, VOL_MDRK2 = 2, VOL_LW_ADAPT = 3, VOL_LW_WALL = 4,
    VOL_LW_SRC = 5, VOL_MDRK1_SRC = 6, VOL_MDRK2_SRC = 7, VOL_MDRK1_WALL = 8, VOL_MDRK2_WALL = 9,
    VOL_LW_ADAPT_WALL = 10
};
constexpr bool vol_is_src(int k) { return k >= VOL_LW_SRC && k <= VOL_MDRK2_SRC; }
constexpr bool vol_is_wall(int k) {
    return k == VOL_LW_WALL || k == VOL_MDRK1_WALL || k == VOL_MDRK2_WALL || k == VOL_LW_ADAPT_WALL;
}
// the plain kind of a source-term or wall instantiation
constexpr int vol_base(int k) {
    return vol_is_src(k)           ? k - VOL_LW_SRC
           : k == VOL_LW_WALL      ? (int)VOL_LW
           : k == VOL_MDRK1_WALL   ? (int)VOL_MDRK1
           : k == VOL_MDRK2_WALL   ? (int)VOL_MDRK2
           : k == VOL_LW_ADAPT_WALL ? (int)VOL_LW_ADAPT
                                   : k;
}
// source-term kinds of DevConst::src_kind
enum SrcKind { SRC_NONE = 0, SRC_CONVERGENCE_TEST = 1, SRC_NS_CONVERGENCE_TEST = 2 };

struct KArgs {
    int E;
    const double* u;      // uⁿ (explicit)
    const double* us;     // MDRK stage state
    double* out;          // volume: result; surface: corrected in place
    double* bufs[2];      // ping-pong state buffers for the device-resident loop
    int in_sel, out_sel;  // BufSel (the surface kernel reads the input at curvilinear boundary faces)
    double* trace;        // [E][4][n][kTraceQ], written for faces leaving a volume workgroup
    const int* nbr;       // [E][4]
    const double* inv_jac;
    const double* center;  // [E][2]
    const double* length;  // [E]
    const DevConst* c;
    DevState* st;
    double* auxU;      // optional full Ũ
    double* auxF;      // optional full F̃
    // per-workgroup partials: the volume kernel's temporal-error sums (adaptive
    // steps), then the node-per-lane kernels' maxima of J⁻¹(max λ1 + max λ2)
    double* partial;
    int vol_grid;      // number of volume workgroups (offset of the node-per-lane partials)
    int ebv;           // elements per volume workgroup
    // Workgroup tables: {first element, elements, instance, 0} per workgroup of the
    // volume grid (vblk) and of the node-per-lane grids (sblk).  Workgroups never
    // straddle two instances of a batch, so each reads one instance's state.
    const int4* vblk;
    const int4* sblk;
    const int* etile;  // [E] volume workgroup of each element
    const double* alpha_tmp;  // indicator values before smoothing [E]
    double* alpha_out;        // smoothed blending factors [E] (optional)
    // hyperbolic-parabolic CFL reduction: the per-workgroup maxima of the
    // viscous bound J⁻² max(4μ/(3ρ), μκ/(ρ Pr)) sit pv_off entries after the
    // advective ones (0: no parabolic terms)
    int pv_off;
    // curvilinear StructuredMesh (null on a TreeMesh): per node the contravariant
    // vectors (Ja11, Ja12, Ja21, Ja22) [E][n²][4] and inverse Jacobian [E][n²],
    // per face node the shared face normal (n1, n2, |n|, 0) [E][4][n][4]
    const double* Ja;
    const double* ijn;
    const double* fnorm;
    const double* fxy;  // boundary face-node coordinates [E][4][n][2] (curvilinear mesh with boundaries)
    const double* nxy;  // node coordinates [E][n*n][2] (curvilinear mesh with source terms)
    double* slots;      // per instance: kSlots CFL maxima, then kSlots viscous maxima (k_finalize resets them)
    // Navier-Stokes walls: per boundary face node the time-averaged viscous flux
    // of the face's orientation (momentum x, momentum y, energy), then the same
    // components at uⁿ [E][4][n][6]
    double* fvb;
    // curvilinear shock capturing: per node the subcell normals of its lower ξ
    // and η subcell faces [E][n²][4], and the calc_fn_low! normals of subcell
    // face n-1 per row (ξ) and column (η) [E][n][4]
    const double* snrm;
    const double* sqrk;
    // caller-defined Dirichlet data (bc_kind kBcTable): per element face the
    // boundary face index or -1 [E][4], and the state at every boundary face node
    // [face][n][levels][4]: levels 1 (steady) or n + 1 (GLL time levels, then t)
    const int* bidx;
    const double* btab;
    int btab_levels;
    double t_arg, dt_arg;
    int use_state_dt;  // 1: t, dt from DevState (device loop); 0: from t_arg/dt_arg
    int step_mode;     // 1: out = u + dt * rhs; 0: out = rhs
    int reduce;        // Reduce
    double bc_scale;   // scaling_factor of the Dirichlet time average
    // L2 mortars of a TreeMesh with refined cells: a face on mortar m in role r
    // (0 lower, 1 upper small element, 2 large element) has neighbour
    // -2 - (3 m + r).  Per mortar (orientation 1|2, large_side 1|2); per role and
    // face node the volume kernel's (jump state, normal flux, λ state) [M][3][n][12]
    // and the mortar kernel's fluxes [M][3][n][4]
    int n_mortars;
    const int2* mort;
    double* mrec;
    double* mflux;
    // P4estMesh (null otherwise): per element face its interface partner (element,
    // face, flipped node order, 1 if this element is the interface's primary);
    // interface faces carry neighbour code -3, so every face record goes to HBM
    const int4* plink;
};

// ---------------------------------------------------------------------------
// 1/x from the hardware reciprocal estimate and two Newton-Raphson steps
// r <- r + r (1 - x r); the result is within a few ulp of the correctly
// rounded reciprocal for the positive, normal densities it is applied to.
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Euler flux, both directions (Trixi flux(u, 1|2, CompressibleEulerEquations2D)).
// The velocities use one reciprocal of ρ instead of two divisions.
__device__ __forceinline__ void euler_fluxes(const double* u, double gm1, double* f, double* g) {
    const double rho = u[0], m1 = u[1], m2 = u[2], E = u[3];
    const double ir = rcp_nr(rho);
    const double v1 = m1 * ir, v2 = m2 * ir;
    const double p = gm1 * (E - 0.5 * (m1 * v1 + m2 * v2));
    f[0] = m1;
    f[1] = m1 * v1 + p;
    f[2] = m1 * v2;
    f[3] = (E + p) * v1;
    g[0] = m2;
    g[1] = m2 * v1;
    g[2] = m2 * v2 + p;
    g[3] = (E + p) * v2;
}

// ---------------------------------------------------------------------------
// Source terms of the LW procedure (dg_2d.jl:89-176).  The device's source is
// Trixi's source_terms_convergence_test for CompressibleEulerEquations2D, the
// residual ∂t u + ∇·f(u) of the manufactured state ρ = ρv1 = ρv2 = φ, ρe = φ²,
// φ = 2 + 0.1 sin(π(x + y - t)): with φ' = 0.1π cos(π(x + y - t)) and
// τ = (2φ - 1)(γ - 1), s = (φ', φ'(1 + τ), φ'(1 + τ), 2φ'(φ + τ)).  It depends
// on (x, t) only, so every time difference of s the reference forms at its
// stencil states is a combination of s(x, t + k dt), k = -2..2, and the order
// in which the reference passes those states does not change it.
struct SrcPt {
    double x, y, t, gamma;
};
__device__ __forceinline__ void src_eval(const SrcPt& p, double toff, double* s) {
    const double w = 3.141592653589793;
    double si, co;
    sincos_cw(w * (p.x + p.y - (p.t + toff)), &si, &co);
    const double rho = 2.0 + 0.1 * si;
    const double rho_x = w * 0.1 * co;
    const double tmp = (2.0 * rho - 1.0) * (p.gamma - 1.0);
    s[0] = rho_x;
    s[1] = rho_x * (1.0 + tmp);
    s[2] = s[1];
    s[3] = 2.0 * rho_x * (rho + tmp);
}
// dt-scaled time differences about p.t + t0 (calc_source_t_N34,
// calc_source_tt_N23, calc_source_ttt_N34, calc_source_tttt_N4)
__device__ __forceinline__ void src_t5(const SrcPt& p, double t0, double dt, double* o) {
    double a[4], b[4];
    src_eval(p, t0 + 2.0 * dt, a);
    src_eval(p, t0 + dt, b);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = 8.0 * b[v] - a[v];
    src_eval(p, t0 - dt, a);
    src_eval(p, t0 - 2.0 * dt, b);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = (1.0 / 12.0) * (o[v] - 8.0 * a[v] + b[v]);
}
__device__ __forceinline__ void src_tt3(const SrcPt& p, double dt, double* o) {
    double a[4], b[4];
    src_eval(p, dt, a);
    src_eval(p, 0.0, b);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = a[v] - 2.0 * b[v];
    src_eval(p, -dt, a);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] += a[v];
}
__device__ __forceinline__ void src_ttt5(const SrcPt& p, double dt, double* o) {
    double a[4], b[4];
    src_eval(p, 2.0 * dt, a);
    src_eval(p, dt, b);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = a[v] - 2.0 * b[v];
    src_eval(p, -dt, a);
    src_eval(p, -2.0 * dt, b);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = 0.5 * (o[v] + 2.0 * a[v] - b[v]);
}
__device__ __forceinline__ void src_tttt5(const SrcPt& p, double dt, double* o) {
    double a[4], b[4];
    src_eval(p, 2.0 * dt, a);
    src_eval(p, dt, b);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = a[v] - 4.0 * b[v];
    src_eval(p, 0.0, a);
    src_eval(p, -dt, b);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = o[v] + 6.0 * a[v] - 4.0 * b[v];
    src_eval(p, -2.0 * dt, a);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] += a[v];
}

// s_tt of calc_source_tt_N4 (five points), which the StructuredMesh kernel 4
// takes: (-s(t+2dt) + 16 s(t+dt) - 30 s(t) + 16 s(t-dt) - s(t-2dt))/12
__device__ __forceinline__ void src_tt5(const SrcPt& p, double dt, double* o) {
    double a[4], b[4];
    src_eval(p, 2.0 * dt, a);
    src_eval(p, dt, b);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = 16.0 * b[v] - a[v];
    src_eval(p, 0.0, a);
    src_eval(p, -dt, b);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = o[v] - 30.0 * a[v] + 16.0 * b[v];
    src_eval(p, -2.0 * dt, a);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = (1.0 / 12.0) * (o[v] - a[v]);
}
// Source level k of the StructuredMesh LW kernels (structured/dg_2d.jl:1440-1666,
// 1750-2058): s^(k) about t into sk (s; s_t of N34; s_tt five-point at polydeg
// 4, three-point at 3; s_ttt of N34; s_tttt of N4), and S += weight·s^(k)
// (S = s at k = 0)
__device__ __forceinline__ void src_level_c(const SrcPt& p, double dt, int k, bool five_tt, double* sk, double* S,
                                            double weight) {
    if (k == 0) src_eval(p, 0.0, sk);
    else if (k == 1) src_t5(p, 0.0, dt, sk);
    else if (k == 2 && five_tt) src_tt5(p, dt, sk);
    else if (k == 2) src_tt3(p, dt, sk);
    else if (k == 3) src_ttt5(p, dt, sk);
    else src_tttt5(p, dt, sk);
#pragma unroll
    for (int v = 0; v < 4; ++v) S[v] = k == 0 ? weight * sk[v] : fma(weight, sk[v], S[v]);
}

// The Navier-Stokes convergence source (the source_terms of
// examples/tree_2d_dgsem/elixir_navierstokes_convergence.jl:49-205): the
// residual ∂t u + ∇·f_a - ∇·f_v of ρ = 2 + ½ sin πx cos πy cos πt,
// v1 = v2 = sin πx g(y) cos πt with g = log(y + 2)(1 - e^{-(y-1)/2}), p = ρ²,
// under Trixi's viscous fluxes (Stokes' hypothesis, heat flux μκ∇T with
// κ = γ/((γ-1)Pr), T = p/ρ), by the product rule from the closed-form
// derivatives.  The node's spatial factors are formed once (NsSrcPt), each
// time level then needs sin and cos of πt only.  It does not depend on u.
struct NsSrcPt {
    double sx, cx, sy, cy, g, g1, g2, t, gamma, mu, prandtl;
};
__device__ __forceinline__ NsSrcPt ns_src_point(double x, double y, double t, const DevConst* c) {
    constexpr double pi = 3.141592653589793, A = 0.5;
    NsSrcPt p;
    sincos_cw(pi * x, &p.sx, &p.cx);
    sincos_cw(pi * y, &p.sy, &p.cy);
    const double e = exp(-A * (y - 1.0)), L = log(y + 2.0), q = 1.0 / (y + 2.0);
    p.g = L * (1.0 - e);
    p.g1 = (1.0 - e) * q + A * L * e;
    p.g2 = -(1.0 - e) * q * q + 2.0 * A * e * q - A * A * L * e;
    p.t = t;
    p.gamma = c->gamma;
    p.mu = c->visc_mu;
    p.prandtl = c->visc_Pr;
    return p;
}
__device__ __forceinline__ void ns_src_eval(const NsSrcPt& p, double toff, double* s) {
    constexpr double pi = 3.141592653589793, A = 0.5;
    double st, ct;
    sincos_cw(pi * (p.t + toff), &st, &ct);
    const double sx = p.sx, cx = p.cx, sy = p.sy, cy = p.cy;
    const double r = 2.0 + A * sx * cy * ct;
    const double r_t = -pi * A * sx * cy * st;
    const double r_x = pi * A * cx * cy * ct;
    const double r_y = -pi * A * sx * sy * ct;
    const double r_xx = -pi * pi * A * sx * cy * ct, r_yy = r_xx;
    const double v = sx * p.g * ct, v_t = -pi * sx * p.g * st;
    const double v_x = pi * cx * p.g * ct, v_y = sx * p.g1 * ct;
    const double v_xx = -pi * pi * sx * p.g * ct, v_xy = pi * cx * p.g1 * ct, v_yy = sx * p.g2 * ct;
    const double pr = r * r, p_x = 2.0 * r * r_x, p_y = 2.0 * r * r_y;
    const double p_xx = 2.0 * (r_x * r_x + r * r_xx), p_yy = 2.0 * (r_y * r_y + r * r_yy);
    const double ig = 1.0 / (p.gamma - 1.0), k2 = v * v + v * v;
    const double E = pr * ig + 0.5 * r * k2;
    const double E_t = 2.0 * r * r_t * ig + 0.5 * r_t * k2 + r * (v * v_t + v * v_t);
    const double E_x = p_x * ig + 0.5 * r_x * k2 + r * (v * v_x + v * v_x);
    const double E_y = p_y * ig + 0.5 * r_y * k2 + r * (v * v_y + v * v_y);
    const double t11 = (4.0 / 3.0) * v_x - (2.0 / 3.0) * v_y, t12 = v_y + v_x;
    const double t22 = (4.0 / 3.0) * v_y - (2.0 / 3.0) * v_x;
    const double t11_x = (4.0 / 3.0) * v_xx - (2.0 / 3.0) * v_xy, t12_x = v_xy + v_xx;
    const double t12_y = v_yy + v_xy, t22_y = (4.0 / 3.0) * v_yy - (2.0 / 3.0) * v_xy;
    const double ir3 = 1.0 / (r * r * r);
    const double T_xx = (p_xx * r * r - 2.0 * p_x * r * r_x + 2.0 * pr * r_x * r_x - pr * r * r_xx) * ir3;
    const double T_yy = (p_yy * r * r - 2.0 * p_y * r * r_y + 2.0 * pr * r_y * r_y - pr * r * r_yy) * ir3;
    const double kappa = p.gamma * ig / p.prandtl, mu = p.mu;
    s[0] = r_t + r_x * v + r * v_x + r_y * v + r * v_y;
    s[1] = r_t * v + r * v_t + r_x * v * v + 2.0 * r * v * v_x + p_x + r_y * v * v + r * v_y * v + r * v * v_y -
           mu * (t11_x + t12_y);
    s[2] = r_t * v + r * v_t + r_x * v * v + r * v_x * v + r * v * v_x + r_y * v * v + 2.0 * r * v * v_y + p_y -
           mu * (t12_x + t22_y);
    const double work_x = v_x * t11 + v * t11_x + v_x * t12 + v * t12_x;
    const double work_y = v_y * t12 + v * t12_y + v_y * t22 + v * t22_y;
    s[3] = E_t + (E_x + p_x) * v + (E + pr) * v_x + (E_y + p_y) * v + (E + pr) * v_y -
           mu * (work_x + work_y + kappa * (T_xx + T_yy));
}
// s_t of calc_source_t_N34 about t: (-s(t+2dt) + 8 s(t+dt) - 8 s(t-dt) + s(t-2dt))/12
__device__ __forceinline__ void ns_src_t5(const NsSrcPt& p, double dt, double* o) {
    double a[4];
    ns_src_eval(p, 2.0 * dt, a);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = -a[v];
    ns_src_eval(p, dt, a);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = fma(8.0, a[v], o[v]);
    ns_src_eval(p, -dt, a);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = fma(-8.0, a[v], o[v]);
    ns_src_eval(p, -2.0 * dt, a);
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = (1.0 / 12.0) * (o[v] + a[v]);
}

// The time difference of order k of the source at the node about t, as the
// reference's LW kernels form it (calc_source / _t_N34 / _tt_N4 (five_tt) or
// _tt_N23 / _ttt_N34 / _tttt_N4, dg_2d.jl:89-176).
__device__ __forceinline__ void ns_src_diff(const NsSrcPt& p, double dt, int order, bool five_tt, double* o) {
    double a[4], b[4], z[4], c[4], d[4];
    ns_src_eval(p, 0.0, z);
    if (order == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) o[v] = z[v];
        return;
    }
    ns_src_eval(p, 2.0 * dt, a);
    ns_src_eval(p, dt, b);
    ns_src_eval(p, -dt, c);
    ns_src_eval(p, -2.0 * dt, d);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        if (order == 1) o[v] = (1.0 / 12.0) * (-a[v] + 8.0 * b[v] - 8.0 * c[v] + d[v]);
        else if (order == 2 && five_tt) o[v] = (1.0 / 12.0) * (-a[v] + 16.0 * b[v] - 30.0 * z[v] + 16.0 * c[v] - d[v]);
        else if (order == 2) o[v] = b[v] - 2.0 * z[v] + c[v];
        else if (order == 3) o[v] = 0.5 * (a[v] - 2.0 * b[v] + 2.0 * c[v] - d[v]);
        else o[v] = a[v] - 4.0 * b[v] + 6.0 * z[v] - 4.0 * c[v] + d[v];
    }
}

__device__ __forceinline__ void euler_flux_dir(const double* u, double gm1, int orientation, double* f) {
    double ff[4], gg[4];
    euler_fluxes(u, gm1, ff, gg);
    for (int v = 0; v < 4; ++v) f[v] = orientation == 1 ? ff[v] : gg[v];
}

__device__ __forceinline__ double dmax(double x, double y) { return (y > x) ? y : x; }

// The CFL maxima of the node-per-lane workgroups meet in kSlots slots per
// instance by an integer atomic maximum: non-negative doubles order like their
// bit patterns, and a maximum does not depend on the order, so the reduction is
// deterministic.  NaN and 0 leave a slot as dmax would.
constexpr int kSlots = 256;
__device__ __forceinline__ void slot_max(double* slots, int inst, int which, double v) {
    if (v > 0.0)
        atomicMax(reinterpret_cast<unsigned long long*>(slots + ((size_t)inst * 2 + which) * kSlots +
                                                        (blockIdx.x & (kSlots - 1))),
                  (unsigned long long)__double_as_longlong(v));
}

// Square root of a positive normal argument: the hardware reciprocal
// square-root estimate refined by two steps of Goldschmidt's iteration and a
// final Newton correction (P. Markstein, "IA-64 and Elementary Functions",
// 2000, ch. 8), within an ulp or two of the rounded root.  Zero, negative and
// NaN arguments give NaN; callers apply it to sound speeds, whose states are
// checked for ρ > 0 and p > 0.
__device__ __forceinline__ double sqrt_gs(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = 0.5 * r;
    double e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    const double d = fma(-g, g, x);
    return fma(d, h, g);
}

// |v_n| and the sound speed: the per-side ingredients of Trixi's
// max_abs_speed_naive (λ = max(|v_n,L|,|v_n,R|) + max(c_L, c_R)).
__device__ __forceinline__ void speed_parts(const double* s, double gamma, int orientation, double& a, double& c) {
    const double ir = rcp_nr(s[0]);
    const double v1 = s[1] * ir, v2 = s[2] * ir;
    const double p = (gamma - 1.0) * (s[3] - 0.5 * (s[1] * v1 + s[2] * v2));
    a = fabs(orientation == 1 ? v1 : v2);
    c = sqrt_gs(gamma * p * ir);
}

// |v1|, |v2| and the sound speed of one state: the λ ingredients of both
// orientations with one reciprocal of ρ and one square root.
__device__ __forceinline__ void speed_parts2(const double* s, double gamma, double& a1, double& a2, double& c) {
    const double ir = rcp_nr(s[0]);
    const double v1 = s[1] * ir, v2 = s[2] * ir;
    const double p = (gamma - 1.0) * (s[3] - 0.5 * (s[1] * v1 + s[2] * v2));
    a1 = fabs(v1);
    a2 = fabs(v2);
    c = sqrt_gs(gamma * p * ir);
}



// Julia's x ≈ y for Float64 (atol = 0, rtol = sqrt(eps)): equal, or both
// finite and |x - y| <= rtol max(|x|, |y|).  isfinished (lwfr.jl:146-148).
__device__ __forceinline__ int julia_isapprox(double x, double y) {
    if (x == y) return 1;
    if (!isfinite(x) || !isfinite(y)) return 0;
    return fabs(x - y) <= 1.4901161193847656e-08 * dmax(fabs(x), fabs(y));
}

// Boundary value functions of the time-averaged Dirichlet condition.
__device__ inline void boundary_value(const DevConst* c, double x, double y, double t, double* u) {
    const double* par = c->bc_par;
    double rho, v1, v2, p, gamma;
    if (c->bc_kind == kBcConvergenceTest) {
        // Trixi's initial_condition_convergence_test: ρ = ρv1 = ρv2 = φ, ρe = φ²
        double si, co;
        sincos_cw(3.141592653589793 * (x + y - t), &si, &co);
        const double phi = 2.0 + 0.1 * si;
        u[0] = u[1] = u[2] = phi;
        u[3] = phi * phi;
        return;
    }
    if (c->bc_kind == 1) {  // isentropic vortex
        const double eps = par[0], mach = par[1], Lx = par[4], Ly = par[5];
        gamma = par[6];
        const double vinf = mach * sqrt(gamma) / sqrt(2.0);
        double dx = x - (par[2] + vinf * t), dy = y - (par[3] + vinf * t);
        dx -= Lx * floor(dx / Lx + 0.5);
        dy -= Ly * floor(dy / Ly + 0.5);
        const double r2 = dx * dx + dy * dy;
        const double pi = 3.141592653589793;
        const double du = eps / (2 * pi) * exp(0.5 * (1 - r2));
        const double T = 1 - (gamma - 1) * eps * eps / (8 * gamma * pi * pi) * exp(1 - r2);
        rho = pow(T, 1 / (gamma - 1));
        v1 = vinf - du * dy;
        v2 = vinf + du * dx;
        p = rho * T;
    } else if (c->bc_kind == 2) {  // constant state
        rho = par[0]; v1 = par[1]; v2 = par[2]; p = par[3]; gamma = par[4];
    } else {  // four-quadrant Riemann data, quadrants NE, NW, SW, SE
        gamma = par[2];
        int q;
        if (x >= par[0]) q = (y >= par[1]) ? 0 : 3;
        else q = (y >= par[1]) ? 1 : 2;
        rho = par[3 + 4 * q]; v1 = par[4 + 4 * q]; v2 = par[5 + 4 * q]; p = par[6 + 4 * q];
    }
    u[0] = rho;
    u[1] = rho * v1;
    u[2] = rho * v2;
    u[3] = p / (gamma - 1) + 0.5 * rho * (v1 * v1 + v2 * v2);
}

// Derivative contraction of one element: out[v] = Σ_k M[i,k] s_f[v][j,k] + Σ_k M[j,k] s_g[v][k,i]
// where s holds f (variables 0..3) and g (4..7) of every node of the workgroup
// with stride S between variables, and M (n x n, row-major) is in LDS.
template <int n, int S>
__device__ __forceinline__ void contract(const double* s, int base, int i, int j, const double* M, double* out) {
    double Di[n], Dj[n];
#pragma unroll
    for (int k = 0; k < n; ++k) {
        Di[k] = M[i * n + k];
        Dj[k] = M[j * n + k];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < n; ++k) acc = fma(Di[k], s[v * S + base + j * n + k], acc);
#pragma unroll
        for (int k = 0; k < n; ++k) acc = fma(Dj[k], s[(4 + v) * S + base + k * n + i], acc);
        out[v] = acc;
    }
}

template <int S> __device__ __forceinline__ void put8(double* s, int idx, const double* f, const double* g) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        s[v * S + idx] = f[v];
        s[(4 + v) * S + idx] = g[v];
    }
}

__device__ __forceinline__ void load4(const double* p, double* u) {
    const double2* p2 = reinterpret_cast<const double2*>(p);
    const double2 a = p2[0], b = p2[1];
    u[0] = a.x; u[1] = a.y; u[2] = b.x; u[3] = b.y;
}

// Boundary state of face node l of element e's face at GLL time level q (q = n:
// the time t itself): the caller's table (kBcTable) or a built-in function.
template <int n>
__device__ __forceinline__ void boundary_state(const DevConst* c, const int* bidx, const double* btab, int levels,
                                               int e, int face, int l, int q, double x, double y, double t,
                                               double* u) {
    if (c->bc_kind == kBcTable) {
        const size_t row = ((size_t)bidx[4 * (size_t)e + face] * n + l) * levels + (levels == 1 ? 0 : q);
        load4(btab + 4 * row, u);
    } else {
        boundary_value(c, x, y, t, u);
    }
}

__device__ __forceinline__ void store4(double* p, const double* u) {
    double2* p2 = reinterpret_cast<double2*>(p);
    p2[0] = make_double2(u[0], u[1]);
    p2[1] = make_double2(u[2], u[3]);
}

template <int RQ = kTraceQ> __device__ __forceinline__ void load_record(const double* p, double* r) {
    const double2* p2 = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int k = 0; k < RQ / 2; ++k) {
        const double2 x = p2[k];
        r[2 * k] = x.x;
        r[2 * k + 1] = x.y;
    }
}
template <int RQ = kTraceQ> __device__ __forceinline__ void store_record(double* p, const double* r) {
    double2* p2 = reinterpret_cast<double2*>(p);
#pragma unroll
    for (int k = 0; k < RQ / 2; ++k) p2[k] = make_double2(r[2 * k], r[2 * k + 1]);
}

// Buffer selection by integer arithmetic on the loaded pointers: a select
// between fields of the argument struct would be folded into a load from a
// computed offset, which moves the whole struct to private memory.
__device__ __forceinline__ uintptr_t pick_addr(uintptr_t p0, uintptr_t p1, bool second) {
    const uintptr_t m = (uintptr_t)0 - (uintptr_t)second;
    return p0 ^ ((p0 ^ p1) & m);
}
__device__ __forceinline__ const double* sel_in(const KArgs& a, int cur) {
    const uintptr_t b0 = (uintptr_t)a.bufs[0], b1 = (uintptr_t)a.bufs[1], ue = (uintptr_t)a.u;
    const int sel = a.in_sel;
    const bool second = (sel == SEL_CUR) != (cur == 0);
    return (const double*)pick_addr(pick_addr(b0, b1, second), ue, sel == SEL_EXPLICIT);
}
__device__ __forceinline__ double* sel_out(const KArgs& a, int cur) {
    const uintptr_t b0 = (uintptr_t)a.bufs[0], b1 = (uintptr_t)a.bufs[1], ue = (uintptr_t)a.out;
    const int sel = a.out_sel;
    const bool second = (sel == SEL_CUR) != (cur == 0);
    return (double*)pick_addr(pick_addr(b0, b1, second), ue, sel == SEL_EXPLICIT);
}

// A workgroup-table entry, read once per workgroup into scalar registers.
// bid is the workgroup's index in its own table (blockIdx.x, or the index
// within its half of a two-context launch).
__device__ __forceinline__ int4 load_blk_at(const int4* t, int bid) {
    const int4 b = t[bid];
    return make_int4(__builtin_amdgcn_readfirstlane(b.x), __builtin_amdgcn_readfirstlane(b.y),
                     __builtin_amdgcn_readfirstlane(b.z), 0);
}
__device__ __forceinline__ int4 load_blk(const int4* t) { return load_blk_at(t, blockIdx.x); }

// Rusanov flux of one face node from the two face records (left = -side):
// F* = ½(F̃n_L + F̃n_R) − ½ λ (q_R − q_L), λ = max(|v_n|) + max(c)
// (FluxPlusDissipation with flux_central and DissipationLocalLaxFriedrichs,
// src/equations/numerical_fluxes.jl:7-38).
__device__ __forceinline__ void face_numflux(const double* L, const double* R, double* Fs) {
    const double lam = dmax(L[8], R[8]) + dmax(L[9], R[9]);
#pragma unroll
    for (int v = 0; v < 4; ++v) Fs[v] = 0.5 * (L[4 + v] + R[4 + v]) + (-0.5 * lam) * (R[v] - L[v]);
}

// Trixi flux_lax_friedrichs (Rusanov with max_abs_speed_naive) of two states.
__device__ __forceinline__ void rusanov(const double* ul, const double* ur, double gamma, int orientation, double* out) {
    const double gm1 = gamma - 1.0;
    double fl[4], fr[4], al, cl, ar, cr;
    euler_flux_dir(ul, gm1, orientation, fl);
    euler_flux_dir(ur, gm1, orientation, fr);
    speed_parts(ul, gamma, orientation, al, cl);
    speed_parts(ur, gamma, orientation, ar, cr);
    const double lam = dmax(al, ar) + dmax(cl, cr);
#pragma unroll
    for (int v = 0; v < 4; ++v) out[v] = 0.5 * (fl[v] + fr[v]) + (-0.5 * lam) * (ur[v] - ul[v]);
}

__device__ __forceinline__ bool admissible(const double* u, double gamma) {
    const double rho = u[0];
    const double v1 = u[1] / rho, v2 = u[2] / rho;
    const double p = (gamma - 1.0) * (u[3] - 0.5 * (u[1] * v1 + u[2] * v2));
    return rho > 0.0 && p > 0.0;
}

// ---------------------------------------------------------------------------
// MUSCL-Hancock subcell reconstruction of a blended element (calcflux_mh!,
// src/solvers/dgsem_structured/dg_2d.jl:313-538, on the TreeMesh's Cartesian
// metric: unit normals, the element's J⁻¹ at every node).

// minmod(a, b, c, β, 0) (structured/dg_2d.jl:83-95) of finite arguments:
// s·min(β|a|, |b|, β|c|) if a, b, c have one sign, else 0
__device__ __forceinline__ double minmod3(double a, double b, double c, double beta) {
    const double m = fmin(fmin(beta * fabs(a), fabs(b)), beta * fabs(c));
    const double sa = (a > 0.0) ? 1.0 : ((a < 0.0) ? -1.0 : 0.0);
    const double sb = (b > 0.0) ? 1.0 : ((b < 0.0) ? -1.0 : 0.0);
    const double sc = (c > 0.0) ? 1.0 : ((c < 0.0) ? -1.0 : 0.0);
    return (sa != sb || sb != sc) ? 0.0 : sa * m;
}

// limit_variable_slope (equations.jl:262-287) for density (var 0) or pressure
// (var 1): when a doubled extrapolation u* drops below 1e-10, the slope is
// scaled by θ = min(|0.1 v - v| / (|v* - v| + 1e-13), 1) and u* recomputed
// with 2θ and the scaled slope.
__device__ __forceinline__ void mh_limit_var(int var, double* s, double* usl, double* usr, const double* u, double xl,
                                             double xr, double gm1) {
    auto val = [&](const double* q) {
        return var == 0 ? q[0] : gm1 * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
    };
    const double vll = val(usl), vrr = val(usr), vlow = val(u);
    if (vll < 1e-10 || vrr < 1e-10) {
        const double thr = 0.1 * vlow;
        const double rll = fabs(thr - vlow) / (fabs(vll - vlow) + 1e-13);
        const double rrr = fabs(thr - vlow) / (fabs(vrr - vlow) + 1e-13);
        const double th = fmin(fmin(rll, rrr), 1.0);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            s[v] *= th;
            usl[v] = u[v] + (2.0 * th) * (xl * s[v]);
            usr[v] = u[v] + (2.0 * th) * (xr * s[v]);
        }
    }
}

// limit_slope (compressible_euler_2d.jl:9-21): density then pressure, face values u + s x
__device__ __forceinline__ void mh_face_pair(double* s, const double* u, double xl, double xr, double gm1, double* qlo,
                                             double* qhi) {
    double usl[4], usr[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        usl[v] = u[v] + 2.0 * s[v] * xl;
        usr[v] = u[v] + 2.0 * s[v] * xr;
    }
    mh_limit_var(0, s, usl, usr, u, xl, xr, gm1);
    mh_limit_var(1, s, usl, usr, u, xl, xr, gm1);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        qlo[v] = u[v] + s[v] * xl;
        qhi[v] = u[v] + s[v] * xr;
    }
}

// Face values of node (i, j) after the Hancock half-step.  s1 holds the
// element's uⁿ (variable-major, stride S).  End nodes (ξ = ±1) keep zero
// slope: the reference's one-sided differences there are 0/0.  Half-step
// terms (all added to the four face values, in the reference's order):
// +Δt/2 J⁻¹ f(left) / hw[i] for 0 < i < n-1, -Δt/2 J⁻¹ f(right) / hw[i] for
// i < n-1, then the same in y.
template <int N, int S>
__device__ __forceinline__ void mh_faces(const double* s1, int sidx, int i, int j, const double* u, const DevConst* c,
                                         double beta, double dtmh, double invJ, double gm1, double* ql, double* qr,
                                         double* qd, double* qu) {
    constexpr int n = N + 1;
    double sx[4] = {0.0, 0.0, 0.0, 0.0}, sy[4] = {0.0, 0.0, 0.0, 0.0};
    if (i > 0 && i < n - 1) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double ul = s1[v * S + sidx - 1], ur = s1[v * S + sidx + 1];
            const double back = (u[v] - ul) / c->mh_dx1[i], fwd = (ur - u[v]) / c->mh_dx2[i];
            const double cent = fma(c->mh_fc[i], ur, fma(c->mh_fa[i], ul, c->mh_fb[i] * u[v]));
            sx[v] = minmod3(back, cent, fwd, beta);
        }
    }
    if (j > 0 && j < n - 1) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double ud = s1[v * S + sidx - n], uu = s1[v * S + sidx + n];
            const double back = (u[v] - ud) / c->mh_dx1[j], fwd = (uu - u[v]) / c->mh_dx2[j];
            const double cent = fma(c->mh_fc[j], uu, fma(c->mh_fa[j], ud, c->mh_fb[j] * u[v]));
            sy[v] = minmod3(back, cent, fwd, beta);
        }
    }
    mh_face_pair(sx, u, c->mh_xl[i], c->mh_xr[i], gm1, ql, qr);
    mh_face_pair(sy, u, c->mh_xl[j], c->mh_xr[j], gm1, qd, qu);
    double fl[4], fr[4], fd[4], fu[4];
    euler_flux_dir(ql, gm1, 1, fl);
    euler_flux_dir(qr, gm1, 1, fr);
    euler_flux_dir(qd, gm1, 2, fd);
    euler_flux_dir(qu, gm1, 2, fu);
    const double hm = (-0.5 * dtmh) * invJ, hp = (0.5 * dtmh) * invJ;
    auto add = [&](double coef, const double* f) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double r = coef * f[v];
            ql[v] += r;
            qr[v] += r;
            qd[v] += r;
            qu[v] += r;
        }
    };
    if (i > 0 && i < n - 1) add(hp / c->mh_hw[i], fl);
    if (i < n - 1) add(hm / c->mh_hw[i], fr);
    if (j > 0 && j < n - 1) add(hp / c->mh_hw[j], fd);
    if (j < n - 1) add(hm / c->mh_hw[j], fu);
}

// Shock-capturing interface flux (dg_2d.jl:1879-1946): the LW flux blended
// with the first-order Rusanov flux of uⁿ by the average blending factor of
// the two elements, falling back to the first-order flux when the low-order
// update of an adjacent node with the blended flux is not admissible.
// Records: [0..3] uⁿ, [4..7] F̃n, [8] |v_n|(Ũ), [9] c(Ũ), [10..13] fn_low, [14] α.
__device__ __forceinline__ void face_numflux_sc(const double* L, const double* R, int orientation, double gamma,
                                                double J, double dt, double w_n, double w_1, double* Fs) {
    double Fh[4], fn[4], Fb[4], low[4];
    face_numflux(L, R, Fh);
    rusanov(L, R, gamma, orientation, fn);
    double alp = 0.5 * (L[14] + R[14]);
#pragma unroll
    for (int v = 0; v < 4; ++v) Fb[v] = (1.0 - alp) * Fh[v] + alp * fn[v];
    const double cl = dt * J / (w_n * 0.5), cr = dt * J / (w_1 * 0.5);
#pragma unroll
    for (int v = 0; v < 4; ++v) low[v] = L[v] - cl * (Fb[v] - L[10 + v]);
    bool ok = admissible(low, gamma);
#pragma unroll
    for (int v = 0; v < 4; ++v) low[v] = R[v] - cr * (R[10 + v] - Fb[v]);
    ok = ok && admissible(low, gamma);
    if (!ok) alp = 1.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) Fs[v] = alp * fn[v] + (1.0 - alp) * Fh[v];
}

// Rusanov flux with a normal direction n (Trixi flux_lax_friedrichs(u_l, u_r,
// n)): Trixi's normal flux (ρv_n, ρv_n v1 + p n1, ρv_n v2 + p n2, (ρe + p) v_n)
// and λ = max|v·n| + max(c)|n|.
__device__ __forceinline__ void rusanov_normal(const double* ul, const double* ur, double gamma, double n1, double n2,
                                               double nn, double* out) {
    const double gm1 = gamma - 1.0;
    double fl[4], fr[4], vl, cl, vr, cr;
    auto side = [&](const double* q, double* f, double& vabs, double& c) {
        const double ir = rcp_nr(q[0]);
        const double v1 = q[1] * ir, v2 = q[2] * ir;
        const double p = gm1 * (q[3] - 0.5 * (q[1] * v1 + q[2] * v2));
        const double vn = fma(v1, n1, v2 * n2), rvn = q[0] * vn;
        f[0] = rvn;
        f[1] = fma(rvn, v1, p * n1);
        f[2] = fma(rvn, v2, p * n2);
        f[3] = (q[3] + p) * vn;
        vabs = fabs(vn);
        c = sqrt_gs(gamma * p * ir) * nn;
    };
    side(ul, fl, vl, cl);
    side(ur, fr, vr, cr);
    const double lam = dmax(vl, vr) + dmax(cl, cr);
#pragma unroll
    for (int v = 0; v < 4; ++v) out[v] = 0.5 * (fl[v] + fr[v]) + (-0.5 * lam) * (ur[v] - ul[v]);
}

// Shock-capturing interface flux on a curvilinear mesh (calc_interface_flux!
// and compute_alp for StructuredMesh, structured/dg_2d.jl:2169-2288).  Records
// (16 doubles): [0..3] Ũ, [4..7] F̃·n, [8..11] uⁿ, [12..15] fn_low.  The LW flux
// takes λ from uⁿ and the jump on Ũ; the low-order flux is the Rusanov flux of
// uⁿ with the same normal; the admissibility test of compute_alp uses each
// side's own node J⁻¹.
__device__ __forceinline__ void face_numflux_sc_curved(const double* L, const double* R, const double* nf, double JL,
                                                       double JR, double aL, double aR, double gamma, double dt,
                                                       double w_n, double w_1, double* Fs) {
    double fn[4], Fh[4], Fb[4], low[4];
    rusanov_normal(L + 8, R + 8, gamma, nf[0], nf[1], nf[2], fn);
    // λ of uⁿ with the normal, as in the Rusanov flux above
    double lam;
    {
        const double gm1 = gamma - 1.0;
        double va[2], ca[2];
        const double* q2[2] = {L + 8, R + 8};
        for (int k = 0; k < 2; ++k) {
            const double* q = q2[k];
            const double ir = rcp_nr(q[0]);
            const double v1 = q[1] * ir, v2 = q[2] * ir;
            const double p = gm1 * (q[3] - 0.5 * (q[1] * v1 + q[2] * v2));
            va[k] = fabs(fma(v1, nf[0], v2 * nf[1]));
            ca[k] = sqrt_gs(gamma * p * ir) * nf[2];
        }
        lam = dmax(va[0], va[1]) + dmax(ca[0], ca[1]);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) Fh[v] = 0.5 * (L[4 + v] + R[4 + v]) + (-0.5 * lam) * (R[v] - L[v]);
    double alp = 0.5 * (aL + aR);
#pragma unroll
    for (int v = 0; v < 4; ++v) Fb[v] = (1.0 - alp) * Fh[v] + alp * fn[v];
    const double cl = dt * JL / (w_n * 0.5), cr = dt * JR / (w_1 * 0.5);
#pragma unroll
    for (int v = 0; v < 4; ++v) low[v] = L[8 + v] - cl * (Fb[v] - L[12 + v]);
    bool ok = admissible(low, gamma);
#pragma unroll
    for (int v = 0; v < 4; ++v) low[v] = R[8 + v] - cr * (R[12 + v] - Fb[v]);
    ok = ok && admissible(low, gamma);
    if (!ok) alp = 1.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) Fs[v] = alp * fn[v] + (1.0 - alp) * Fh[v];
}

// Face record of a face node: jump state, normal component of F̃ and the λ
// ingredients |v_n| and c of the λ state (TreeMesh order: λ from Ũ, jump on uⁿ).
// A face node on a mortar (neighbour code -2 - (3 m + role)): the jump state,
// the normal flux and the λ state the mortar kernel interpolates or pairs.
__device__ __forceinline__ void mortar_side(const KArgs& a, int code, int n, int l, const double* jump,
                                            const double* Fn, const double* lam) {
    double* q = a.mrec + ((size_t)(-2 - code) * n + l) * 12;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        q[v] = jump[v];
        q[4 + v] = Fn[v];
        q[8 + v] = lam[v];
    }
}

__device__ __forceinline__ void make_record(double* r, const double* jump, const double* Fn, double vn_abs, double c) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        r[v] = jump[v];
        r[4 + v] = Fn[v];
    }
    r[8] = vn_abs;
    r[9] = c;
}

// Per-node maxima λ1 = |v1| + c, λ2 = |v2| + c (Trixi max_abs_speeds).
__device__ __forceinline__ void node_speeds(const double* un, double gamma, double& l1, double& l2, bool& bad) {
    const double ir = rcp_nr(un[0]);
    const double v1 = un[1] * ir, v2 = un[2] * ir;
    const double p = (gamma - 1.0) * (un[3] - 0.5 * (un[1] * v1 + un[2] * v2));
    bad = !(un[0] > 0.0) || !(p > 0.0);
    const double cs = sqrt_gs(gamma * p * ir);
    l1 = fabs(v1) + cs;
    l2 = fabs(v2) + cs;
}

// ---------------------------------------------------------------------------
// Element volume computations.  Each fills uN (state), Uc (Ũ), F, G (F̃) and
// vol (D̂F̃ contraction) for the calling node; the final F̃ exchange is in s2,
// so s0 and s1 are free once every thread has passed the last barrier.

// lw_volume_kernel_4!, dg_2d.jl:918-1267 (the temporal-error estimate at
// 1183-1199, 1251-1265 is an adaptive-mode output and is not computed).
// With KEEP, tder receives the time derivatives ut, utt, ... (4 per order),
// the intermediates an adjoint sweep needs.
// With terr, the temporal error of the node for the adaptive controller
// (dg_2d.jl:1183-1199, 1251-1265): u_np1 = u - dt J⁻¹ D F̃ and u_np1_low, the
// same with F̃ before the f_tttt term, compared variable by variable as
// ((u_np1 - u_np1_low) / (abstol + reltol max(|u_np1|, |u_np1_low|)))², summed.
// terr must be uniformly null or non-null over the workgroup (it adds barriers).
// With SRC, the source terms of the node (*sp) enter every time derivative as
// dt times the source's own difference (dg_2d.jl:980-987, 1033-1045, 1093-1103,
// 1148-1160) and Ssrc receives S = s + s_t/2 + s_tt/6 + s_ttt/24 + s_tttt/120
// (1228-1247), which the caller adds to the residual.
template <int N, int EB, bool KEEP = false, bool SRC = false>
__device__ __forceinline__ void lw4_element(const double* uN, double dt, double invJ, double gm1, const DevConst* c,
                                            int i, int j, bool act, int sidx, int sbase, double* s0, double* s1,
                                            double* s2, const double* sDm, const double* sDh, double* Uc, double* F,
                                            double* G, double* vol, double* tder = nullptr, double* terr = nullptr,
                                            double abstol = 0.0, double reltol = 0.0, const SrcPt* sp = nullptr,
                                            double* Ssrc = nullptr) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, S = G_::S;
    euler_fluxes(uN, gm1, F, G);
    if (act) put8<S>(s0, sidx, F, G);
    __syncthreads();
    double ut[4], sk[4];
    contract<n, S>(s0, sbase, i, j, sDm, ut);
    if constexpr (SRC) src_eval(*sp, 0.0, sk);
    double up[4], um[4], upp[4], umm[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        ut[v] *= invJ;
        if constexpr (SRC) {
            ut[v] = fma(dt, sk[v], ut[v]);
            Ssrc[v] = sk[v];
        }
        if constexpr (KEEP) tder[v] = ut[v];
        Uc[v] = fma(0.5, ut[v], uN[v]);
        up[v] = uN[v] + ut[v];
        um[v] = uN[v] - ut[v];
        umm[v] = fma(-2.0, ut[v], uN[v]);
        upp[v] = fma(2.0, ut[v], uN[v]);
    }
    double fa[4], ga[4], dx[4], dy[4];
    // stage 1: f_t = (-f(u+2) + 8 f(u+1) - 8 f(u-1) + f(u-2)) / 12
    euler_fluxes(upp, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = -fa[v]; dy[v] = -ga[v]; }
    euler_fluxes(up, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(8.0, fa[v], dx[v]); dy[v] = fma(8.0, ga[v], dy[v]); }
    euler_fluxes(um, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(-8.0, fa[v], dx[v]); dy[v] = fma(-8.0, ga[v], dy[v]); }
    euler_fluxes(umm, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = (1.0 / 12.0) * (dx[v] + fa[v]);
        dy[v] = (1.0 / 12.0) * (dy[v] + ga[v]);
        F[v] = fma(0.5, dx[v], F[v]);
        G[v] = fma(0.5, dy[v], G[v]);
    }
    if (act) put8<S>(s1, sidx, dx, dy);
    __syncthreads();
    double uk[4];
    contract<n, S>(s1, sbase, i, j, sDm, uk);
    if constexpr (SRC) src_t5(*sp, 0.0, dt, sk);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double w = uk[v] * invJ;  // utt
        if constexpr (SRC) {
            w = fma(dt, sk[v], w);
            Ssrc[v] = fma(0.5, sk[v], Ssrc[v]);
        }
        if constexpr (KEEP) tder[4 + v] = w;
        Uc[v] = fma(1.0 / 6.0, w, Uc[v]);
        up[v] = fma(0.5, w, up[v]);
        um[v] = fma(0.5, w, um[v]);
        upp[v] = fma(2.0, w, upp[v]);
        umm[v] = fma(2.0, w, umm[v]);
    }
    // stage 2: f_tt = (-f(u+2) + 16 f(u+1) - 30 f + 16 f(u-1) - f(u-2)) / 12
    euler_fluxes(upp, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = -fa[v]; dy[v] = -ga[v]; }
    euler_fluxes(up, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = fma(16.0, fa[v], dx[v]);
        dy[v] = fma(16.0, ga[v], dy[v]);
        dx[v] = fma(-30.0, s0[v * S + sidx], dx[v]);
        dy[v] = fma(-30.0, s0[(4 + v) * S + sidx], dy[v]);
    }
    euler_fluxes(um, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(16.0, fa[v], dx[v]); dy[v] = fma(16.0, ga[v], dy[v]); }
    euler_fluxes(umm, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = (1.0 / 12.0) * (dx[v] - fa[v]);
        dy[v] = (1.0 / 12.0) * (dy[v] - ga[v]);
        F[v] = fma(1.0 / 6.0, dx[v], F[v]);
        G[v] = fma(1.0 / 6.0, dy[v], G[v]);
    }
    if (act) put8<S>(s2, sidx, dx, dy);
    __syncthreads();
    contract<n, S>(s2, sbase, i, j, sDm, uk);
    if constexpr (SRC) src_tt3(*sp, dt, sk);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double w = uk[v] * invJ;  // uttt
        if constexpr (SRC) {
            w = fma(dt, sk[v], w);
            Ssrc[v] = fma(1.0 / 6.0, sk[v], Ssrc[v]);
        }
        if constexpr (KEEP) tder[8 + v] = w;
        Uc[v] = fma(1.0 / 24.0, w, Uc[v]);
        um[v] = fma(-1.0 / 6.0, w, um[v]);
        up[v] = fma(1.0 / 6.0, w, up[v]);
        umm[v] = fma(-4.0 / 3.0, w, umm[v]);
        upp[v] = fma(4.0 / 3.0, w, upp[v]);
    }
    // stage 3: f_ttt = (f(u+2) - 2 f(u+1) + 2 f(u-1) - f(u-2)) / 2
    euler_fluxes(upp, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fa[v]; dy[v] = ga[v]; }
    euler_fluxes(up, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(-2.0, fa[v], dx[v]); dy[v] = fma(-2.0, ga[v], dy[v]); }
    euler_fluxes(um, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(2.0, fa[v], dx[v]); dy[v] = fma(2.0, ga[v], dy[v]); }
    euler_fluxes(umm, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = 0.5 * (dx[v] - fa[v]);
        dy[v] = 0.5 * (dy[v] - ga[v]);
        F[v] = fma(1.0 / 24.0, dx[v], F[v]);
        G[v] = fma(1.0 / 24.0, dy[v], G[v]);
    }
    if (act) put8<S>(s1, sidx, dx, dy);
    __syncthreads();
    contract<n, S>(s1, sbase, i, j, sDm, uk);
    if constexpr (SRC) src_ttt5(*sp, dt, sk);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double w = uk[v] * invJ;  // utttt
        if constexpr (SRC) {
            w = fma(dt, sk[v], w);
            Ssrc[v] = fma(1.0 / 24.0, sk[v], Ssrc[v]);
        }
        if constexpr (KEEP) tder[12 + v] = w;
        Uc[v] = fma(1.0 / 120.0, w, Uc[v]);
        um[v] = fma(1.0 / 24.0, w, um[v]);
        up[v] = fma(1.0 / 24.0, w, up[v]);
        umm[v] = fma(2.0 / 3.0, w, umm[v]);
        upp[v] = fma(2.0 / 3.0, w, upp[v]);
    }
    // stage 4: f_tttt = 0.5 (f(u+2) - 4 f(u+1) + 6 f - 4 f(u-1) + f(u-2)); the
    // factor 0.5 is the reference's (dg_2d.jl:1203-1206)
    euler_fluxes(upp, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fa[v]; dy[v] = ga[v]; }
    euler_fluxes(up, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = fma(-4.0, fa[v], dx[v]);
        dy[v] = fma(-4.0, ga[v], dy[v]);
        dx[v] = fma(6.0, s0[v * S + sidx], dx[v]);
        dy[v] = fma(6.0, s0[(4 + v) * S + sidx], dy[v]);
    }
    euler_fluxes(um, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(-4.0, fa[v], dx[v]); dy[v] = fma(-4.0, ga[v], dy[v]); }
    euler_fluxes(umm, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = 0.5 * (dx[v] + fa[v]);
        dy[v] = 0.5 * (dy[v] + ga[v]);
        F[v] = fma(1.0 / 120.0, dx[v], F[v]);
        G[v] = fma(1.0 / 120.0, dy[v], G[v]);
    }
    if (act) put8<S>(s2, sidx, F, G);
    // s0 is read only at the own node since the first contraction: it takes f_tttt
    if (terr && act) put8<S>(s0, sidx, dx, dy);
    __syncthreads();
    contract<n, S>(s2, sbase, i, j, sDh, vol);
    if constexpr (SRC) {
        src_tttt5(*sp, dt, sk);
#pragma unroll
        for (int v = 0; v < 4; ++v) Ssrc[v] = fma(1.0 / 120.0, sk[v], Ssrc[v]);
    }
    if (terr) {
        double a1[4], a2[4];
        contract<n, S>(s2, sbase, i, j, sDm, a1);  // -dt D F̃
        contract<n, S>(s0, sbase, i, j, sDm, a2);  // -dt D f_tttt
        // u_np1 - u_np1_low is formed as the contraction of the f_tttt term itself,
        // not as a difference of the two updates (which cancels to ~1e-7 relative)
        double e = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double d = invJ * ((1.0 / 120.0) * a2[v]);
            const double hi = fma(invJ, a1[v], uN[v]);
            const double lo = hi - d;
            const double r = d / (abstol + reltol * dmax(fabs(hi), fabs(lo)));
            e = fma(r, r, e);
        }
        *terr = e;
        __syncthreads();  // s0 is reused for the face records
    }
}

// lw_volume_kernel_3!, dg_2d.jl:674-916
// With KEEP, tder receives the time derivatives ut, utt, ... (4 per order),
// the intermediates an adjoint sweep needs.  With SRC, source terms as in
// lw4_element (732-740, 785-797, 841-851, 891-911): S = s + s_t/2 + s_tt/6 +
// s_ttt/24.
template <int N, int EB, bool KEEP = false, bool SRC = false>
__device__ __forceinline__ void lw3_element(const double* uN, double dt, double invJ, double gm1, const DevConst* c,
                                            int i, int j, bool act, int sidx, int sbase, double* s0, double* s1,
                                            double* s2, const double* sDm, const double* sDh, double* Uc, double* F,
                                            double* G, double* vol, double* tder = nullptr,
                                            const SrcPt* sp = nullptr, double* Ssrc = nullptr) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, S = G_::S;
    euler_fluxes(uN, gm1, F, G);
    if (act) put8<S>(s0, sidx, F, G);
    __syncthreads();
    double ut[4], sk[4];
    contract<n, S>(s0, sbase, i, j, sDm, ut);
    if constexpr (SRC) src_eval(*sp, 0.0, sk);
    double up[4], um[4], upp[4], umm[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        ut[v] *= invJ;
        if constexpr (SRC) {
            ut[v] = fma(dt, sk[v], ut[v]);
            Ssrc[v] = sk[v];
        }
        if constexpr (KEEP) tder[v] = ut[v];
        Uc[v] = fma(0.5, ut[v], uN[v]);
        up[v] = uN[v] + ut[v];
        um[v] = uN[v] - ut[v];
        umm[v] = fma(-2.0, ut[v], uN[v]);
        upp[v] = fma(2.0, ut[v], uN[v]);
    }
    double fa[4], ga[4], dx[4], dy[4];
    euler_fluxes(upp, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = -fa[v]; dy[v] = -ga[v]; }
    euler_fluxes(up, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(8.0, fa[v], dx[v]); dy[v] = fma(8.0, ga[v], dy[v]); }
    euler_fluxes(um, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(-8.0, fa[v], dx[v]); dy[v] = fma(-8.0, ga[v], dy[v]); }
    euler_fluxes(umm, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = (1.0 / 12.0) * (dx[v] + fa[v]);
        dy[v] = (1.0 / 12.0) * (dy[v] + ga[v]);
        F[v] = fma(0.5, dx[v], F[v]);
        G[v] = fma(0.5, dy[v], G[v]);
    }
    if (act) put8<S>(s1, sidx, dx, dy);
    __syncthreads();
    double uk[4];
    contract<n, S>(s1, sbase, i, j, sDm, uk);
    if constexpr (SRC) src_t5(*sp, 0.0, dt, sk);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double w = uk[v] * invJ;  // utt
        if constexpr (SRC) {
            w = fma(dt, sk[v], w);
            Ssrc[v] = fma(0.5, sk[v], Ssrc[v]);
        }
        if constexpr (KEEP) tder[4 + v] = w;
        Uc[v] = fma(1.0 / 6.0, w, Uc[v]);
        up[v] = fma(0.5, w, up[v]);
        um[v] = fma(0.5, w, um[v]);
        upp[v] = fma(2.0, w, upp[v]);
        umm[v] = fma(2.0, w, umm[v]);
    }
    // f_tt = f(u+1) - 2 f + f(u-1)
    euler_fluxes(up, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = fma(-2.0, s0[v * S + sidx], fa[v]);
        dy[v] = fma(-2.0, s0[(4 + v) * S + sidx], ga[v]);
    }
    euler_fluxes(um, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = dx[v] + fa[v];
        dy[v] = dy[v] + ga[v];
        F[v] = fma(1.0 / 6.0, dx[v], F[v]);
        G[v] = fma(1.0 / 6.0, dy[v], G[v]);
    }
    if (act) put8<S>(s2, sidx, dx, dy);
    __syncthreads();
    contract<n, S>(s2, sbase, i, j, sDm, uk);
    if constexpr (SRC) src_tt3(*sp, dt, sk);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double w = uk[v] * invJ;  // uttt
        if constexpr (SRC) {
            w = fma(dt, sk[v], w);
            Ssrc[v] = fma(1.0 / 6.0, sk[v], Ssrc[v]);
        }
        if constexpr (KEEP) tder[8 + v] = w;
        Uc[v] = fma(1.0 / 24.0, w, Uc[v]);
        um[v] = fma(-1.0 / 6.0, w, um[v]);
        up[v] = fma(1.0 / 6.0, w, up[v]);
        umm[v] = fma(-4.0 / 3.0, w, umm[v]);
        upp[v] = fma(4.0 / 3.0, w, upp[v]);
    }
    // f_ttt = (f(u+2) - 2 f(u+1) + 2 f(u-1) - f(u-2)) / 2
    euler_fluxes(upp, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fa[v]; dy[v] = ga[v]; }
    euler_fluxes(up, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(-2.0, fa[v], dx[v]); dy[v] = fma(-2.0, ga[v], dy[v]); }
    euler_fluxes(um, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(2.0, fa[v], dx[v]); dy[v] = fma(2.0, ga[v], dy[v]); }
    euler_fluxes(umm, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = 0.5 * (dx[v] - fa[v]);
        dy[v] = 0.5 * (dy[v] - ga[v]);
        F[v] = fma(1.0 / 24.0, dx[v], F[v]);
        G[v] = fma(1.0 / 24.0, dy[v], G[v]);
    }
    __syncthreads();  // every thread has finished reading s2 before it is reused
    if (act) put8<S>(s2, sidx, F, G);
    __syncthreads();
    contract<n, S>(s2, sbase, i, j, sDh, vol);
    if constexpr (SRC) {
        src_ttt5(*sp, dt, sk);
#pragma unroll
        for (int v = 0; v < 4; ++v) Ssrc[v] = fma(1.0 / 24.0, sk[v], Ssrc[v]);
    }
}

// f_t of the two-level Taylor states u ± ut, u ± 2ut:
// (-f(u+2) + 8 f(u+1) - 8 f(u-1) + f(u-2)) / 12
__device__ __forceinline__ void mdrk_ft(const double* u, const double* ut, double gm1, double* dx, double* dy) {
    double w[4], fa[4], ga[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) w[v] = fma(2.0, ut[v], u[v]);
    euler_fluxes(w, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = -fa[v]; dy[v] = -ga[v]; w[v] = u[v] + ut[v]; }
    euler_fluxes(w, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(8.0, fa[v], dx[v]); dy[v] = fma(8.0, ga[v], dy[v]); w[v] = u[v] - ut[v]; }
    euler_fluxes(w, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = fma(-8.0, fa[v], dx[v]);
        dy[v] = fma(-8.0, ga[v], dy[v]);
        w[v] = fma(-2.0, ut[v], u[v]);
    }
    euler_fluxes(w, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = (1.0 / 12.0) * (dx[v] + fa[v]);
        dy[v] = (1.0 / 12.0) * (dy[v] + ga[v]);
    }
}

// mdrk_kernel_1!, tree/multi_derivative_2d.jl:263-420: F̃ = f/2 + f_t/8,
// Ũ = u/2 + ut/8.  F2 = f + f_t/6 and U2 = u + ut/6 are recomputed by the
// stage-2 kernel instead of being stored.  With SRC (323-379): ut gets
// dt s(t), and Ssrc = s/2 + s_t/8.
template <int N, int EB, bool SRC = false>
__device__ __forceinline__ void mdrk1_element(const double* uN, double dt, double invJ, double gm1, const DevConst* c,
                                              int i, int j, bool act, int sidx, int sbase, double* s0, double* s1,
                                              double* s2, const double* sDm, const double* sDh, double* Uc,
                                              double* F, double* G, double* vol, const SrcPt* sp = nullptr,
                                              double* Ssrc = nullptr) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, S = G_::S;
    double f[4], g[4];
    euler_fluxes(uN, gm1, f, g);
    if (act) put8<S>(s0, sidx, f, g);
    __syncthreads();
    double ut[4];
    contract<n, S>(s0, sbase, i, j, sDm, ut);
    double sk[4];
    if constexpr (SRC) src_eval(*sp, 0.0, sk);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        ut[v] *= invJ;
        if constexpr (SRC) {
            ut[v] = fma(dt, sk[v], ut[v]);
            Ssrc[v] = 0.5 * sk[v];
        }
    }
    if constexpr (SRC) {
        src_t5(*sp, 0.0, dt, sk);
#pragma unroll
        for (int v = 0; v < 4; ++v) Ssrc[v] = fma(0.125, sk[v], Ssrc[v]);
    }
    double dx[4], dy[4];
    mdrk_ft(uN, ut, gm1, dx, dy);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F[v] = fma(0.125, dx[v], 0.5 * f[v]);
        G[v] = fma(0.125, dy[v], 0.5 * g[v]);
        Uc[v] = fma(0.125, ut[v], 0.5 * uN[v]);
    }
    if (act) put8<S>(s2, sidx, F, G);
    __syncthreads();
    contract<n, S>(s2, sbase, i, j, sDh, vol);
}

// mdrk_kernel_2!, tree/multi_derivative_2d.jl:423-555: F̃ = F2 + f_t(us)/3,
// Ũ = U2 + ust/3 with F2, U2 recomputed from uⁿ.  With SRC (487-522): the
// recomputed stage-1 ut gets dt s(t) and S2 = s + s_t/6; ust gets
// dt s(t + dt/2), and Ssrc = S2 + s_t(t + dt/2)/3.
template <int N, int EB, bool SRC = false>
__device__ __forceinline__ void mdrk2_element(const double* uN, const double* usn, double dt, double invJ, double gm1,
                                              const DevConst* c, int i, int j, bool act, int sidx, int sbase,
                                              double* s0, double* s1, double* s2, const double* sDm,
                                              const double* sDh, double* Uc, double* F, double* G, double* vol,
                                              const SrcPt* sp = nullptr, double* Ssrc = nullptr) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, S = G_::S;
    double f[4], g[4], fs[4], gs[4];
    euler_fluxes(uN, gm1, f, g);
    euler_fluxes(usn, gm1, fs, gs);
    if (act) {
        put8<S>(s0, sidx, f, g);
        put8<S>(s1, sidx, fs, gs);
    }
    __syncthreads();
    double ut[4], ust[4];
    contract<n, S>(s0, sbase, i, j, sDm, ut);
    contract<n, S>(s1, sbase, i, j, sDm, ust);
#pragma unroll
    for (int v = 0; v < 4; ++v) { ut[v] *= invJ; ust[v] *= invJ; }
    if constexpr (SRC) {
        double sk[4], sh[4];
        src_eval(*sp, 0.0, sk);
        src_eval(*sp, 0.5 * dt, sh);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            ut[v] = fma(dt, sk[v], ut[v]);
            ust[v] = fma(dt, sh[v], ust[v]);
            Ssrc[v] = sk[v];
        }
        src_t5(*sp, 0.0, dt, sk);
        src_t5(*sp, 0.5 * dt, dt, sh);
#pragma unroll
        for (int v = 0; v < 4; ++v) Ssrc[v] = fma(1.0 / 3.0, sh[v], fma(1.0 / 6.0, sk[v], Ssrc[v]));
    }
    double dx[4], dy[4];
    mdrk_ft(uN, ut, gm1, dx, dy);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F[v] = fma(1.0 / 6.0, dx[v], f[v]);   // F2
        G[v] = fma(1.0 / 6.0, dy[v], g[v]);
        Uc[v] = fma(1.0 / 6.0, ut[v], uN[v]);  // U2
    }
    mdrk_ft(usn, ust, gm1, dx, dy);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F[v] = fma(1.0 / 3.0, dx[v], F[v]);
        G[v] = fma(1.0 / 3.0, dy[v], G[v]);
        Uc[v] = fma(1.0 / 3.0, ust[v], Uc[v]);
    }
    if (act) put8<S>(s2, sidx, F, G);
    __syncthreads();
    contract<n, S>(s2, sbase, i, j, sDh, vol);
}

// ---------------------------------------------------------------------------
// Hyperbolic-parabolic element procedures (compressible Navier-Stokes with
// conservative gradient variables).  The BR1 gradient of the input state is
// formed in the volume workgroup itself (weak-form D̂ volume term, central
// face value ½(u + u_neighbour) lifted by 1/w at the end nodes, ×J⁻¹), so no
// gradient array goes through HBM.  The LW procedure then runs on the total
// flux f_a(u) - f_v(u, ∇u) with Taylor states of u and of ∇u, the time
// derivatives' gradients being element-local (J⁻¹ D ∂ₜᵏu).  Only the total
// F̃ = F̃a - F̃v is kept: the interface flux of the reference,
// LW-Rusanov(F̃a) - ½(F̃v_L + F̃v_R), equals ½(F̃_L + F̃_R) - ½λ(u_R - u_L),
// so the face records and the surface kernel are those of the Euler path.
// Reference: lw_volume_kernel_3!/4!, dg_2d_parabolic.jl:725-1860;
// calc_interface_flux_hyperbolic_parabolic!, dg_2d_parabolic.jl:1911-1948;
// mdrk_kernel_1!/2!, multi_derivative_2d_parabolic.jl:221-711.

// Separate x and y contractions of the 4 variables in s (stride S):
// ox[v] = Σ_k M[i,k] s[v][j,k], oy[v] = Σ_k M[j,k] s[v][k,i].
template <int n, int S>
__device__ __forceinline__ void contract_xy(const double* s, int base, int i, int j, const double* M, double* ox,
                                            double* oy) {
    double Mi[n], Mj[n];
#pragma unroll
    for (int k = 0; k < n; ++k) {
        Mi[k] = M[i * n + k];
        Mj[k] = M[j * n + k];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double ax = 0.0, ay = 0.0;
#pragma unroll
        for (int k = 0; k < n; ++k) {
            ax = fma(Mi[k], s[v * S + base + j * n + k], ax);
            ay = fma(Mj[k], s[v * S + base + k * n + i], ay);
        }
        ox[v] = ax;
        oy[v] = ay;
    }
}

template <int S> __device__ __forceinline__ void put4(double* s, int idx, const double* w) {
#pragma unroll
    for (int v = 0; v < 4; ++v) s[v * S + idx] = w[v];
}

// Viscous fluxes of CompressibleNavierStokesDiffusion2D from the conservative
// state and conservative gradients (convert_derivative_to_primitive,
// compressible_navier_stokes_2d.jl:37-57; Trixi flux(u, ∇u, orientation)).
__device__ __forceinline__ void ns_visc_fluxes(const double* u, const double* gx, const double* gy, double gm1,
                                               const DevConst* c, double* fx, double* fy) {
    const double ir = rcp_nr(u[0]);
    const double v1 = u[1] * ir, v2 = u[2] * ir;
    const double p = gm1 * (u[3] - 0.5 * (u[1] * v1 + u[2] * v2));
    const double T = p * ir;
    const double vv = v1 * v1 + v2 * v2;
    auto prim = [&](const double* W, double& v1d, double& v2d, double& Td) {
        const double rd = W[0];
        v1d = (W[1] - rd * v1) * ir;
        v2d = (W[2] - rd * v2) * ir;
        const double pd = (W[3] - 0.5 * rd * vv - u[0] * (v1 * v1d + v2 * v2d)) * gm1;
        Td = (pd - rd * T) * ir;
    };
    double v1x, v2x, Tx, v1y, v2y, Ty;
    prim(gx, v1x, v2x, Tx);
    prim(gy, v1y, v2y, Ty);
    const double mu = c->visc_mu, kappa = c->visc_kappa;
    const double t11 = (4.0 / 3.0) * v1x - (2.0 / 3.0) * v2y;
    const double t12 = v1y + v2x;
    const double t22 = (4.0 / 3.0) * v2y - (2.0 / 3.0) * v1x;
    fx[0] = 0.0;
    fx[1] = t11 * mu;
    fx[2] = t12 * mu;
    fx[3] = (v1 * t11 + v2 * t12 + kappa * Tx) * mu;
    fy[0] = 0.0;
    fy[1] = t12 * mu;
    fy[2] = t22 * mu;
    fy[3] = (v1 * t12 + v2 * t22 + kappa * Ty) * mu;
}

// f = f_a(u) - f_v(u, ∇u)
__device__ __forceinline__ void ns_total_fluxes(const double* u, const double* gx, const double* gy, double gm1,
                                                const DevConst* c, double* f, double* g) {
    double fx[4], fy[4];
    euler_fluxes(u, gm1, f, g);
    ns_visc_fluxes(u, gx, gy, gm1, c, fx, fy);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        f[v] -= fx[v];
        g[v] -= fy[v];
    }
}

// Viscous time-step bound of one node, J⁻² max(4μ/(3ρ), μκ/(ρ Pr))
// (compressible_navier_stokes_2d.jl:230-237), with IEEE divisions.
__device__ __forceinline__ double lam_visc(const double* u, double invJ, const DevConst* c) {
    const double mu = c->visc_mu, rho = u[0];
    return invJ * invJ * dmax(4.0 * mu / (3.0 * rho), mu * c->visc_kappa / (rho * c->visc_Pr));
}
// The same with the node's J⁻¹ of a curvilinear mesh in place of J⁻² (the 2-D
// Jacobian already scales as h²)
__device__ __forceinline__ double lam_visc_curved(const double* u, double invJ, const DevConst* c) {
    const double mu = c->visc_mu, rho = u[0];
    return invJ * dmax(4.0 * mu / (3.0 * rho), mu * c->visc_kappa / (rho * c->visc_Pr));
}

// Contravariant fluxes of a Cartesian pair: Ja¹·(f, g), Ja²·(f, g)
// (contravariant_flux, structured/dg_2d.jl:550-556).
__device__ __forceinline__ void to_contravariant(const double* J, const double* f, const double* g, double* fc,
                                                 double* gc) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        fc[v] = fma(J[0], f[v], J[1] * g[v]);
        gc[v] = fma(J[2], f[v], J[3] * g[v]);
    }
}

// Reference-space gradient (w_ξ, w_η) → Cartesian with the transposed
// contravariant vectors J = (Ja11, Ja12, Ja21, Ja22) and the node's J⁻¹
// (compute_local_grad_u!, p4est/dg_2d_parabolic.jl:51-104).
__device__ __forceinline__ void grad_to_cartesian(const double* J, double invJ, double* gx, double* gy) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double g1 = gx[v], g2 = gy[v];
        gx[v] = fma(J[0], g1, J[2] * g2) * invJ;
        gy[v] = fma(J[1], g1, J[3] * g2) * invJ;
    }
}

// BR1 gradient of the node's state: s holds the element's state (put4 before
// the barrier), src the state array the face neighbours are read from.
template <int N, int EB, bool CURV = false>
__device__ __forceinline__ void br1_gradient(const double* s, const double* src, const int* nbr, const DevConst* c,
                                             int e, int i, int j, int sbase, double invJ, const double* sDh,
                                             const double* uown, double* gx, double* gy, const double* Jn = nullptr) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, NN = G_::NN, S = G_::S;
    contract_xy<n, S>(s, sbase, i, j, sDh, gx, gy);
    // at a no-slip wall the interface value is the wall state prim2cons(ρ, v_wall, p)
    // of the node (Trixi calc_boundary_flux_by_direction_gradient! with
    // compressible_navier_stokes_2d.jl:113-124)
    // the gradient's boundary value at a no-slip wall: prim2cons(ρ, v_wall, p) with the
    // node's pressure (Adiabatic) or p = ρ T_wall (Isothermal, the reference's R = 1;
    // compressible_navier_stokes_2d.jl:93-101)
    auto wall_state = [&](int f, double* uo) {
        const double vw1 = c->wall_v[2 * f], vw2 = c->wall_v[2 * f + 1];
        const double rho = uown[0];
        const double p = c->wall_iso[f] ? rho * c->wall_T[f]
                                        : (c->gamma - 1.0) * (uown[3] - 0.5 * (uown[1] * uown[1] + uown[2] * uown[2]) / rho);
        uo[0] = rho;
        uo[1] = rho * vw1;
        uo[2] = rho * vw2;
        uo[3] = p / (c->gamma - 1.0) + 0.5 * rho * (vw1 * vw1 + vw2 * vw2);
    };
    if (i == 0 || i == n - 1) {
        const int f = (i == 0) ? 0 : 1;
        const int nb = nbr[4 * e + f];
        const double wf = f ? c->bnd_hi : -c->bnd_lo;
        if (nb >= 0) {
            double un[4];
            load4(src + ((size_t)nb * NN + j * n + (n - 1 - i)) * 4, un);
#pragma unroll
            for (int v = 0; v < 4; ++v) gx[v] = fma(0.5 * (uown[v] + un[v]), wf, gx[v]);
        } else if (c->ns_walls) {
            double uo[4];
            wall_state(f, uo);
#pragma unroll
            for (int v = 0; v < 4; ++v) gx[v] = fma(uo[v], wf, gx[v]);
        }
    }
    if (j == 0 || j == n - 1) {
        const int f = (j == 0) ? 2 : 3;
        const int nb = nbr[4 * e + f];
        const double wf = (f & 1) ? c->bnd_hi : -c->bnd_lo;
        if (nb >= 0) {
            double un[4];
            load4(src + ((size_t)nb * NN + (n - 1 - j) * n + i) * 4, un);
#pragma unroll
            for (int v = 0; v < 4; ++v) gy[v] = fma(0.5 * (uown[v] + un[v]), wf, gy[v]);
        } else if (c->ns_walls) {
            double uo[4];
            wall_state(f, uo);
#pragma unroll
            for (int v = 0; v < 4; ++v) gy[v] = fma(uo[v], wf, gy[v]);
        }
    }
    if constexpr (CURV) {
        // curvilinear BR1 (Trixi calc_gradient! for P4estMesh): the reference-space
        // volume and lifted interface terms above, turned with the transposed
        // metric; the face terms of the element's own outward normal ±Ja are the
        // same linear map
        grad_to_cartesian(Jn, invJ, gx, gy);
        return;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        gx[v] *= invJ;
        gy[v] *= invJ;
    }
}

// Taylor states of u and ∇u: q = 0..3 for u+, u-, u++, u--.
struct ParStates {
    double u[4][4], x[4][4], y[4][4];
    __device__ __forceinline__ void init(const double* w, const double* wx, const double* wy) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                u[q][v] = w[v];
                x[q][v] = wx[v];
                y[q][v] = wy[v];
            }
    }
    // states += c[q] · (w, wx, wy)
    __device__ __forceinline__ void add(double c0, double c1, double c2, double c3, const double* w,
                                        const double* wx, const double* wy) {
        const double cq[4] = {c0, c1, c2, c3};
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                u[q][v] = fma(cq[q], w[v], u[q][v]);
                x[q][v] = fma(cq[q], wx[v], x[q][v]);
                y[q][v] = fma(cq[q], wy[v], y[q][v]);
            }
    }
    __device__ __forceinline__ void flux(int q, double gm1, const DevConst* c, double* f, double* g) const {
        ns_total_fluxes(u[q], x[q], y[q], gm1, c, f, g);
    }
    // momentum components of the viscous fluxes of state q: (fv1, fv2, gv1, gv2)
    // components 1..3 of the viscous fluxes of state q: (fx1, fx2, fx3, fy1, fy2, fy3)
    __device__ __forceinline__ void visc6(int q, double gm1, const DevConst* c, double* out) const {
        double fx[4], fy[4];
        ns_visc_fluxes(u[q], x[q], y[q], gm1, c, fx, fy);
        out[0] = fx[1];
        out[1] = fx[2];
        out[2] = fx[3];
        out[3] = fy[1];
        out[4] = fy[2];
        out[5] = fy[3];
    }
};

// Time derivative of the next level and its element-local gradient:
// w = J⁻¹ (-dt D) · (dx, dy) (+ add, the level's source term dt·s) through sa,
// then (wx, wy) = J⁻¹ D w through sg.
template <int N, int EB, bool CURV = false>
__device__ __forceinline__ void par_level(bool act, int sidx, int sbase, int i, int j, double invJ, double* sa,
                                          double* sg, const double* sDm, const double* sD, const double* dx,
                                          const double* dy, double* w, double* wx, double* wy,
                                          const double* Jn = nullptr, const double* add = nullptr) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, S = G_::S;
    if constexpr (CURV) {
        // contravariant pair for the derivative contraction
        double cx[4], cy[4];
        to_contravariant(Jn, dx, dy, cx, cy);
        if (act) put8<S>(sa, sidx, cx, cy);
    } else {
        if (act) put8<S>(sa, sidx, dx, dy);
    }
    __syncthreads();
    contract<n, S>(sa, sbase, i, j, sDm, w);
#pragma unroll
    for (int v = 0; v < 4; ++v) w[v] *= invJ;
    if (add) {
#pragma unroll
        for (int v = 0; v < 4; ++v) w[v] += add[v];
    }
    if (act) put4<S>(sg, sidx, w);
    __syncthreads();
    contract_xy<n, S>(sg, sbase, i, j, sD, wx, wy);
    if constexpr (CURV) {
        grad_to_cartesian(Jn, invJ, wx, wy);
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            wx[v] *= invJ;
            wy[v] *= invJ;
        }
    }
}

// lw_volume_kernel_3!/4! with viscous terms (dg_2d_parabolic.jl:725-1189,
// 1191-1860): the Euler procedure of lw3_element/lw4_element on the total
// flux, with the states of ∇u advanced by the same Taylor coefficients.
// Buffer use: s0 keeps f(uⁿ) for the 5- and 3-point stencils, s1 takes each
// flux derivative, s2 each time derivative; the final F̃ exchange is in s2.
template <int N, int EB, bool CURV = false, bool WALL = false>
__device__ __forceinline__ void lwp_element(const double* uN, const double* gxN, const double* gyN, double invJ,
                                            double gm1, const DevConst* c, int i, int j, bool act, int sidx,
                                            int sbase, double* s0, double* s1, double* s2, const double* sDm,
                                            const double* sD, const double* sDh, double* Uc, double* F, double* G,
                                            double* vol, const double* Jn = nullptr, double* const* wq = nullptr,
                                            double* terr = nullptr, double abstol = 0.0, double reltol = 0.0,
                                            bool has_src = false, const NsSrcPt& sp = NsSrcPt{},
                                            double* Ssrc = nullptr, double dtv = 0.0) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, S = G_::S;
    ns_total_fluxes(uN, gxN, gyN, gm1, c, F, G);
    // the Navier-Stokes convergence source (the wall instantiation, has_src;
    // dg_2d_parabolic.jl:725-1860): level k of the Taylor procedure gets dt·s^(k)
    // before its gradient is formed, S the same weights as F̃ (s^(k) the time
    // difference of ns_src_diff; kernel 4 takes the five-point s_tt)
    auto src_lvl = [&](int order, double weight, double* ds) {
        if constexpr (WALL) {
            double v4[4] = {0.0, 0.0, 0.0, 0.0};
            if (has_src) ns_src_diff(sp, dtv, order, N == 4, v4);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (ds) ds[v] = dtv * v4[v];
                if (has_src) Ssrc[v] = order == 0 ? weight * v4[v] : fma(weight, v4[v], Ssrc[v]);
            }
        }
    };
    // with walls: the time average of the viscous flux of a wall face node in the
    // face's orientation, alongside F̃, accumulated in the node's wall record
    // wq[o] (o = 0: its x face, 1: its y face; null where that face is no wall):
    // [time average (3), value at uⁿ (3)].  Only wall lanes touch memory here.
    if constexpr (WALL) {
        double fx[4], fy[4];
        ns_visc_fluxes(uN, gxN, gyN, gm1, c, fx, fy);
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            double* q = wq[o];
            if (!q) continue;
            const double* f = o ? fy : fx;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                q[k] = f[1 + k];
                q[3 + k] = f[1 + k];
            }
        }
    }
    // one stage of the wall average: weight · scale · (cv v(uⁿ) + Σ c_s v(state s)),
    // the states u+, u-, u++, u-- one at a time, both orientations from one
    // viscous-flux evaluation; only wall lanes enter
    auto wall_stage = [&](const ParStates& Ps, double c0, double c1, double c2, double c3, double cv, double scale,
                          double weight, bool outer) {
        if constexpr (WALL) {
            if (!wq[0] && !wq[1]) return;
            double d[6], a6[6];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                d[k] = wq[0] ? cv * wq[0][3 + k] : 0.0;
                d[3 + k] = wq[1] ? cv * wq[1][3 + k] : 0.0;
            }
            Ps.visc6(0, gm1, c, a6);
#pragma unroll
            for (int k = 0; k < 6; ++k) d[k] = fma(c0, a6[k], d[k]);
            Ps.visc6(1, gm1, c, a6);
#pragma unroll
            for (int k = 0; k < 6; ++k) d[k] = fma(c1, a6[k], d[k]);
            if (outer) {
                Ps.visc6(2, gm1, c, a6);
#pragma unroll
                for (int k = 0; k < 6; ++k) d[k] = fma(c2, a6[k], d[k]);
                Ps.visc6(3, gm1, c, a6);
#pragma unroll
                for (int k = 0; k < 6; ++k) d[k] = fma(c3, a6[k], d[k]);
            }
#pragma unroll
            for (int o = 0; o < 2; ++o) {
                double* q = wq[o];
                if (!q) continue;
#pragma unroll
                for (int k = 0; k < 3; ++k) q[k] = fma(weight, scale * d[3 * o + k], q[k]);
            }
        }
    };
    double w[4], wx[4], wy[4];
    if constexpr (WALL) {
        double ds[4];
        src_lvl(0, 1.0, ds);
        par_level<N, EB, CURV>(act, sidx, sbase, i, j, invJ, s0, s2, sDm, sD, F, G, w, wx, wy, Jn, ds);  // ut
    } else {
        par_level<N, EB, CURV>(act, sidx, sbase, i, j, invJ, s0, s2, sDm, sD, F, G, w, wx, wy, Jn);  // ut
    }
    // on a curvilinear mesh s0 took the contravariant pair; the centred stencils
    // read the node's own Cartesian f(uⁿ), written once every lane has passed
    // the contraction of s0
    if constexpr (CURV) {
        if (act) put8<S>(s0, sidx, F, G);
    }
    ParStates P;
    P.init(uN, gxN, gyN);
    P.add(1.0, -1.0, 2.0, -2.0, w, wx, wy);
#pragma unroll
    for (int v = 0; v < 4; ++v) Uc[v] = fma(0.5, w[v], uN[v]);
    double fa[4], ga[4], dx[4], dy[4];
    // stage 1: f_t = (-f(u++) + 8 f(u+) - 8 f(u-) + f(u--)) / 12
    P.flux(2, gm1, c, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = -fa[v]; dy[v] = -ga[v]; }
    P.flux(0, gm1, c, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(8.0, fa[v], dx[v]); dy[v] = fma(8.0, ga[v], dy[v]); }
    P.flux(1, gm1, c, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(-8.0, fa[v], dx[v]); dy[v] = fma(-8.0, ga[v], dy[v]); }
    P.flux(3, gm1, c, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = (1.0 / 12.0) * (dx[v] + fa[v]);
        dy[v] = (1.0 / 12.0) * (dy[v] + ga[v]);
        F[v] = fma(0.5, dx[v], F[v]);
        G[v] = fma(0.5, dy[v], G[v]);
    }
    // f_t of the viscous part, weight 1/2
    wall_stage(P, 8.0, -8.0, -1.0, 1.0, 0.0, 1.0 / 12.0, 0.5, true);
    if constexpr (WALL) {
        double ds[4];
        src_lvl(1, 0.5, ds);
        par_level<N, EB, CURV>(act, sidx, sbase, i, j, invJ, s1, s2, sDm, sD, dx, dy, w, wx, wy, Jn, ds);  // utt
    } else {
        par_level<N, EB, CURV>(act, sidx, sbase, i, j, invJ, s1, s2, sDm, sD, dx, dy, w, wx, wy, Jn);  // utt
    }
    P.add(0.5, 0.5, 2.0, 2.0, w, wx, wy);
#pragma unroll
    for (int v = 0; v < 4; ++v) Uc[v] = fma(1.0 / 6.0, w[v], Uc[v]);
    if constexpr (N == 4) {
        // stage 2: f_tt = (-f(u++) + 16 f(u+) - 30 f + 16 f(u-) - f(u--)) / 12
        P.flux(2, gm1, c, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) { dx[v] = -fa[v]; dy[v] = -ga[v]; }
        P.flux(0, gm1, c, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dx[v] = fma(16.0, fa[v], dx[v]);
            dy[v] = fma(16.0, ga[v], dy[v]);
            dx[v] = fma(-30.0, s0[v * S + sidx], dx[v]);
            dy[v] = fma(-30.0, s0[(4 + v) * S + sidx], dy[v]);
        }
        P.flux(1, gm1, c, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) { dx[v] = fma(16.0, fa[v], dx[v]); dy[v] = fma(16.0, ga[v], dy[v]); }
        P.flux(3, gm1, c, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dx[v] = (1.0 / 12.0) * (dx[v] - fa[v]);
            dy[v] = (1.0 / 12.0) * (dy[v] - ga[v]);
        }
    } else {
        // stage 2: f_tt = f(u+) - 2 f + f(u-)
        P.flux(0, gm1, c, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dx[v] = fma(-2.0, s0[v * S + sidx], fa[v]);
            dy[v] = fma(-2.0, s0[(4 + v) * S + sidx], ga[v]);
        }
        P.flux(1, gm1, c, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) { dx[v] += fa[v]; dy[v] += ga[v]; }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F[v] = fma(1.0 / 6.0, dx[v], F[v]);
        G[v] = fma(1.0 / 6.0, dy[v], G[v]);
    }
    // f_tt of the viscous part, weight 1/6 (five points at N = 4, three at N = 3)
    if constexpr (N == 4) wall_stage(P, 16.0, 16.0, -1.0, -1.0, -30.0, 1.0 / 12.0, 1.0 / 6.0, true);
    else wall_stage(P, 1.0, 1.0, 0.0, 0.0, -2.0, 1.0, 1.0 / 6.0, false);
    if constexpr (WALL) {
        double ds[4];
        src_lvl(2, 1.0 / 6.0, ds);
        par_level<N, EB, CURV>(act, sidx, sbase, i, j, invJ, s1, s2, sDm, sD, dx, dy, w, wx, wy, Jn, ds);  // uttt
    } else {
        par_level<N, EB, CURV>(act, sidx, sbase, i, j, invJ, s1, s2, sDm, sD, dx, dy, w, wx, wy, Jn);  // uttt
    }
    P.add(1.0 / 6.0, -1.0 / 6.0, 4.0 / 3.0, -4.0 / 3.0, w, wx, wy);
#pragma unroll
    for (int v = 0; v < 4; ++v) Uc[v] = fma(1.0 / 24.0, w[v], Uc[v]);
    // stage 3: f_ttt = (f(u++) - 2 f(u+) + 2 f(u-) - f(u--)) / 2
    P.flux(2, gm1, c, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fa[v]; dy[v] = ga[v]; }
    P.flux(0, gm1, c, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(-2.0, fa[v], dx[v]); dy[v] = fma(-2.0, ga[v], dy[v]); }
    P.flux(1, gm1, c, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(2.0, fa[v], dx[v]); dy[v] = fma(2.0, ga[v], dy[v]); }
    P.flux(3, gm1, c, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = 0.5 * (dx[v] - fa[v]);
        dy[v] = 0.5 * (dy[v] - ga[v]);
        F[v] = fma(1.0 / 24.0, dx[v], F[v]);
        G[v] = fma(1.0 / 24.0, dy[v], G[v]);
    }
    // f_ttt of the viscous part, weight 1/24
    wall_stage(P, -2.0, 2.0, 1.0, -1.0, 0.0, 0.5, 1.0 / 24.0, true);
    if constexpr (N == 4) {
        if constexpr (WALL) {
            double ds[4];
            src_lvl(3, 1.0 / 24.0, ds);
            par_level<N, EB, CURV>(act, sidx, sbase, i, j, invJ, s1, s2, sDm, sD, dx, dy, w, wx, wy, Jn, ds);
        } else {
            par_level<N, EB, CURV>(act, sidx, sbase, i, j, invJ, s1, s2, sDm, sD, dx, dy, w, wx, wy, Jn);  // utttt
        }
        P.add(1.0 / 24.0, 1.0 / 24.0, 2.0 / 3.0, 2.0 / 3.0, w, wx, wy);
#pragma unroll
        for (int v = 0; v < 4; ++v) Uc[v] = fma(1.0 / 120.0, w[v], Uc[v]);
        // stage 4: f_tttt = 0.5 (f(u++) - 4 f(u+) + 6 f - 4 f(u-) + f(u--)), with
        // the reference's factor 0.5 (dg_2d_parabolic.jl:1746-1747)
        P.flux(2, gm1, c, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) { dx[v] = fa[v]; dy[v] = ga[v]; }
        P.flux(0, gm1, c, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dx[v] = fma(-4.0, fa[v], dx[v]);
            dy[v] = fma(-4.0, ga[v], dy[v]);
            dx[v] = fma(6.0, s0[v * S + sidx], dx[v]);
            dy[v] = fma(6.0, s0[(4 + v) * S + sidx], dy[v]);
        }
        P.flux(1, gm1, c, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) { dx[v] = fma(-4.0, fa[v], dx[v]); dy[v] = fma(-4.0, ga[v], dy[v]); }
        P.flux(3, gm1, c, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dx[v] = 0.5 * (dx[v] + fa[v]);
            dy[v] = 0.5 * (dy[v] + ga[v]);
            F[v] = fma(1.0 / 120.0, dx[v], F[v]);
            G[v] = fma(1.0 / 120.0, dy[v], G[v]);
        }
        // s0 is read only at the own node since the first contraction: it takes
        // the total f_tttt for the temporal error
        if (terr && act) put8<S>(s0, sidx, dx, dy);
    }
    // f_tttt of the viscous part (with the 0.5), weight 1/120
    if constexpr (N == 4) wall_stage(P, -4.0, -4.0, 1.0, 1.0, 6.0, 0.5, 1.0 / 120.0, true);
    // the source's last difference enters S only: s_tttt/120 (kernel 4), s_ttt/24 (kernel 3)
    if constexpr (N == 4) src_lvl(4, 1.0 / 120.0, nullptr);
    else src_lvl(3, 1.0 / 24.0, nullptr);
    __syncthreads();  // every thread has finished reading s2 before it is reused
    if constexpr (CURV) {
        double Fc[4], Gc[4];
        to_contravariant(Jn, F, G, Fc, Gc);
        if (act) put8<S>(s2, sidx, Fc, Gc);
    } else {
        if (act) put8<S>(s2, sidx, F, G);
    }
    __syncthreads();
    contract<n, S>(s2, sbase, i, j, sDh, vol);
    if (terr) {
        // temporal error of kernel 4 (dg_2d_parabolic.jl:1749-1775, 1841-1855): as in
        // lw4_element, with the total F̃ and f_tttt
        double a1[4], a2[4];
        contract<n, S>(s2, sbase, i, j, sDm, a1);  // -dt D F̃
        contract<n, S>(s0, sbase, i, j, sDm, a2);  // -dt D f_tttt
        double e = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double d = invJ * ((1.0 / 120.0) * a2[v]);
            const double hi = fma(invJ, a1[v], uN[v]);
            const double lo = hi - d;
            const double r = d / (abstol + reltol * dmax(fabs(hi), fabs(lo)));
            e = fma(r, r, e);
        }
        *terr = e;
        __syncthreads();  // s0 is reused for the face records
    }
}

// f_t of the states w ± wt, w ± 2wt (and their gradients) for the MDRK stages;
// WALL: also f_t of components 1..3 of the viscous fluxes in v6 (x, then y),
// from the same evaluations
template <bool WALL = false>
__device__ __forceinline__ void mdrkp_ft(const double* u, const double* ut, const double* gx, const double* gtx,
                                         const double* gy, const double* gty, double gm1, const DevConst* c,
                                         double* dx, double* dy, double* v6 = nullptr) {
    const double cq[4] = {2.0, 1.0, -1.0, -2.0};
    const double wq[4] = {-1.0, 8.0, -8.0, 1.0};
    double a6[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 4; ++v) dx[v] = dy[v] = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double s[4], sx[4], sy[4], fa[4], ga[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            s[v] = fma(cq[q], ut[v], u[v]);
            sx[v] = fma(cq[q], gtx[v], gx[v]);
            sy[v] = fma(cq[q], gty[v], gy[v]);
        }
        if constexpr (WALL) {
            double fx[4], fy[4];
            euler_fluxes(s, gm1, fa, ga);
            ns_visc_fluxes(s, sx, sy, gm1, c, fx, fy);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                fa[v] -= fx[v];
                ga[v] -= fy[v];
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                a6[k] = fma(wq[q], fx[1 + k], a6[k]);
                a6[3 + k] = fma(wq[q], fy[1 + k], a6[3 + k]);
            }
        } else {
            ns_total_fluxes(s, sx, sy, gm1, c, fa, ga);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dx[v] = fma(wq[q], fa[v], dx[v]);
            dy[v] = fma(wq[q], ga[v], dy[v]);
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] *= 1.0 / 12.0;
        dy[v] *= 1.0 / 12.0;
    }
    if constexpr (WALL) {
#pragma unroll
        for (int k = 0; k < 6; ++k) v6[k] = a6[k] * (1.0 / 12.0);
    }
}

// No-slip walls: the wall records of a face node, wq[o] (o = 0: its x face,
// 1: its y face; null where that face is no wall), take the time average of
// components 1..3 of the viscous flux in the face's orientation: q = w·v6
// (first term) or q += w·v6.  Only wall lanes touch memory.
__device__ __forceinline__ void wall_acc(double* const* wq, double w, const double* v6, bool first) {
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        double* q = wq[o];
        if (!q) continue;
#pragma unroll
        for (int k = 0; k < 3; ++k) q[k] = first ? w * v6[3 * o + k] : fma(w, v6[3 * o + k], q[k]);
    }
}
// components 1..3 of the viscous fluxes of a wall face node's state into its
// wall records with weight w (the first term of the average)
__device__ __forceinline__ void wall_first(double* const* wq, double w, const double* u, const double* gx,
                                           const double* gy, double gm1, const DevConst* c) {
    if (!wq[0] && !wq[1]) return;
    double fx[4], fy[4], v6[6];
    ns_visc_fluxes(u, gx, gy, gm1, c, fx, fy);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v6[k] = fx[1 + k];
        v6[3 + k] = fy[1 + k];
    }
    wall_acc(wq, w, v6, true);
}

// mdrk_kernel_1! with viscous terms (multi_derivative_2d_parabolic.jl:221-464):
// F̃ = f/2 + f_t/8, Ũ = u/2 + ut/8 on the total flux.  WALL: the wall records
// of the node take the stage's viscous average f_v/2 + f_v,t/8 (the Fv the
// reference's prolong2boundaries_visc_lw! reads).  With the Navier-Stokes
// convergence source (has_src; 290-300, 380-392, 441-446): ut += dt s(t) before
// its gradient, and Ssrc = s/2 + s_t/8 for the residual.
template <int N, int EB, bool CURV = false, bool WALL = false>
__device__ __forceinline__ void mdrkp1_element(const double* uN, const double* gxN, const double* gyN, double invJ,
                                               double gm1, const DevConst* c, int i, int j, bool act, int sidx,
                                               int sbase, double* s0, double* s1, double* s2, const double* sDm,
                                               const double* sD, const double* sDh, double* Uc, double* F,
                                               double* G, double* vol, const double* Jn = nullptr,
                                               double* const* wq = nullptr, bool has_src = false,
                                               const NsSrcPt& sp = NsSrcPt{}, double dt = 0.0,
                                               double* Ssrc = nullptr) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, S = G_::S;
    double f[4], g[4];
    ns_total_fluxes(uN, gxN, gyN, gm1, c, f, g);
    if constexpr (WALL) wall_first(wq, 0.5, uN, gxN, gyN, gm1, c);
    double ut[4], utx[4], uty[4];
    if (has_src) {
        double s[4], st[4], ds[4];
        ns_src_eval(sp, 0.0, s);
        ns_src_t5(sp, dt, st);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            ds[v] = dt * s[v];
            Ssrc[v] = fma(0.125, st[v], 0.5 * s[v]);
        }
        par_level<N, EB, CURV>(act, sidx, sbase, i, j, invJ, s0, s2, sDm, sD, f, g, ut, utx, uty, Jn, ds);
    } else {
        par_level<N, EB, CURV>(act, sidx, sbase, i, j, invJ, s0, s2, sDm, sD, f, g, ut, utx, uty, Jn);
    }
    double dx[4], dy[4];
    if constexpr (WALL) {
        double v6[6];
        mdrkp_ft<true>(uN, ut, gxN, utx, gyN, uty, gm1, c, dx, dy, v6);
        wall_acc(wq, 0.125, v6, false);
    } else {
        mdrkp_ft(uN, ut, gxN, utx, gyN, uty, gm1, c, dx, dy);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F[v] = fma(0.125, dx[v], 0.5 * f[v]);
        G[v] = fma(0.125, dy[v], 0.5 * g[v]);
        Uc[v] = fma(0.125, ut[v], 0.5 * uN[v]);
    }
    __syncthreads();  // every thread has finished reading s2 before it is reused
    if constexpr (CURV) {
        double Fc[4], Gc[4];
        to_contravariant(Jn, F, G, Fc, Gc);
        if (act) put8<S>(s2, sidx, Fc, Gc);
    } else {
        if (act) put8<S>(s2, sidx, F, G);
    }
    __syncthreads();
    contract<n, S>(s2, sbase, i, j, sDh, vol);
}

// mdrk_kernel_2! with viscous terms (multi_derivative_2d_parabolic.jl:506-711):
// F̃ = F2 + f_t(us)/3, Ũ = U2 + ust/3, with F2 = f + f_t/6 and U2 = u + ut/6
// recomputed from uⁿ and its BR1 gradient.  WALL: the wall records take the
// stage's viscous average f_v + f_v,t/6 + f_v,t(us)/3.  With the Navier-Stokes
// convergence source (has_src; 576-601, 668-674): ut and ust both += dt s(t)
// before their gradients (the reference evaluates stage 2's source at the
// call's t), Ssrc = S2 + s_t/3 with S2 = s + s_t/6.
template <int N, int EB, bool CURV = false, bool WALL = false>
__device__ __forceinline__ void mdrkp2_element(const double* uN, const double* gxN, const double* gyN,
                                               const double* usn, const double* gxS, const double* gyS,
                                               double invJ, double gm1, const DevConst* c, int i, int j, bool act,
                                               int sidx, int sbase, double* s0, double* s1, double* s2,
                                               const double* sDm, const double* sD, const double* sDh, double* Uc,
                                               double* F, double* G, double* vol, const double* Jn = nullptr,
                                               double* const* wq = nullptr, bool has_src = false,
                                               const NsSrcPt& sp = NsSrcPt{}, double dt = 0.0,
                                               double* Ssrc = nullptr) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, S = G_::S;
    double f[4], g[4], fs[4], gs[4];
    ns_total_fluxes(uN, gxN, gyN, gm1, c, f, g);
    if constexpr (WALL) wall_first(wq, 1.0, uN, gxN, gyN, gm1, c);
    ns_total_fluxes(usn, gxS, gyS, gm1, c, fs, gs);
    if constexpr (CURV) {
        double cx[4], cy[4];
        to_contravariant(Jn, f, g, cx, cy);
        if (act) put8<S>(s0, sidx, cx, cy);
        to_contravariant(Jn, fs, gs, cx, cy);
        if (act) put8<S>(s1, sidx, cx, cy);
    } else if (act) {
        put8<S>(s0, sidx, f, g);
        put8<S>(s1, sidx, fs, gs);
    }
    __syncthreads();
    double ut[4], ust[4];
    contract<n, S>(s0, sbase, i, j, sDm, ut);
    contract<n, S>(s1, sbase, i, j, sDm, ust);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        ut[v] *= invJ;
        ust[v] *= invJ;
    }
    if (has_src) {
        double s[4], st[4];
        ns_src_eval(sp, 0.0, s);
        ns_src_t5(sp, dt, st);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            ut[v] = fma(dt, s[v], ut[v]);
            ust[v] = fma(dt, s[v], ust[v]);
            Ssrc[v] = fma(1.0 / 3.0, st[v], fma(1.0 / 6.0, st[v], s[v]));
        }
    }
    if (act) {
        put4<S>(s2, sidx, ut);
        put4<S>(s2 + 4 * S, sidx, ust);
    }
    __syncthreads();
    double utx[4], uty[4], stx[4], sty[4];
    contract_xy<n, S>(s2, sbase, i, j, sD, utx, uty);
    contract_xy<n, S>(s2 + 4 * S, sbase, i, j, sD, stx, sty);
    if constexpr (CURV) {
        grad_to_cartesian(Jn, invJ, utx, uty);
        grad_to_cartesian(Jn, invJ, stx, sty);
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            utx[v] *= invJ;
            uty[v] *= invJ;
            stx[v] *= invJ;
            sty[v] *= invJ;
        }
    }
    double dx[4], dy[4];
    if constexpr (WALL) {
        double v6[6];
        mdrkp_ft<true>(uN, ut, gxN, utx, gyN, uty, gm1, c, dx, dy, v6);
        wall_acc(wq, 1.0 / 6.0, v6, false);
    } else {
        mdrkp_ft(uN, ut, gxN, utx, gyN, uty, gm1, c, dx, dy);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F[v] = fma(1.0 / 6.0, dx[v], f[v]);
        G[v] = fma(1.0 / 6.0, dy[v], g[v]);
        Uc[v] = fma(1.0 / 6.0, ut[v], uN[v]);
    }
    if constexpr (WALL) {
        double v6[6];
        mdrkp_ft<true>(usn, ust, gxS, stx, gyS, sty, gm1, c, dx, dy, v6);
        wall_acc(wq, 1.0 / 3.0, v6, false);
    } else {
        mdrkp_ft(usn, ust, gxS, stx, gyS, sty, gm1, c, dx, dy);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F[v] = fma(1.0 / 3.0, dx[v], F[v]);
        G[v] = fma(1.0 / 3.0, dy[v], G[v]);
        Uc[v] = fma(1.0 / 3.0, ust[v], Uc[v]);
    }
    __syncthreads();  // every thread has finished reading s2 before it is reused
    if constexpr (CURV) {
        double Fc[4], Gc[4];
        to_contravariant(Jn, F, G, Fc, Gc);
        if (act) put8<S>(s2, sidx, Fc, Gc);
    } else {
        if (act) put8<S>(s2, sidx, F, G);
    }
    __syncthreads();
    contract<n, S>(s2, sbase, i, j, sDh, vol);
}

// ---------------------------------------------------------------------------
// Curvilinear StructuredMesh (lw_volume_kernel_3!/4! for StructuredMesh,
// src/solvers/dgsem_structured/dg_2d.jl:1379-2105): the LW procedure with the
// node's contravariant vectors J = (Ja11, Ja12, Ja21, Ja22) and inverse
// Jacobian.  The flux time derivatives are formed for the Cartesian fluxes
// (their time averages F̃, G̃ feed the face records, F̃·n) and turned into
// contravariant ones, Ja¹·(f, g) and Ja²·(f, g), for the derivative
// contractions; the volume term is D̂ of the contravariant time average.
// f(uⁿ) is kept in registers (f0x, f0y) for the centred stencils.

// SRC: with the source terms of the StructuredMesh kernels (src_level_c; sp the
// node's point, Ssrc the S of the residual, dtv the step).
template <int N, int EB, bool SRC = false>
__device__ __forceinline__ void lwc_element(const double* uN, const double* J, double invJ, double gm1, int i, int j,
                                            bool act, int sidx, int sbase, double* s0, double* s1, double* s2,
                                            const double* sDm, const double* sDh, double* Uc, double* F, double* G,
                                            double* vol, double* terr = nullptr, double abstol = 0.0,
                                            double reltol = 0.0, const SrcPt* sp = nullptr, double* Ssrc = nullptr,
                                            double dtv = 0.0) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, S = G_::S;
    double f0x[4], f0y[4], cx[4], cy[4];
    euler_fluxes(uN, gm1, f0x, f0y);
#pragma unroll
    for (int v = 0; v < 4; ++v) { F[v] = f0x[v]; G[v] = f0y[v]; }
    to_contravariant(J, f0x, f0y, cx, cy);
    if (act) put8<S>(s0, sidx, cx, cy);
    __syncthreads();
    double w[4];
    contract<n, S>(s0, sbase, i, j, sDm, w);
    double up[4], um[4], upp[4], umm[4];
    double sk[4] = {0.0, 0.0, 0.0, 0.0};
    if constexpr (SRC) src_level_c(*sp, dtv, 0, N == 4, sk, Ssrc, 1.0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        w[v] *= invJ;  // ut
        if constexpr (SRC) w[v] = fma(dtv, sk[v], w[v]);
        Uc[v] = fma(0.5, w[v], uN[v]);
        up[v] = uN[v] + w[v];
        um[v] = uN[v] - w[v];
        upp[v] = fma(2.0, w[v], uN[v]);
        umm[v] = fma(-2.0, w[v], uN[v]);
    }
    double fa[4], ga[4], dx[4], dy[4];
    // stage 1: f_t = (-f(u++) + 8 f(u+) - 8 f(u-) + f(u--)) / 12
    euler_fluxes(upp, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = -fa[v]; dy[v] = -ga[v]; }
    euler_fluxes(up, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(8.0, fa[v], dx[v]); dy[v] = fma(8.0, ga[v], dy[v]); }
    euler_fluxes(um, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(-8.0, fa[v], dx[v]); dy[v] = fma(-8.0, ga[v], dy[v]); }
    euler_fluxes(umm, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = (1.0 / 12.0) * (dx[v] + fa[v]);
        dy[v] = (1.0 / 12.0) * (dy[v] + ga[v]);
        F[v] = fma(0.5, dx[v], F[v]);
        G[v] = fma(0.5, dy[v], G[v]);
    }
    to_contravariant(J, dx, dy, cx, cy);
    if (act) put8<S>(s1, sidx, cx, cy);
    __syncthreads();
    contract<n, S>(s1, sbase, i, j, sDm, w);
    if constexpr (SRC) src_level_c(*sp, dtv, 1, N == 4, sk, Ssrc, 0.5);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        w[v] *= invJ;  // utt
        if constexpr (SRC) w[v] = fma(dtv, sk[v], w[v]);
        Uc[v] = fma(1.0 / 6.0, w[v], Uc[v]);
        up[v] = fma(0.5, w[v], up[v]);
        um[v] = fma(0.5, w[v], um[v]);
        upp[v] = fma(2.0, w[v], upp[v]);
        umm[v] = fma(2.0, w[v], umm[v]);
    }
    if constexpr (N == 4) {
        // stage 2: f_tt = (-f(u++) + 16 f(u+) - 30 f + 16 f(u-) - f(u--)) / 12
        euler_fluxes(upp, gm1, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) { dx[v] = -fa[v]; dy[v] = -ga[v]; }
        euler_fluxes(up, gm1, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dx[v] = fma(16.0, fa[v], dx[v]);
            dy[v] = fma(16.0, ga[v], dy[v]);
            dx[v] = fma(-30.0, f0x[v], dx[v]);
            dy[v] = fma(-30.0, f0y[v], dy[v]);
        }
        euler_fluxes(um, gm1, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) { dx[v] = fma(16.0, fa[v], dx[v]); dy[v] = fma(16.0, ga[v], dy[v]); }
        euler_fluxes(umm, gm1, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dx[v] = (1.0 / 12.0) * (dx[v] - fa[v]);
            dy[v] = (1.0 / 12.0) * (dy[v] - ga[v]);
        }
    } else {
        // stage 2: f_tt = f(u+) - 2 f + f(u-)
        euler_fluxes(up, gm1, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dx[v] = fma(-2.0, f0x[v], fa[v]);
            dy[v] = fma(-2.0, f0y[v], ga[v]);
        }
        euler_fluxes(um, gm1, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) { dx[v] += fa[v]; dy[v] += ga[v]; }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F[v] = fma(1.0 / 6.0, dx[v], F[v]);
        G[v] = fma(1.0 / 6.0, dy[v], G[v]);
    }
    to_contravariant(J, dx, dy, cx, cy);
    if (act) put8<S>(s2, sidx, cx, cy);
    __syncthreads();
    contract<n, S>(s2, sbase, i, j, sDm, w);
    if constexpr (SRC) src_level_c(*sp, dtv, 2, N == 4, sk, Ssrc, 1.0 / 6.0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        w[v] *= invJ;  // uttt
        if constexpr (SRC) w[v] = fma(dtv, sk[v], w[v]);
        Uc[v] = fma(1.0 / 24.0, w[v], Uc[v]);
        um[v] = fma(-1.0 / 6.0, w[v], um[v]);
        up[v] = fma(1.0 / 6.0, w[v], up[v]);
        umm[v] = fma(-4.0 / 3.0, w[v], umm[v]);
        upp[v] = fma(4.0 / 3.0, w[v], upp[v]);
    }
    // stage 3: f_ttt = (f(u++) - 2 f(u+) + 2 f(u-) - f(u--)) / 2
    euler_fluxes(upp, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fa[v]; dy[v] = ga[v]; }
    euler_fluxes(up, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(-2.0, fa[v], dx[v]); dy[v] = fma(-2.0, ga[v], dy[v]); }
    euler_fluxes(um, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dx[v] = fma(2.0, fa[v], dx[v]); dy[v] = fma(2.0, ga[v], dy[v]); }
    euler_fluxes(umm, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = 0.5 * (dx[v] - fa[v]);
        dy[v] = 0.5 * (dy[v] - ga[v]);
        F[v] = fma(1.0 / 24.0, dx[v], F[v]);
        G[v] = fma(1.0 / 24.0, dy[v], G[v]);
    }
    if constexpr (N == 4) {
        to_contravariant(J, dx, dy, cx, cy);
        if (act) put8<S>(s1, sidx, cx, cy);
        __syncthreads();
        contract<n, S>(s1, sbase, i, j, sDm, w);
        if constexpr (SRC) src_level_c(*sp, dtv, 3, true, sk, Ssrc, 1.0 / 24.0);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            w[v] *= invJ;  // utttt
            if constexpr (SRC) w[v] = fma(dtv, sk[v], w[v]);
            Uc[v] = fma(1.0 / 120.0, w[v], Uc[v]);
            um[v] = fma(1.0 / 24.0, w[v], um[v]);
            up[v] = fma(1.0 / 24.0, w[v], up[v]);
            umm[v] = fma(2.0 / 3.0, w[v], umm[v]);
            upp[v] = fma(2.0 / 3.0, w[v], upp[v]);
        }
        // stage 4: f_tttt = 0.5 (f(u++) - 4 f(u+) + 6 f - 4 f(u-) + f(u--)), with the
        // reference's factor 0.5 (dg_2d.jl:2018-2021)
        euler_fluxes(upp, gm1, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) { dx[v] = fa[v]; dy[v] = ga[v]; }
        euler_fluxes(up, gm1, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dx[v] = fma(-4.0, fa[v], dx[v]);
            dy[v] = fma(-4.0, ga[v], dy[v]);
            dx[v] = fma(6.0, f0x[v], dx[v]);
            dy[v] = fma(6.0, f0y[v], dy[v]);
        }
        euler_fluxes(um, gm1, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) { dx[v] = fma(-4.0, fa[v], dx[v]); dy[v] = fma(-4.0, ga[v], dy[v]); }
        euler_fluxes(umm, gm1, fa, ga);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dx[v] = 0.5 * (dx[v] + fa[v]);
            dy[v] = 0.5 * (dy[v] + ga[v]);
            F[v] = fma(1.0 / 120.0, dx[v], F[v]);
            G[v] = fma(1.0 / 120.0, dy[v], G[v]);
        }
    }
    // the source's last difference enters S only: s_tttt/120 (kernel 4), s_ttt/24 (kernel 3)
    if constexpr (SRC) src_level_c(*sp, dtv, N == 4 ? 4 : 3, N == 4, sk, Ssrc, N == 4 ? 1.0 / 120.0 : 1.0 / 24.0);
    to_contravariant(J, F, G, cx, cy);
    // polydeg 3 read s2 in its last contraction: every thread has finished
    // before it is reused (polydeg 4 passed a barrier since)
    if constexpr (N == 3) __syncthreads();
    if (act) put8<S>(s2, sidx, cx, cy);
    // the last stage's contravariant term for the temporal error; s0 is read
    // only by the first contraction
    if (terr) {
        double tx[4], ty[4];
        to_contravariant(J, dx, dy, tx, ty);
        if (act) put8<S>(s0, sidx, tx, ty);
    }
    __syncthreads();
    contract<n, S>(s2, sbase, i, j, sDh, vol);
    if (terr) {
        // temporal error of the node (structured dg_2d.jl:1595-1681 at N=3, 1983-2077 at N=4):
        // u_np1 = u + J⁻¹(-dt D)·F̃c, u_np1_low without the last stage's term, whose update
        // itself is the numerator
        constexpr double wlast = (N == 4) ? 1.0 / 120.0 : 1.0 / 24.0;
        double a1[4], a2[4];
        contract<n, S>(s2, sbase, i, j, sDm, a1);
        contract<n, S>(s0, sbase, i, j, sDm, a2);
        double e = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double d = invJ * (wlast * a2[v]);
            const double hi = fma(invJ, a1[v], uN[v]);
            const double lo = hi - d;
            const double r = d / (abstol + reltol * dmax(fabs(hi), fabs(lo)));
            e = fma(r, r, e);
        }
        *terr = e;
        __syncthreads();  // s0 is reused for the face records
    }
}

// mdrk_kernel_1! for StructuredMesh (structured/multi_derivative_2d.jl:82-243):
// F̃ = f/2 + f_t/8 (Cartesian, for the records), its contravariant image for
// the D̂ contraction, Ũ = u/2 + ut/8, ut from the contravariant fluxes.  SRC
// (146-155, 179-200): ut += dt s, Ssrc = s/2 + s_t/8.
template <int N, int EB, bool SRC = false>
__device__ __forceinline__ void mdrkc1_element(const double* uN, const double* J, double invJ, double gm1, int i,
                                               int j, bool act, int sidx, int sbase, double* s0, double* s2,
                                               const double* sDm, const double* sDh, double* Uc, double* F,
                                               double* G, double* vol, const SrcPt* sp = nullptr,
                                               double* Ssrc = nullptr, double dtv = 0.0) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, S = G_::S;
    double f[4], g[4], cx[4], cy[4];
    euler_fluxes(uN, gm1, f, g);
    to_contravariant(J, f, g, cx, cy);
    if (act) put8<S>(s0, sidx, cx, cy);
    __syncthreads();
    double ut[4];
    contract<n, S>(s0, sbase, i, j, sDm, ut);
#pragma unroll
    for (int v = 0; v < 4; ++v) ut[v] *= invJ;
    if constexpr (SRC) {
        double s[4], st[4];
        src_eval(*sp, 0.0, s);
        src_t5(*sp, 0.0, dtv, st);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            ut[v] = fma(dtv, s[v], ut[v]);
            Ssrc[v] = fma(0.125, st[v], 0.5 * s[v]);
        }
    }
    double dx[4], dy[4];
    mdrk_ft(uN, ut, gm1, dx, dy);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F[v] = fma(0.125, dx[v], 0.5 * f[v]);
        G[v] = fma(0.125, dy[v], 0.5 * g[v]);
        Uc[v] = fma(0.125, ut[v], 0.5 * uN[v]);
    }
    to_contravariant(J, F, G, cx, cy);
    if (act) put8<S>(s2, sidx, cx, cy);
    __syncthreads();
    contract<n, S>(s2, sbase, i, j, sDh, vol);
}

// mdrk_kernel_2! for StructuredMesh (structured/multi_derivative_2d.jl:245-445):
// F̃ = F2 + f_t(us)/3, Ũ = U2 + ust/3, with F2 = f + f_t/6 and U2 = u + ut/6
// recomputed from uⁿ instead of being stored by stage 1.  SRC (326-339,
// 358-385): ut and ust += dt s(t); Ssrc = s + s_t/3 about t, the reference's
// stage 2 overwriting the S2 it loads (stage 1's s + s_t/6) with s.
template <int N, int EB, bool SRC = false>
__device__ __forceinline__ void mdrkc2_element(const double* uN, const double* usn, const double* J, double invJ,
                                               double gm1, int i, int j, bool act, int sidx, int sbase, double* s0,
                                               double* s1, double* s2, const double* sDm, const double* sDh,
                                               double* Uc, double* F, double* G, double* vol,
                                               const SrcPt* sp = nullptr, double* Ssrc = nullptr, double dtv = 0.0) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, S = G_::S;
    double f[4], g[4], cx[4], cy[4];
    euler_fluxes(uN, gm1, f, g);
    to_contravariant(J, f, g, cx, cy);
    if (act) put8<S>(s0, sidx, cx, cy);
    euler_fluxes(usn, gm1, F, G);
    to_contravariant(J, F, G, cx, cy);
    if (act) put8<S>(s1, sidx, cx, cy);
    __syncthreads();
    double ut[4], ust[4];
    contract<n, S>(s0, sbase, i, j, sDm, ut);
    contract<n, S>(s1, sbase, i, j, sDm, ust);
#pragma unroll
    for (int v = 0; v < 4; ++v) { ut[v] *= invJ; ust[v] *= invJ; }
    if constexpr (SRC) {
        double s[4], st[4];
        src_eval(*sp, 0.0, s);
        src_t5(*sp, 0.0, dtv, st);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            ut[v] = fma(dtv, s[v], ut[v]);
            ust[v] = fma(dtv, s[v], ust[v]);
            Ssrc[v] = fma(1.0 / 3.0, st[v], s[v]);
        }
    }
    double dx[4], dy[4];
    mdrk_ft(uN, ut, gm1, dx, dy);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F[v] = fma(1.0 / 6.0, dx[v], f[v]);    // F2
        G[v] = fma(1.0 / 6.0, dy[v], g[v]);
        Uc[v] = fma(1.0 / 6.0, ut[v], uN[v]);  // U2
    }
    mdrk_ft(usn, ust, gm1, dx, dy);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F[v] = fma(1.0 / 3.0, dx[v], F[v]);
        G[v] = fma(1.0 / 3.0, dy[v], G[v]);
        Uc[v] = fma(1.0 / 3.0, ust[v], Uc[v]);
    }
    to_contravariant(J, F, G, cx, cy);
    if (act) put8<S>(s2, sidx, cx, cy);
    __syncthreads();
    contract<n, S>(s2, sbase, i, j, sDh, vol);
}

// Curvilinear CFL speeds of a node (Trixi max_dt for curvilinear meshes): the
// speeds transformed to the reference element, |Ja^k · (λ1, λ2)| |J⁻¹|.
__device__ __forceinline__ void curved_speeds(const double* u, const double* J, double invJ, double gamma, double& m1,
                                              double& m2, bool& bad) {
    double l1, l2;
    node_speeds(u, gamma, l1, l2, bad);
    const double aj = fabs(invJ);
    m1 = fabs(fma(J[0], l1, J[1] * l2)) * aj;
    m2 = fabs(fma(J[2], l1, J[3] * l2)) * aj;
}

// ---------------------------------------------------------------------------
// Volume kernel: element computation, then the interfaces inside the
// workgroup, the lifted surface terms (weak form, ℓ̂ = 1/w at the end nodes),
// the Jacobian and the update.  Faces leaving the workgroup write their face
// record for the surface kernel.  The CFL reduction of a step is left to the
// surface kernel, which reads every node once after its corrections (a
// reduction here costs barriers and LDS passes worth 15% of this kernel).
// Waves per SIMD the register budget is sized for: 3 (168 VGPRs) for the
// plain scheme (TLW_VOL_WAVES at build time); 2 (256 VGPRs) with shock
// capturing, whose subcell FV and blended face fluxes would otherwise spill,
// and with source terms (the source's time differences at every stage).
#ifndef TLW_VOL_WAVES
#define TLW_VOL_WAVES 3
#endif
template <bool TWO> constexpr int kVolWaves = TWO ? 2 : TLW_VOL_WAVES;

// PAR: hyperbolic-parabolic (compressible Navier-Stokes) element procedures,
// same register budget as shock capturing.
// The volume workgroup `bid` of a's table, with its LDS: lds [G_::LDS], sDm,
// sDh, sD [NN] (-dt D, D̂ and D, row-major) and s_alpha [EB] (smoothed blending
// factors of the curvilinear interfaces).
template <int N, int EB, int KIND, bool SC, bool PAR, bool CURV>
__device__ __forceinline__ void volume_body(const KArgs& a, int bid, double* lds, double* sDm, double* sDh, double* sD,
                                            double* s_alpha) {
    static_assert(!(SC && PAR), "shock capturing has no hyperbolic-parabolic form");
    static_assert(!CURV || !PAR || KIND != VOL_LW_ADAPT, "adaptive steps: Euler LW, or Navier-Stokes LW on a TreeMesh");
    static_assert(!PAR || vol_base(KIND) != VOL_LW_ADAPT || N == 4, "the parabolic temporal error is kernel 4's");
    static_assert(!vol_is_wall(KIND) || (PAR && !CURV && !SC), "walls: the TreeMesh Navier-Stokes kernels");
    static_assert(!(CURV && PAR && SC), "no shock capturing with parabolic terms");
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, NN = G_::NN, S = G_::S, NT = G_::NT;
    double* s0 = lds;
    double* s1 = lds + 8 * S;
    double* s2 = lds + 16 * S;
    const int4 blk = load_blk_at(a.vblk, bid);
    DevState* st = a.st + blk.z;
    if (a.use_state_dt && st->finished) return;
    const int tid = threadIdx.x;
    const int le = tid / NN, node = tid - (tid / NN) * NN;
    const int i = node % n, j = node / n;
    const int e0 = blk.x;
    const int nel = blk.y;
    const bool act = le < nel;
    const int lec = le < nel ? le : nel - 1;
    const int e = e0 + lec;
    const int sidx = lec * NN + node, sbase = lec * NN;
    const int cur = a.use_state_dt ? st->cur : 0;
    const double* uin = sel_in(a, cur);
    double* out = sel_out(a, cur);
    const double dt = a.use_state_dt ? st->dt : a.dt_arg;
    const DevConst* c = a.c;
    const double gamma = c->gamma, gm1 = gamma - 1.0;
    // inverse Jacobian: per element on a TreeMesh, per node on a curvilinear mesh
    const double invJ = CURV ? a.ijn[(size_t)e * NN + node] : a.inv_jac[e];
    const size_t off = ((size_t)e * NN + node) * 4;
    // source terms: separate instantiations (KIND = VOL_LW_SRC + the plain kind);
    // the node's coordinates from the element's centre and edge length, the
    // time of the call; Ssrc = S, added to the residual at the end
    constexpr bool SRC = vol_is_src(KIND);
    constexpr bool WALL = vol_is_wall(KIND);
    constexpr int KB = vol_base(KIND);
    static_assert(!SRC || (!SC && !PAR), "source terms: the Euler kernels");
    double Ssrc[4] = {0.0, 0.0, 0.0, 0.0};
    SrcPt sp{0.0, 0.0, 0.0, gamma};
    if constexpr (SRC && CURV) {  // the node's coordinates of the curvilinear mesh
        sp.x = a.nxy[((size_t)e * NN + node) * 2];
        sp.y = a.nxy[((size_t)e * NN + node) * 2 + 1];
        sp.t = a.use_state_dt ? st->t : a.t_arg;
    } else if constexpr (SRC) {
        const double hh = 0.5 * a.length[e];
        sp.x = a.center[2 * e] + hh * c->nodes[i];
        sp.y = a.center[2 * e + 1] + hh * c->nodes[j];
        sp.t = a.use_state_dt ? st->t : a.t_arg;
    }

    if (tid < NN) {  // published by the first barrier of the element computation
        sDm[tid] = -dt * c->D[tid / n][tid % n];
        sDh[tid] = c->Dhat[tid / n][tid % n];
        if constexpr (PAR) sD[tid] = c->D[tid / n][tid % n];
    }
    double uN[4], Uc[4], F[4], G[4], vol[4];
    double terr = 0.0;  // temporal error of the node (adaptive steps, LW N=4)
    load4(uin + off, uN);
    double Jn[4] = {0.0, 0.0, 0.0, 0.0};  // Ja11, Ja12, Ja21, Ja22 of the node: same layout as the state
    // walls: the node's wall records (its x face, its y face; null where that face is no wall)
    double* wq[2] = {nullptr, nullptr};
    if constexpr (WALL) {
        const int wfx = (i == 0) ? 0 : ((i == n - 1) ? 1 : -1), wfy = (j == 0) ? 2 : ((j == n - 1) ? 3 : -1);
        if (act && wfx >= 0 && a.nbr[4 * e + wfx] == -1) wq[0] = a.fvb + (((size_t)e * 4 + wfx) * n + j) * 6;
        if (act && wfy >= 0 && a.nbr[4 * e + wfy] == -1) wq[1] = a.fvb + (((size_t)e * 4 + wfy) * n + i) * 6;
    }
    if constexpr (CURV) load4(a.Ja + off, Jn);
    if constexpr (CURV && !PAR) {
        if constexpr (KIND == VOL_LW) {
            lwc_element<N, EB>(uN, Jn, invJ, gm1, i, j, act, sidx, sbase, s0, s1, s2, sDm, sDh, Uc, F, G, vol);
        } else if constexpr (KIND == VOL_LW_SRC) {
            lwc_element<N, EB, true>(uN, Jn, invJ, gm1, i, j, act, sidx, sbase, s0, s1, s2, sDm, sDh, Uc, F, G, vol,
                                     nullptr, 0.0, 0.0, &sp, Ssrc, dt);
        } else if constexpr (KIND == VOL_MDRK1_SRC) {
            mdrkc1_element<N, EB, true>(uN, Jn, invJ, gm1, i, j, act, sidx, sbase, s0, s2, sDm, sDh, Uc, F, G, vol,
                                        &sp, Ssrc, dt);
        } else if constexpr (KIND == VOL_MDRK2_SRC) {
            double usn[4];
            load4(a.us + off, usn);
            mdrkc2_element<N, EB, true>(uN, usn, Jn, invJ, gm1, i, j, act, sidx, sbase, s0, s1, s2, sDm, sDh, Uc, F,
                                        G, vol, &sp, Ssrc, dt);
        } else if constexpr (KIND == VOL_LW_ADAPT) {
            lwc_element<N, EB>(uN, Jn, invJ, gm1, i, j, act, sidx, sbase, s0, s1, s2, sDm, sDh, Uc, F, G, vol,
                               &terr, st->abstol, st->reltol);
        } else if constexpr (KIND == VOL_MDRK1) {
            mdrkc1_element<N, EB>(uN, Jn, invJ, gm1, i, j, act, sidx, sbase, s0, s2, sDm, sDh, Uc, F, G, vol);
        } else {
            double usn[4];
            load4(a.us + off, usn);
            mdrkc2_element<N, EB>(uN, usn, Jn, invJ, gm1, i, j, act, sidx, sbase, s0, s1, s2, sDm, sDh, Uc, F, G,
                                  vol);
        }
    } else if constexpr (PAR) {
        // BR1 gradients of uⁿ (and of us in MDRK stage 2), exchanged through a
        // buffer the element procedure writes only after its first barrier
        double* sg = (KB == VOL_MDRK2) ? s2 : s1;
        double gxN[4], gyN[4];
        if (act) put4<S>(sg, sidx, uN);
        // the Navier-Stokes convergence source (MDRK on the TreeMesh): the node's
        // spatial factors once, at the call's t
        NsSrcPt nsp{};
        bool nsrc = false;
        if constexpr (!CURV && (KB == VOL_MDRK1 || KB == VOL_MDRK2 || (WALL && KB == VOL_LW))) {
            nsrc = c->src_kind == SRC_NS_CONVERGENCE_TEST;
            if (nsrc) {
                const double hh = 0.5 * a.length[e];
                nsp = ns_src_point(a.center[2 * e] + hh * c->nodes[i], a.center[2 * e + 1] + hh * c->nodes[j],
                                   a.use_state_dt ? st->t : a.t_arg, c);
            }
        }
        if constexpr (KB == VOL_MDRK2) {
            double usn[4], gxS[4], gyS[4];
            load4(a.us + off, usn);
            if (act) put4<S>(sg + 4 * S, sidx, usn);
            __syncthreads();
            br1_gradient<N, EB, CURV>(sg, uin, a.nbr, c, e, i, j, sbase, invJ, sDh, uN, gxN, gyN, Jn);
            br1_gradient<N, EB, CURV>(sg + 4 * S, a.us, a.nbr, c, e, i, j, sbase, invJ, sDh, usn, gxS, gyS, Jn);
            mdrkp2_element<N, EB, CURV, WALL>(uN, gxN, gyN, usn, gxS, gyS, invJ, gm1, c, i, j, act, sidx, sbase, s0,
                                              s1, s2, sDm, sD, sDh, Uc, F, G, vol, Jn, wq, nsrc, nsp, dt, Ssrc);
        } else {
            __syncthreads();
            br1_gradient<N, EB, CURV>(sg, uin, a.nbr, c, e, i, j, sbase, invJ, sDh, uN, gxN, gyN, Jn);
            if constexpr (KB == VOL_LW)
                lwp_element<N, EB, CURV, WALL>(uN, gxN, gyN, invJ, gm1, c, i, j, act, sidx, sbase, s0, s1, s2, sDm,
                                               sD, sDh, Uc, F, G, vol, Jn, wq, nullptr, 0.0, 0.0, nsrc, nsp, Ssrc,
                                               dt);
            else if constexpr (KB == VOL_LW_ADAPT)
                lwp_element<N, EB, CURV, WALL>(uN, gxN, gyN, invJ, gm1, c, i, j, act, sidx, sbase, s0, s1, s2, sDm,
                                               sD, sDh, Uc, F, G, vol, Jn, wq, &terr, st->abstol, st->reltol);
            else
                mdrkp1_element<N, EB, CURV, WALL>(uN, gxN, gyN, invJ, gm1, c, i, j, act, sidx, sbase, s0, s1, s2, sDm,
                                                  sD, sDh, Uc, F, G, vol, Jn, wq, nsrc, nsp, dt, Ssrc);
        }
    } else if (KB == VOL_LW || KB == VOL_LW_ADAPT) {
        if constexpr (N == 4 && KB == VOL_LW_ADAPT)
            lw4_element<N, EB>(uN, dt, invJ, gm1, c, i, j, act, sidx, sbase, s0, s1, s2, sDm, sDh, Uc, F, G, vol,
                               nullptr, &terr, st->abstol, st->reltol);
        else if constexpr (N == 4)
            lw4_element<N, EB, false, SRC>(uN, dt, invJ, gm1, c, i, j, act, sidx, sbase, s0, s1, s2, sDm, sDh, Uc, F,
                                           G, vol, nullptr, nullptr, 0.0, 0.0, &sp, Ssrc);
        else
            lw3_element<N, EB, false, SRC>(uN, dt, invJ, gm1, c, i, j, act, sidx, sbase, s0, s1, s2, sDm, sDh, Uc, F,
                                           G, vol, nullptr, &sp, Ssrc);
    } else if (KB == VOL_MDRK1) {
        mdrk1_element<N, EB, SRC>(uN, dt, invJ, gm1, c, i, j, act, sidx, sbase, s0, s1, s2, sDm, sDh, Uc, F, G, vol,
                                  &sp, Ssrc);
    } else {
        double usn[4];
        load4(a.us + off, usn);
        mdrk2_element<N, EB, SRC>(uN, usn, dt, invJ, gm1, c, i, j, act, sidx, sbase, s0, s1, s2, sDm, sDh, Uc, F, G,
                                  vol, &sp, Ssrc);
    }
    if (act && a.auxU) {
        store4(a.auxU + off, Uc);
        store4(a.auxF + ((size_t)e * NN + node) * 8, F);
        store4(a.auxF + ((size_t)e * NN + node) * 8 + 4, G);
    }

    const int fx = (i == 0) ? 0 : ((i == n - 1) ? 1 : -1);
    const int fy = (j == 0) ? 2 : ((j == n - 1) ? 3 : -1);
    constexpr int RQ = kRQ<SC>;
    double fnlx[4] = {0, 0, 0, 0}, fnly[4] = {0, 0, 0, 0};
    double alpha_e = 0.0;
    if constexpr (SC) {
        // blending factor, smoothed over the face neighbours (Trixi apply_smoothing!)
        alpha_e = a.alpha_tmp[e];
        if (c->alpha_smooth) {
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const int nb = a.nbr[4 * e + f];
                if (nb >= 0) alpha_e = dmax(alpha_e, 0.5 * a.alpha_tmp[nb]);
            }
        }
        if (act && node == 0 && a.alpha_out) a.alpha_out[e] = alpha_e;
        if constexpr (CURV) {
            if (act && node == 0) s_alpha[lec] = alpha_e;
        }
        // subcell FV (fv_kernel!, dg_2d.jl:251-324): Rusanov fluxes between the
        // face values of neighbouring subcells, zero at the element faces.  First
        // order (Trixi calcflux_fv!): the face values are the node values.  A
        // blended element away from the domain boundary with MUSCL-Hancock takes
        // the reconstructed, half-step-evolved face values (mh_faces).
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (act) s1[v * S + sidx] = uN[v];
        __syncthreads();
        const bool blended = !(fabs(alpha_e) <= 1e-12);  // pure_and_blended_element_ids!
        double ql[4], qr[4], qd[4], qu[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) ql[v] = qr[v] = qd[v] = qu[v] = uN[v];
        if (!CURV && c->recon == 1 && blended && a.nbr[4 * e] >= 0 && a.nbr[4 * e + 1] >= 0 && a.nbr[4 * e + 2] >= 0 &&
            a.nbr[4 * e + 3] >= 0) {
            const double dtmh = (KIND == VOL_MDRK1) ? 0.5 * dt : dt;  // tree/multi_derivative_2d.jl:199
            mh_faces<N, S>(s1, sidx, i, j, uN, c, 2.0 - alpha_e, dtmh, invJ, gm1, ql, qr, qd, qu);
        }
        // the right and upper face values go to the neighbouring subcells
        if (act) put8<S>(s0, sidx, qr, qu);
        __syncthreads();
        double fxl[4] = {0, 0, 0, 0}, gyl[4] = {0, 0, 0, 0}, w[4];
        if (i > 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) w[v] = s0[v * S + sidx - 1];
            if constexpr (CURV) {
                // Trixi's subcell normal, or for a pure-DG element at subcell face n-1 the
                // normal of the reference's calc_fn_low! (structured/dg_2d.jl:111-173)
                const double* nq = (!blended && i == n - 1) ? a.sqrk + ((size_t)e * n + j) * 4
                                                             : a.snrm + ((size_t)e * NN + node) * 4;
                rusanov_normal(w, ql, gamma, nq[0], nq[1], sqrt(fma(nq[0], nq[0], nq[1] * nq[1])), fxl);
            } else {
                rusanov(w, ql, gamma, 1, fxl);
            }
        }
        if (j > 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) w[v] = s0[(4 + v) * S + sidx - n];
            if constexpr (CURV) {
                const double* nq = (!blended && j == n - 1) ? a.sqrk + ((size_t)e * n + i) * 4 + 2
                                                             : a.snrm + ((size_t)e * NN + node) * 4 + 2;
                rusanov_normal(w, qd, gamma, nq[0], nq[1], sqrt(fma(nq[0], nq[0], nq[1] * nq[1])), gyl);
            } else {
                rusanov(w, qd, gamma, 2, gyl);
            }
        }
        if (act) put8<S>(s2, sidx, fxl, gyl);
        __syncthreads();
        const double iwi = 1.0 / c->weights[i], iwj = 1.0 / c->weights[j];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double fxr = (i < n - 1) ? s2[v * S + sidx + 1] : 0.0;
            const double gyr = (j < n - 1) ? s2[(4 + v) * S + sidx + n] : 0.0;
            const double fv = iwi * (fxr - fxl[v]) + iwj * (gyr - gyl[v]);
            vol[v] = (1.0 - alpha_e) * vol[v] + (blended ? alpha_e * fv : 0.0);
            // fn_low: the subcell fluxes next to the element faces (load_fn_low!)
            fnlx[v] = (fx == 0) ? s2[v * S + sidx + 1] : fxl[v];
            fnly[v] = (fy == 2) ? s2[(4 + v) * S + sidx + n] : gyl[v];
        }
        __syncthreads();  // s0, s1 and s2 are reused for the face records and interface sums
    }
    // --- face records into LDS (s0, s1 are free), one per face the node lies on.
    // λ ingredients: |v1|, |v2| and c of the λ state, once per node.
    double* rec = lds;  // [EB][4][n][RQ]
    const bool tree = c->flux_order == 0;
    const double* jump = tree ? uN : Uc;
    const double* lamst = tree ? Uc : uN;
    const int nbx = fx >= 0 ? a.nbr[4 * e + fx] : -1;
    const int nby = fy >= 0 ? a.nbr[4 * e + fy] : -1;
    if constexpr (CURV) {
        // structured order with the interface's normal n (shared by both
        // elements): F̃·n, |v(uⁿ)·n| and c(uⁿ)|n| (dg_2d.jl:2229-2264)
        const double ir = rcp_nr(uN[0]);
        const double v1 = uN[1] * ir, v2 = uN[2] * ir;
        const double p = gm1 * (uN[3] - 0.5 * (uN[1] * v1 + uN[2] * v2));
        const double cs = sqrt_gs(gamma * p * ir);
        // (a domain-boundary face is rewritten to the TreeMesh order of the
        // Dirichlet call by the surface kernel, which has uⁿ at hand)
        // with shock capturing: Ũ, F̃·n, uⁿ and fn_low (face_numflux_sc_curved)
        auto face = [&](int f, int l, const double* fnl) {
            double nf[4], Fn[4], r[RQ];
            load4(a.fnorm + (((size_t)e * 4 + f) * n + l) * 4, nf);
#pragma unroll
            for (int v = 0; v < 4; ++v) Fn[v] = fma(F[v], nf[0], G[v] * nf[1]);
            if constexpr (SC) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    r[v] = Uc[v];
                    r[4 + v] = Fn[v];
                    r[8 + v] = uN[v];
                    r[12 + v] = fnl[v];
                }
            } else {
                make_record(r, Uc, Fn, fabs(fma(v1, nf[0], v2 * nf[1])), cs * nf[2]);
            }
            store_record<RQ>(rec + ((lec * 4 + f) * n + l) * RQ, r);
        };
        if (act && fx >= 0) face(fx, j, fnlx);
        if (act && fy >= 0) face(fy, i, fnly);
    } else {
        double av1, av2, cs;
        speed_parts2(lamst, gamma, av1, av2, cs);
        auto fill = [&](double* r, const double* fnl) {
            if constexpr (SC) {
#pragma unroll
                for (int v = 0; v < 4; ++v) r[10 + v] = fnl[v];
                r[14] = alpha_e;
                r[15] = 0.0;
            }
        };
        if (act && fx >= 0) {
            double r[RQ];
            make_record(r, jump, F, av1, cs);
            fill(r, fnlx);
            store_record<RQ>(rec + ((lec * 4 + fx) * n + j) * RQ, r);
            if (nbx <= -2) mortar_side(a, nbx, n, j, jump, F, lamst);
        }
        if (act && fy >= 0) {
            double r[RQ];
            make_record(r, jump, G, av2, cs);
            fill(r, fnly);
            store_record<RQ>(rec + ((lec * 4 + fy) * n + i) * RQ, r);
            if (nby <= -2) mortar_side(a, nby, n, i, jump, G, lamst);
        }
    }
    __syncthreads();
    // --- one lane per face slot (element, face, node).  An interface inside the
    // workgroup is evaluated once, by the slot of its left element's + face, and
    // its flux goes to both adjacent nodes (accx for x faces, accy for y faces);
    // the record of a face leaving the workgroup goes to HBM for the surface kernel.
    double* accx = lds + 16 * S;  // [4][S], in s2 (free after the last contraction)
    double* accy = accx + 4 * S;
    constexpr int NSLOT = EB * 4 * n;
    static_assert(NSLOT <= NT, "one face slot per lane");
    static_assert(EB * 4 * n * RQ <= 16 * S, "records fit in s0 and s1");
    if (tid < NSLOT) {
        const int ls = tid / (4 * n), f = (tid / n) % 4, l = tid % n;
        if (ls < nel) {
            const int es = e0 + ls;
            const int ne = a.nbr[4 * es + f];
            const double* own = rec + tid * RQ;
            if (!(ne >= e0 && ne < e0 + nel)) {
                double r[RQ];
                load_record<RQ>(own, r);
                store_record<RQ>(a.trace + ((size_t)es * 4 * n + f * n + l) * RQ, r);
            } else if (f & 1) {
                double Lr[RQ], Rr[RQ], Fs[4];
                load_record<RQ>(own, Lr);
                load_record<RQ>(rec + (((ne - e0) * 4 + (f ^ 1)) * n + l) * RQ, Rr);
                // + face of ls: node (n-1, l) or (l, n-1); - face of ne: node (0, l) or (l, 0)
                const int nodeL = (f == 1) ? l * n + (n - 1) : (n - 1) * n + l;
                const int nodeR = (f == 1) ? l * n : l;
                if constexpr (SC && CURV) {
                    double nf[4];
                    load4(a.fnorm + (((size_t)es * 4 + f) * n + l) * 4, nf);
                    face_numflux_sc_curved(Lr, Rr, nf, a.ijn[(size_t)es * NN + nodeL], a.ijn[(size_t)ne * NN + nodeR],
                                           s_alpha[ls], s_alpha[ne - e0], gamma, dt, c->weights[n - 1], c->weights[0],
                                           Fs);
                } else if constexpr (SC) {
                    face_numflux_sc(Lr, Rr, f < 2 ? 1 : 2, gamma, st->invJ0, dt, c->weights[n - 1], c->weights[0], Fs);
                } else {
                    face_numflux(Lr, Rr, Fs);
                }
                double* acc = (f == 1) ? accx : accy;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    acc[v * S + ls * NN + nodeL] = Fs[v];
                    acc[v * S + (ne - e0) * NN + nodeR] = Fs[v];
                }
            }
        }
    }
    __syncthreads();
    // --- lifted internal fluxes, Jacobian, output
    const bool intx = fx >= 0 && nbx >= e0 && nbx < e0 + nel;
    const bool inty = fy >= 0 && nby >= e0 && nby < e0 + nel;
    double surf[4] = {0.0, 0.0, 0.0, 0.0};
    if (intx) {
        const double wf = (fx & 1) ? c->bnd_hi : -c->bnd_lo;
#pragma unroll
        for (int v = 0; v < 4; ++v) surf[v] = accx[v * S + sidx] * wf;
    }
    if (inty) {
        const double wf = (fy & 1) ? c->bnd_hi : -c->bnd_lo;
#pragma unroll
        for (int v = 0; v < 4; ++v) surf[v] = fma(accy[v * S + sidx], wf, surf[v]);
    }
    double val[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double d = (-invJ) * (vol[v] + surf[v]);
        if constexpr (SRC || (PAR && !CURV && (KB == VOL_MDRK1 || KB == VOL_MDRK2 || (WALL && KB == VOL_LW))))
            d += Ssrc[v];
        val[v] = a.step_mode ? fma(dt, d, uN[v]) : d;
    }
    if (act) store4(out + off, val);
    if constexpr (KB == VOL_LW_ADAPT) {
        if (a.reduce != RED_ADAPT) return;
        // the workgroup's sum of its elements' temporal errors (adaptive.jl:48-64)
        double* se = lds;
        double* rs = lds + S;
        double* es = rs + EB * n;
        if (act) se[sidx] = terr;
        __syncthreads();
        if (act && i == 0) {
            double r = 0.0;
#pragma unroll
            for (int k = 0; k < n; ++k) r += se[sbase + j * n + k];
            rs[lec * n + j] = r;
        }
        __syncthreads();
        if (act && node == 0) {
            double r = 0.0;
#pragma unroll
            for (int k = 0; k < n; ++k) r += rs[lec * n + k];
            es[lec] = r;
        }
        __syncthreads();
        if (tid == 0) {
            double r = 0.0;
            for (int k = 0; k < nel; ++k) r += es[k];
            a.partial[bid] = r;
        }
    }
    // the CFL reduction of a step (and the admissibility test) is the surface
    // kernel's: it reads every node once after the last correction
}

// the Navier-Stokes LW and MDRK stage-2 wall kernels (a rarely used
// configuration: the lid-driven cavity's walls) run at one wave per SIMD, so
// their extra wall terms live in the accumulation registers instead of private
// memory; the stage-1 wall kernel fits two waves
template <int KIND, bool SC, bool PAR>
constexpr int kVolWavesOf =
    (KIND == VOL_LW_WALL || KIND == VOL_MDRK2_WALL || KIND == VOL_LW_ADAPT_WALL)
        ? 1
        : kVolWaves<SC || PAR || vol_is_src(KIND)>;
template <int N, int EB, int KIND, bool SC, bool PAR = false, bool CURV = false>
__global__ __launch_bounds__((VGeo<N, EB>::NT), (kVolWavesOf<KIND, SC, PAR>)) void k_volume(KArgs a) {
    using G_ = VGeo<N, EB>;
    __shared__ double lds[G_::LDS];
    __shared__ double sDm[G_::NN], sDh[G_::NN], sD[G_::NN];
    __shared__ double s_alpha[(SC && CURV) ? EB : 1];
    volume_body<N, EB, KIND, SC, PAR, CURV>(a, blockIdx.x, lds, sDm, sDh, sD, s_alpha);
}

// Outer record of a domain-boundary face node: time average of the boundary
// value over the step with the GLL rule (equations.jl:6-37).
template <int N>
__device__ inline void boundary_record(const KArgs& a, int e, int face, int l, double t, double dt, double* r) {
    constexpr int n = N + 1;
    const DevConst* c = a.c;
    const int orientation = face < 2 ? 1 : 2;
    const double half = 0.5 * a.length[e];
    double x, y;
    if (orientation == 1) {
        x = a.center[2 * e] + half * c->nodes[face == 0 ? 0 : n - 1];
        y = a.center[2 * e + 1] + half * c->nodes[l];
    } else {
        x = a.center[2 * e] + half * c->nodes[l];
        y = a.center[2 * e + 1] + half * c->nodes[face == 2 ? 0 : n - 1];
    }
    const double sf = a.bc_scale;
    const double dts = sf * dt;
    const double gm1 = c->gamma - 1.0;
    double Uo[4] = {0, 0, 0, 0}, Fo[4] = {0, 0, 0, 0}, uo[4];
    for (int q = 0; q < n; ++q) {
        const double ts = t + 0.5 * dts * (c->nodes[q] + 1.0);
        double ub[4], fb[4];
        boundary_state<n>(c, a.bidx, a.btab, a.btab_levels, e, face, l, q, x, y, ts, ub);
        euler_flux_dir(ub, gm1, orientation, fb);
        for (int v = 0; v < 4; ++v) {
            Uo[v] += 0.5 * sf * ub[v] * c->weights[q];
            Fo[v] += 0.5 * sf * fb[v] * c->weights[q];
        }
    }
    boundary_state<n>(c, a.bidx, a.btab, a.btab_levels, e, face, l, n, x, y, t, uo);
    const bool tree = c->flux_order == 0;
    double av1, av2, cs;
    speed_parts2(tree ? Uo : uo, c->gamma, av1, av2, cs);
    make_record(r, tree ? uo : Uo, Fo, orientation == 1 ? av1 : av2, cs);
}

// Outer record of a domain-boundary face node of a curvilinear mesh
// (calc_boundary_flux_by_direction!, structured/dg_2d.jl:2352-2391, into the
// Dirichlet call of equations.jl:6-37): the GLL time averages (over
// [t, t + s dt], s = a.bc_scale: 1 on a StructuredMesh, 0.5 in MDRK stage 1 of
// a P4estMesh) of the boundary value and of its flux in the normal direction
// (Trixi flux(u, n): (ρv_n, ρv_n v1 + p n1, ρv_n v2 + p n2, (ρe + p) v_n)) with
// the element's own face normal, and λ's parts of the average.  The record's
// jump state is u_outer(t) (the StructuredMesh call: TreeMesh order, jump on
// uⁿ, λ from Ũ) or, with p4est_form, the time average itself (the P4est
// method, equations.jl:66-93).
template <int N>
__device__ inline void boundary_record_curved(const KArgs& a, int e, int face, int l, double t, double dt, double* r,
                                              bool p4est_form = false) {
    constexpr int n = N + 1;
    const DevConst* c = a.c;
    const size_t fo = ((size_t)e * 4 + face) * n + l;
    const double x = a.fxy[2 * fo], y = a.fxy[2 * fo + 1];
    double nf[4];
    load4(a.fnorm + 4 * fo, nf);
    const double gm1 = c->gamma - 1.0;
    const double sf = a.bc_scale, dts = sf * dt;
    double Uo[4] = {0, 0, 0, 0}, Fo[4] = {0, 0, 0, 0}, uo[4];
    for (int q = 0; q < n; ++q) {
        const double ts = t + 0.5 * dts * (c->nodes[q] + 1.0);
        double ub[4];
        boundary_state<n>(c, a.bidx, a.btab, a.btab_levels, e, face, l, q, x, y, ts, ub);
        const double rho = ub[0], v1 = ub[1] / rho, v2 = ub[2] / rho;
        const double p = gm1 * (ub[3] - 0.5 * (ub[1] * v1 + ub[2] * v2));
        const double vn = v1 * nf[0] + v2 * nf[1], rvn = rho * vn;
        const double fb[4] = {rvn, rvn * v1 + p * nf[0], rvn * v2 + p * nf[1], (ub[3] + p) * vn};
        for (int v = 0; v < 4; ++v) {
            Uo[v] += 0.5 * sf * ub[v] * c->weights[q];
            Fo[v] += 0.5 * sf * fb[v] * c->weights[q];
        }
    }
    if (!p4est_form) boundary_state<n>(c, a.bidx, a.btab, a.btab_levels, e, face, l, n, x, y, t, uo);
    const double ir = rcp_nr(Uo[0]);
    const double b1 = Uo[1] * ir, b2 = Uo[2] * ir;
    const double pb = gm1 * (Uo[3] - 0.5 * (Uo[1] * b1 + Uo[2] * b2));
    make_record(r, p4est_form ? Uo : uo, Fo, fabs(fma(b1, nf[0], b2 * nf[1])), sqrt_gs(c->gamma * pb * ir) * nf[2]);
}

// Surface kernel: the faces that leave a volume workgroup (a.ebv elements).
// Each interface is evaluated by both neighbours from identical inputs, so
// both see the same bits.  In step mode every lane then takes its node's
// final value (corrected here or finalised by the volume kernel) into the
// CFL maxima and the admissibility test.
// BC: the mesh has domain-boundary faces; MORT: it has faces on L2 mortars (a
// mesh with mortars and no boundary face compiles no boundary code).
template <int N, bool BC, bool SC, bool CURV = false, bool MORT = false>
__device__ __forceinline__ void surface_body(const KArgs& a, int bid) {
    using G_ = SGeo<N>;
    constexpr int n = G_::n, NN = G_::NN, EB = G_::EB, S = G_::S;
    __shared__ double s_l1[S], s_l2[S], s_m[EB], s_lv[S], s_mv[EB];
    const int4 blk = load_blk_at(a.sblk, bid);
    DevState* st = a.st + blk.z;
    if (a.use_state_dt && st->finished) return;
    const int tid = threadIdx.x;
    const int le = tid / NN, node = tid - (tid / NN) * NN;
    const int i = node % n, j = node / n;
    const int e0 = blk.x;
    const int nel = blk.y;
    const bool act = le < nel;
    const int lec = le < nel ? le : nel - 1;
    const int e = e0 + lec;
    const int cur = a.use_state_dt ? st->cur : 0;
    double* out = sel_out(a, cur);
    const double dt = a.use_state_dt ? st->dt : a.dt_arg;
    const double t = a.use_state_dt ? st->t : a.t_arg;
    const DevConst* c = a.c;
    const double invJ = CURV ? a.ijn[(size_t)e * NN + node] : a.inv_jac[e];
    const int eblk = a.etile[e];
    const int fx = (i == 0) ? 0 : ((i == n - 1) ? 1 : -1);
    const int fy = (j == 0) ? 2 : ((j == n - 1) ? 3 : -1);
    const int nbx = fx >= 0 ? a.nbr[4 * e + fx] : 0;
    const int nby = fy >= 0 ? a.nbr[4 * e + fy] : 0;
    const bool extx = fx >= 0 && (nbx < 0 || a.etile[nbx] != eblk);
    const bool exty = fy >= 0 && (nby < 0 || a.etile[nby] != eblk);
    const bool any_ext = extx || exty;
    double val[4];
    if (act && any_ext) {
        const size_t off = ((size_t)e * NN + node) * 4;
        load4(out + off, val);
        double surf[4] = {0.0, 0.0, 0.0, 0.0};
        constexpr int RQ = kRQ<SC>;
        auto do_face = [&](int f, int l, int nb) {
            double own[RQ], oth[RQ], Fs[4];
            load_record<RQ>(a.trace + (((size_t)e * 4 + f) * n + l) * RQ, own);
            const double* Lr = (f & 1) ? own : oth;
            const double* Rr = (f & 1) ? oth : own;
            if ((!BC && !MORT) || nb >= 0) {
                load_record<RQ>(a.trace + (((size_t)nb * 4 + (f ^ 1)) * n + l) * RQ, oth);
                if constexpr (SC && CURV) {
                    // the neighbour's mirrored face node and the element-wise blending factors
                    const int nnode = (f == 0) ? l * n + (n - 1) : (f == 1) ? l * n : (f == 2) ? (n - 1) * n + l : l;
                    const double Jo = a.ijn[(size_t)e * NN + node], Jn = a.ijn[(size_t)nb * NN + nnode];
                    const double ao = a.alpha_out[e], an = a.alpha_out[nb];
                    double nf[4];
                    load4(a.fnorm + (((size_t)e * 4 + f) * n + l) * 4, nf);
                    if (f & 1)
                        face_numflux_sc_curved(Lr, Rr, nf, Jo, Jn, ao, an, c->gamma, dt, c->weights[n - 1],
                                               c->weights[0], Fs);
                    else
                        face_numflux_sc_curved(Lr, Rr, nf, Jn, Jo, an, ao, c->gamma, dt, c->weights[n - 1],
                                               c->weights[0], Fs);
                } else if constexpr (SC) {
                    face_numflux_sc(Lr, Rr, f < 2 ? 1 : 2, c->gamma, st->invJ0, dt, c->weights[n - 1], c->weights[0],
                                    Fs);
                } else {
                    face_numflux(Lr, Rr, Fs);
                }
            } else if (MORT && nb <= -2) {
                // a face on an L2 mortar: the mortar kernel's flux of this element's role
                load4(a.mflux + ((size_t)(-2 - nb) * n + l) * 4, Fs);
            } else if constexpr (BC) {
                // domain boundary: time-averaged Dirichlet state, no blending
                // (calc_boundary_flux!, dg_2d.jl:1379-1454; structured/dg_2d.jl:2297-2391)
                if constexpr (CURV) {
                    // the volume kernel's record holds Ũ and F̃·n (and uⁿ with shock
                    // capturing); the Dirichlet call takes the TreeMesh order: jump on
                    // uⁿ, λ from Ũ; no blending (structured/dg_2d.jl:2352-2391)
                    double nf[4], in10[kTraceQ], out10[kTraceQ];
                    load4(a.fnorm + (((size_t)e * 4 + f) * n + l) * 4, nf);
                    if (nb == -3) {
                        // P4estMesh interface (prolong2interfaces!/calc_interface_flux!,
                        // p4est/dg_2d.jl:8-192): each record holds F̃ in its element's +ξ/+η
                        // direction (F̃·Ja of its own node); in the primary's outward
                        // direction that is s_p·F̃_p and -s_s·F̃_s (s = +1 on + faces, -1 on
                        // - faces); λ's parts from uⁿ, the jump on Ũ.  Both elements form
                        // the same primary-outward flux; each keeps its +ξ/+η version.
                        const int4 pl = a.plink[4 * e + f];
                        const int lp = pl.z ? n - 1 - l : l;
                        load_record<RQ>(a.trace + (((size_t)pl.x * 4 + pl.y) * n + lp) * RQ, oth);
                        const double so = (f & 1) ? 1.0 : -1.0, sq = (pl.y & 1) ? 1.0 : -1.0;
                        const double* P = pl.w ? own : oth;
                        const double* Q = pl.w ? oth : own;
                        const double sP = pl.w ? so : sq, sQ = pl.w ? sq : so;
                        double Lq[kTraceQ], Rq[kTraceQ], Fo_[4];
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            Lq[v] = P[v];
                            Rq[v] = Q[v];
                            Lq[4 + v] = sP * P[4 + v];
                            Rq[4 + v] = -sQ * Q[4 + v];
                        }
                        Lq[8] = P[8];
                        Lq[9] = P[9];
                        Rq[8] = Q[8];
                        Rq[9] = Q[9];
                        face_numflux(Lq, Rq, Fo_);
#pragma unroll
                        for (int v = 0; v < 4; ++v) Fs[v] = pl.w ? so * Fo_[v] : -so * Fo_[v];
                        const double wf = (f & 1) ? c->bnd_hi : -c->bnd_lo;
#pragma unroll
                        for (int v = 0; v < 4; ++v) surf[v] = fma(Fs[v], wf, surf[v]);
                        return;
                    }
                    if (a.plink) {
                        // P4estMesh boundary with the time-averaged Dirichlet value (the
                        // P4est method, equations.jl:66-93): the volume kernel's record as
                        // the inner state (jump on Ũ, λ from uⁿ), the outer record with the
                        // time average as both jump state and λ state; inner first along
                        // the outward normal, i.e. the outer record on the - side in +ξ/+η
                        if (c->bc_kind != kBcOutflow && c->bc_kind != kBcSlipWallApprox) {
                            boundary_record_curved<N>(a, e, f, l, t, dt, out10, true);
                            if (f & 1) face_numflux(own, out10, Fs);
                            else face_numflux(out10, own, Fs);
                            const double wf = (f & 1) ? c->bnd_hi : -c->bnd_lo;
#pragma unroll
                            for (int v = 0; v < 4; ++v) surf[v] = fma(Fs[v], wf, surf[v]);
                            return;
                        }
                    }
                    if (c->bc_kind == kBcOutflow || c->bc_kind == kBcSlipWallApprox) {
                        // conditions on uⁿ with the node's contravariant vector as normal
                        // (boundary_condition_outflow, slip_wall_approximate,
                        // equations.jl:185-197, 251-260): the normal flux of uⁿ, or the
                        // Rusanov flux between uⁿ and its reflection with the reflected
                        // state on the side of the boundary
                        double un[4], ur[4];
                        if constexpr (SC) {
#pragma unroll
                            for (int v = 0; v < 4; ++v) un[v] = own[8 + v];
                        } else {
                            load4(sel_in(a, cur) + ((size_t)e * NN + node) * 4, un);
                        }
                        if (c->bc_kind == kBcOutflow) {
                            rusanov_normal(un, un, c->gamma, nf[0], nf[1], nf[2], Fs);
                        } else {
                            const double ir = rcp_nr(un[0]);
                            const double v1 = un[1] * ir, v2 = un[2] * ir;
                            const double p = (c->gamma - 1.0) * (un[3] - 0.5 * (un[1] * v1 + un[2] * v2));
                            const double vn = fma(v1, nf[0], v2 * nf[1]) / (nf[2] * nf[2]);
                            const double w1 = v1 - 2.0 * vn * nf[0], w2 = v2 - 2.0 * vn * nf[1];
                            ur[0] = un[0];
                            ur[1] = un[0] * w1;
                            ur[2] = un[0] * w2;
                            ur[3] = p / (c->gamma - 1.0) + 0.5 * un[0] * (w1 * w1 + w2 * w2);
                            if (f & 1) rusanov_normal(un, ur, c->gamma, nf[0], nf[1], nf[2], Fs);
                            else rusanov_normal(ur, un, c->gamma, nf[0], nf[1], nf[2], Fs);
                        }
                        const double wf = (f & 1) ? c->bnd_hi : -c->bnd_lo;
#pragma unroll
                        for (int v = 0; v < 4; ++v) surf[v] = fma(Fs[v], wf, surf[v]);
                        return;
                    }
                    const double ir = rcp_nr(own[0]);
                    const double b1 = own[1] * ir, b2 = own[2] * ir;
                    const double pb = (c->gamma - 1.0) * (own[3] - 0.5 * (own[1] * b1 + own[2] * b2));
                    if constexpr (SC) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) in10[v] = own[8 + v];
                    } else {
                        load4(sel_in(a, cur) + ((size_t)e * NN + node) * 4, in10);
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) in10[4 + v] = own[4 + v];
                    in10[8] = fabs(fma(b1, nf[0], b2 * nf[1]));
                    in10[9] = sqrt_gs(c->gamma * pb * ir) * nf[2];
                    boundary_record_curved<N>(a, e, f, l, t, dt, out10);
                    if (f & 1) face_numflux(in10, out10, Fs);
                    else face_numflux(out10, in10, Fs);
                } else if (c->bc_kind == kBcSlipWall) {
                    // reflecting wall (equations.jl:95-139): the mirrored record, whatever
                    // the flux order (both states are mirrored, the speeds kept)
                    const int m = (f < 2) ? 1 : 2;
                    double fvw[3] = {0.0, 0.0, 0.0};
                    if (c->ns_walls) {
                        // Navier-Stokes: the record holds F̃a - F̃v; the wall flux is built
                        // from F̃a (its normal-momentum component: the others cancel against
                        // their mirror images), then the wall's viscous flux is subtracted
                        const double* q = a.fvb + (((size_t)e * 4 + f) * n + l) * 6;
                        fvw[0] = q[0];
                        fvw[1] = q[1];
                        fvw[2] = q[2];
                        own[4 + m] += fvw[m - 1];
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        oth[v] = (v == m) ? -own[v] : own[v];
                        oth[4 + v] = (v == m) ? own[4 + v] : -own[4 + v];
                    }
                    oth[8] = own[8];
                    oth[9] = own[9];
                    face_numflux(Lr, Rr, Fs);
                    if (c->ns_walls) {
                        // NoSlip/Adiabatic (compressible_navier_stokes_2d.jl:81-91):
                        // (F̃v1, F̃v2, F̃v3, v_wall · (F̃v2, F̃v3)), F̃v1 = 0; NoSlip/Isothermal
                        // (103-109): the inner viscous flux itself
                        Fs[1] -= fvw[0];
                        Fs[2] -= fvw[1];
                        Fs[3] -= c->wall_iso[f] ? fvw[2]
                                                : fma(c->wall_v[2 * f], fvw[0], c->wall_v[2 * f + 1] * fvw[1]);
                    }
                } else {
                    boundary_record<N>(a, e, f, l, t, dt, oth);
                    face_numflux(Lr, Rr, Fs);
                }
            }
            const double wf = (f & 1) ? c->bnd_hi : -c->bnd_lo;
#pragma unroll
            for (int v = 0; v < 4; ++v) surf[v] = fma(Fs[v], wf, surf[v]);
        };
        if (extx) do_face(fx, j, nbx);
        if (exty) do_face(fy, i, nby);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double d = (-invJ) * surf[v];
            val[v] = a.step_mode ? fma(dt, d, val[v]) : val[v] + d;
        }
        store4(out + off, val);
    }
    if (a.reduce == RED_NONE) return;
    // Every node's final value: the corrected ones are in registers, the volume
    // kernel finalised the others.  CFL maxima (Trixi max_dt) and the
    // admissibility test of the updated state cover all nodes here.
    if (act && !any_ext) load4(out + ((size_t)e * NN + node) * 4, val);
    if (a.reduce == RED_ADAPT) {  // test_updated_solution (adaptive.jl:30-46)
        if (act && !admissible(val, c->gamma)) st->bad_state = 1;
        return;
    }
    const bool par = a.pv_off > 0;
    if (act) {
        double l1, l2;
        bool bad;
        if constexpr (CURV) {
            double Jn[4];
            load4(a.Ja + ((size_t)e * NN + node) * 4, Jn);
            curved_speeds(val, Jn, invJ, c->gamma, l1, l2, bad);
        } else {
            node_speeds(val, c->gamma, l1, l2, bad);
        }
        if (bad) st->bad_state = 1;
        s_l1[lec * NN + node] = l1;
        s_l2[lec * NN + node] = l2;
        s_lv[lec * NN + node] = par ? (CURV ? lam_visc_curved(val, invJ, c) : lam_visc(val, invJ, c)) : 0.0;
    }
    __syncthreads();
    if (act && node == 0) {
        double m1 = 0.0, m2 = 0.0, mv = 0.0;
        for (int k = 0; k < NN; ++k) {
            m1 = dmax(m1, s_l1[lec * NN + k]);
            m2 = dmax(m2, s_l2[lec * NN + k]);
            mv = dmax(mv, s_lv[lec * NN + k]);
        }
        // curvilinear speeds carry the node's |J⁻¹| already
        s_m[lec] = CURV ? m1 + m2 : invJ * (m1 + m2);
        s_mv[lec] = mv;
    }
    __syncthreads();
    if (tid == 0) {
        double bm = 0.0, bv = 0.0;
        for (int k = 0; k < nel; ++k) {
            bm = dmax(bm, s_m[k]);
            bv = dmax(bv, s_mv[k]);
        }
        slot_max(a.slots, blk.z, 0, bm);
        if (par) slot_max(a.slots, blk.z, 1, bv);
    }
}

template <int N, bool BC, bool SC, bool CURV = false, bool MORT = false>
__global__ __launch_bounds__(kSurfBlock) void k_surface(KArgs a) {
    surface_body<N, BC, SC, CURV, MORT>(a, blockIdx.x);
}

// AMR solution transfer between two TreeMesh leaf sets (Trixi's refine! and
// coarsen! of the DG solution, src/callbacks_step/amr_dg2d.jl in Trixi.jl,
// which the reference's AMR callback runs before resizing its caches,
// src/callbacks/amr.jl:2-20): one lane per node of the new mesh.  Plan
// [E_new][5]: (0, e) copies leaf e; (1, p, child) interpolates parent p to its
// child (0 lower left, 1 lower right, 2 upper left, 3 upper right) with the
// forward mortar operators in x, then y; (2, c0, c1, c2, c3) projects four
// children onto their parent with the reverse operators, summed over children.
constexpr int kTransferBlock = 256;
template <int N>
__global__ __launch_bounds__(kTransferBlock) void k_transfer(const double* __restrict__ uold, double* __restrict__ unew,
                                                            const int* __restrict__ plan, int E_new, const DevConst* c) {
    constexpr int n = N + 1, NN = n * n;
    const long long gid = (long long)blockIdx.x * kTransferBlock + threadIdx.x;
    if (gid >= (long long)E_new * NN) return;
    const int e = (int)(gid / NN), node = (int)(gid - (long long)e * NN);
    const int i = node % n, j = node / n;
    const int* p = plan + 5 * (size_t)e;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (p[0] == 0) {
        load4(uold + ((size_t)p[1] * NN + node) * 4, acc);
    } else {
        const int nsrc = (p[0] == 1) ? 1 : 4;
        for (int q = 0; q < nsrc; ++q) {
            const int src = (p[0] == 1) ? p[1] : p[1 + q];
            const int child = (p[0] == 1) ? p[2] : q;
            const double* Mx = (p[0] == 1) ? c->mfwd[child & 1][i] : c->mrev[child & 1][i];
            const double* My = (p[0] == 1) ? c->mfwd[child >> 1][j] : c->mrev[child >> 1][j];
            const double* us = uold + (size_t)src * NN * 4;
            for (int l = 0; l < n; ++l) {
                double t[4] = {0.0, 0.0, 0.0, 0.0};
                for (int k = 0; k < n; ++k) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) t[v] = fma(Mx[k], us[((size_t)l * n + k) * 4 + v], t[v]);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[v] = fma(My[l], t[v], acc[v]);
            }
        }
    }
    store4(unew + ((size_t)e * NN + node) * 4, acc);
}

// L2 mortars of a TreeMesh with refined cells (prolong2mortars! and
// calc_mortar_flux!, dg_2d.jl:1456-1646, with Trixi's
// mortar_fluxes_to_elements!): one lane per mortar and face node.  On each
// half the large element's jump state, normal flux and λ state are
// interpolated to the node (forward operator of the half) and paired with
// the small element's, the LW flux taken in the interface order (the large
// element on the left when large_side is 1) and kept for the small element;
// the large element's flux at the node is reverse_upper·F*_upper +
// reverse_lower·F*_lower.  The surface kernel lifts both.
constexpr int kMortarBlock = 64;
template <int N>
__global__ __launch_bounds__(kMortarBlock) void k_mortar(KArgs a) {
    constexpr int n = N + 1, MPB = kMortarBlock / n;
    __shared__ double fs[MPB][2][n][4];
    if (a.use_state_dt && a.st[0].finished) return;
    const DevConst* c = a.c;
    const int tid = threadIdx.x, lm = tid / n, l = tid - lm * n;
    const int m = blockIdx.x * MPB + lm;
    const bool act = lm < MPB && m < a.n_mortars;
    if (act) {
        const int2 mo = a.mort[m];
        const double* base = a.mrec + (size_t)m * 3 * n * 12;
        const double* lg = base + 2 * n * 12;
        for (int half = 0; half < 2; ++half) {
            double jl[4] = {0, 0, 0, 0}, Fl[4] = {0, 0, 0, 0}, Ul[4] = {0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < n; ++k) {
                const double p = c->mfwd[half][l][k];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    jl[v] = fma(p, lg[k * 12 + v], jl[v]);
                    Fl[v] = fma(p, lg[k * 12 + 4 + v], Fl[v]);
                    Ul[v] = fma(p, lg[k * 12 + 8 + v], Ul[v]);
                }
            }
            const double* sm = base + ((size_t)half * n + l) * 12;
            double rL[kTraceQ], rS[kTraceQ], Fs[4];
            double aL, cL, aS, cS;
            speed_parts(Ul, c->gamma, mo.x, aL, cL);
            speed_parts(sm + 8, c->gamma, mo.x, aS, cS);
            make_record(rL, jl, Fl, aL, cL);
            make_record(rS, sm, sm + 4, aS, cS);
            if (mo.y == 1) face_numflux(rL, rS, Fs);
            else face_numflux(rS, rL, Fs);
            store4(a.mflux + (((size_t)m * 3 + half) * n + l) * 4, Fs);
#pragma unroll
            for (int v = 0; v < 4; ++v) fs[lm][half][l][v] = Fs[v];
        }
    }
    __syncthreads();
    if (act) {
        double Fb[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < n; ++k) acc = fma(c->mrev[1][l][k], fs[lm][1][k][v], acc);
#pragma unroll
            for (int k = 0; k < n; ++k) acc = fma(c->mrev[0][l][k], fs[lm][0][k][v], acc);
            Fb[v] = acc;
        }
        store4(a.mflux + (((size_t)m * 3 + 2) * n + l) * 4, Fb);
    }
}

// CFL reduction over a whole state (Trixi max_dt) without an update: first dt
// of a solve (RED_INIT) or a max_dt query (RED_QUERY).
template <int N>
__global__ __launch_bounds__(kSurfBlock) void k_max_speed(KArgs a) {
    using G_ = SGeo<N>;
    constexpr int NN = G_::NN, EB = G_::EB, S = G_::S;
    __shared__ double s_l1[S], s_l2[S], s_m[EB], s_lv[S], s_mv[EB];
    const int4 blk = load_blk(a.sblk);
    DevState* st = a.st + blk.z;
    const int tid = threadIdx.x;
    const int le = tid / NN, node = tid - (tid / NN) * NN;
    const int e0 = blk.x;
    const int nel = blk.y;
    const bool act = le < nel;
    const int lec = le < nel ? le : nel - 1;
    const int e = e0 + lec;
    const double* uin = sel_in(a, st->cur);
    double val[4];
    load4(uin + ((size_t)e * NN + node) * 4, val);
    double l1, l2;
    bool bad;
    if (a.Ja) {  // curvilinear mesh: speeds in reference space, |J⁻¹| included
        double Jn[4];
        load4(a.Ja + ((size_t)e * NN + node) * 4, Jn);
        curved_speeds(val, Jn, a.ijn[(size_t)e * NN + node], a.c->gamma, l1, l2, bad);
    } else {
        node_speeds(val, a.c->gamma, l1, l2, bad);
    }
    if (act && bad) st->bad_state = 1;
    const bool par = a.pv_off > 0;
    if (act) {
        s_l1[lec * NN + node] = l1;
        s_l2[lec * NN + node] = l2;
        s_lv[lec * NN + node] = par ? (a.Ja ? lam_visc_curved(val, a.ijn[(size_t)e * NN + node], a.c)
                                            : lam_visc(val, a.inv_jac[e], a.c))
                                    : 0.0;
    }
    __syncthreads();
    if (act && node == 0) {
        double m1 = 0.0, m2 = 0.0, mv = 0.0;
        for (int k = 0; k < NN; ++k) {
            m1 = dmax(m1, s_l1[lec * NN + k]);
            m2 = dmax(m2, s_l2[lec * NN + k]);
            mv = dmax(mv, s_lv[lec * NN + k]);
        }
        s_m[lec] = a.Ja ? m1 + m2 : a.inv_jac[e] * (m1 + m2);
        s_mv[lec] = mv;
    }
    __syncthreads();
    if (tid == 0) {
        double bm = 0.0, bv = 0.0;
        for (int k = 0; k < nel; ++k) {
            bm = dmax(bm, s_m[k]);
            bv = dmax(bv, s_mv[k]);
        }
        slot_max(a.slots, blk.z, 0, bm);
        if (par) slot_max(a.slots, blk.z, 1, bv);
    }
}

// Hennemann-Gassner indicator (Trixi IndicatorHennemannGassner, Hennemann et
// al., JCP 426, 2021) on density_pressure, before smoothing: modal energy of
// the highest and second-highest modes, logistic blending with threshold
// 0.5·10^(-1.8 n^0.25) and sharpness ln((1-1e-4)/1e-4), clipped below
// alpha_min to 0, above 1-alpha_min to 1, capped at alpha_max.
template <int N>
__global__ __launch_bounds__(kSurfBlock) void k_indicator(KArgs a, double* alpha_tmp) {
    using G_ = SGeo<N>;
    constexpr int n = G_::n, NN = G_::NN, EB = G_::EB, S = G_::S;
    __shared__ double s_a[S], s_b[S], s_e[3][S];
    const int tid = threadIdx.x;
    const int le = tid / NN, node = tid - (tid / NN) * NN;
    const int i = node % n, j = node / n;
    const int4 blk = load_blk(a.sblk);
    const DevState* st = a.st + blk.z;
    const int e0 = blk.x;
    const int nel = blk.y;
    const bool act = le < nel;
    const int lec = le < nel ? le : nel - 1;
    const int e = e0 + lec;
    const DevConst* c = a.c;
    const double* uin = sel_in(a, a.use_state_dt ? st->cur : 0);
    if (a.use_state_dt && st->finished) return;
    double un[4];
    load4(uin + ((size_t)e * NN + node) * 4, un);
    const double rp = (c->gamma - 1.0) * (un[0] * un[3] - 0.5 * (un[1] * un[1] + un[2] * un[2]));
    if (act) s_a[lec * NN + node] = rp;
    __syncthreads();
    double acc = 0.0;  // x transform: tmp[j][i] = Σ_ii Vinv[i][ii] ind[j][ii]
#pragma unroll
    for (int k = 0; k < n; ++k) acc = fma(c->Vinv[i][k], s_a[lec * NN + j * n + k], acc);
    if (act) s_b[lec * NN + node] = acc;
    __syncthreads();
    double m = 0.0;  // y transform: modal[j][i] = Σ_jj Vinv[j][jj] tmp[jj][i]
#pragma unroll
    for (int k = 0; k < n; ++k) m = fma(c->Vinv[j][k], s_b[lec * NN + k * n + i], m);
    const double m2 = m * m;
    if (act) {
        s_e[0][lec * NN + node] = m2;
        s_e[1][lec * NN + node] = (i < n - 1 && j < n - 1) ? m2 : 0.0;
        s_e[2][lec * NN + node] = (i < n - 2 && j < n - 2) ? m2 : 0.0;
    }
    __syncthreads();
    if (act && node == 0) {
        double tot = 0.0, c1 = 0.0, c2 = 0.0;
        for (int k = 0; k < NN; ++k) {
            tot += s_e[0][lec * NN + k];
            c1 += s_e[1][lec * NN + k];
            c2 += s_e[2][lec * NN + k];
        }
        const double f1 = tot != 0.0 ? (tot - c1) / tot : 0.0;
        const double f2 = c1 != 0.0 ? (c1 - c2) / c1 : 0.0;
        const double energy = dmax(f1, f2);
        const double threshold = 0.5 * pow(10.0, -1.8 * pow((double)n, 0.25));
        const double par_s = log((1.0 - 0.0001) / 0.0001);
        double al = 1.0 / (1.0 + exp(-par_s / threshold * (energy - threshold)));
        if (al < c->alpha_min) al = 0.0;
        if (al > 1.0 - c->alpha_min) al = 1.0;
        alpha_tmp[e] = al < c->alpha_max ? al : c->alpha_max;
    }
}

// Zhang-Shu positivity limiter on (density, pressure), in that order
// (my_limiter_zhang_shu!, src/callbacks/callbacks_stage/positivity_zhang_shu_dg2d.jl:8-45,
// positivity_zhang_shu.jl:26-41): where the variable's minimum over an
// element's nodes is below its threshold, u <- θ u + (1-θ) ū with
// θ = (var(ū) - threshold) / (var(ū) - min var) and ū = Σ u w_i w_j / 4 summed
// in the reference's node order.  Limits a.out (explicit) or the new state of
// the device loop in place and, in the loop, takes the CFL reduction of the
// limited state (the next step's compute_dt).
template <int N>
__global__ __launch_bounds__(kSurfBlock) void k_limiter(KArgs a) {
    using G_ = SGeo<N>;
    constexpr int n = G_::n, NN = G_::NN, EB = G_::EB, S = G_::S;
    __shared__ double s_var[S], s_q[4][S];
    __shared__ double s_theta[EB], s_mean[EB][4];
    __shared__ int s_flag[EB];
    const int4 blk = load_blk(a.sblk);
    DevState* st = a.st + blk.z;
    if (a.use_state_dt && st->finished) return;
    const int tid = threadIdx.x;
    const int le = tid / NN, node = tid - (tid / NN) * NN;
    const int i = node % n, j = node / n;
    const int e0 = blk.x;
    const int nel = blk.y;
    const bool act = le < nel;
    const int lec = le < nel ? le : nel - 1;
    const int e = e0 + lec;
    const int sidx = lec * NN + node, sbase = lec * NN;
    const int cur = a.use_state_dt ? st->cur : 0;
    double* u = sel_out(a, cur);
    const DevConst* c = a.c;
    const double gm1 = c->gamma - 1.0;
    const size_t off = ((size_t)e * NN + node) * 4;
    double un[4];
    load4(u + off, un);
    bool changed = false;
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        const double thr = c->pos_threshold[pass];
        const double var = pass == 0 ? un[0] : gm1 * (un[3] - 0.5 * (un[1] * un[1] + un[2] * un[2]) / un[0]);
        if (act) {
            s_var[sidx] = var;
#pragma unroll
            for (int v = 0; v < 4; ++v) s_q[v][sidx] = un[v] * c->weights[i] * c->weights[j];
        }
        __syncthreads();
        if (act && node == 0) {
            double vmin = s_var[sbase];
            for (int k = 1; k < NN; ++k) vmin = s_var[sbase + k] < vmin ? s_var[sbase + k] : vmin;
            int flag = vmin < thr;
            if (flag) {
                double m[4] = {0.0, 0.0, 0.0, 0.0};
                for (int k = 0; k < NN; ++k)
#pragma unroll
                    for (int v = 0; v < 4; ++v) m[v] += s_q[v][sbase + k];
#pragma unroll
                for (int v = 0; v < 4; ++v) m[v] = m[v] / 4.0;
                const double vmean = pass == 0 ? m[0] : gm1 * (m[3] - 0.5 * (m[1] * m[1] + m[2] * m[2]) / m[0]);
                s_theta[le] = (vmean - thr) / (vmean - vmin);
#pragma unroll
                for (int v = 0; v < 4; ++v) s_mean[le][v] = m[v];
            }
            s_flag[le] = flag;
        }
        __syncthreads();
        if (s_flag[lec]) {
            const double th = s_theta[lec];
#pragma unroll
            for (int v = 0; v < 4; ++v) un[v] = th * un[v] + (1.0 - th) * s_mean[lec][v];
            changed = true;
        }
        __syncthreads();  // s_var, s_q, s_flag are rewritten by the next pass
    }
    if (act && changed) store4(u + off, un);
    if (a.reduce == RED_NONE) return;
    // per-element max of λ1, λ2 of the limited state (curvilinear speeds carry the
    // node's |J⁻¹|; Navier-Stokes adds the viscous bound), one partial per workgroup
    double l1, l2;
    bool bad;
    if (a.Ja) {
        double Jn[4];
        load4(a.Ja + off, Jn);
        curved_speeds(un, Jn, a.ijn[(size_t)e * NN + node], c->gamma, l1, l2, bad);
    } else {
        node_speeds(un, c->gamma, l1, l2, bad);
    }
    if (act && bad) st->bad_state = 1;
    const bool par = a.pv_off > 0;
    double* s_l1 = s_var;
    double* s_l2 = s_q[0];
    double* s_lv = s_q[1];
    double* s_mv = s_q[2];
    if (act) {
        s_l1[sidx] = l1;
        s_l2[sidx] = l2;
        s_lv[sidx] = par ? (a.Ja ? lam_visc_curved(un, a.ijn[(size_t)e * NN + node], c) : lam_visc(un, a.inv_jac[e], c))
                         : 0.0;
    }
    __syncthreads();
    if (act && node == 0) {
        double m1 = 0.0, m2 = 0.0, mv = 0.0;
        for (int k = 0; k < NN; ++k) {
            m1 = dmax(m1, s_l1[sbase + k]);
            m2 = dmax(m2, s_l2[sbase + k]);
            mv = dmax(mv, s_lv[sbase + k]);
        }
        s_theta[lec] = a.Ja ? m1 + m2 : a.inv_jac[e] * (m1 + m2);
        s_mv[lec] = mv;
    }
    __syncthreads();
    if (tid == 0) {
        double bm = 0.0, bv = 0.0;
        for (int k = 0; k < nel; ++k) {
            bm = dmax(bm, s_theta[k]);
            bv = dmax(bv, s_mv[k]);
        }
        slot_max(a.slots, blk.z, 0, bm);
        if (par) slot_max(a.slots, blk.z, 1, bv);
    }
}

// Step control of the adaptive LW loop (perform_step! for Adaptive and LW,
// adaptive.jl:48-64, 95-173): total error sqrt(Σ temporal_errors / ndofs),
// epsilon shift on a first attempt, factor 1 + atan(ε1^(β1/k) ε0^(β2/k) ε-1^(β3/k) - 1)
// with (β1, β2, β3) = (0.6, -0.2, 0) (lwfr.jl:162) and k = nnodes; the step is
// rejected when the updated state is not admissible or factor <= 0.81 (retry
// from uⁿ with min(factor, 0.95)·dt), else accepted with next dt min(factor, 1.5)·dt.
// A non-finite factor on a rejected step shrinks dt by 0.95.
__device__ inline void adaptive_control(DevState* st, double total, double k) {
    const double err = sqrt(total / st->ndofs);
    if (st->nfail == 0) {
        st->eps[0] = st->eps[1];
        st->eps[1] = st->eps[2];
    }
    st->eps[2] = 1.0 / (err + 1e-12);
    const double fac = pow(st->eps[2], 0.6 / k) * pow(st->eps[1], -0.2 / k) * pow(st->eps[0], 0.0 / k);
    const double factor = 1.0 + atan(fac - 1.0);
    const double t = st->t, tf = st->t_final;
    double dt = st->dt;
    if (st->bad_state || !(factor > 0.81)) {
        st->nfail = st->nfail + 1;
        dt = (factor < 0.95 ? factor : 0.95) * dt;
        if (t + dt > tf) dt = tf - t;  // set_dt!
        st->dt = dt;
        st->bad_state = 0;
        return;
    }
    st->nreject = st->nreject + st->nfail;
    st->nfail = 0;
    double dtn = (factor < 1.5 ? factor : 1.5) * dt;
    double dts = dt;
    if (t + dts > tf) dts = tf - t;  // set_t_and_iter!
    const double t1 = t + dts;
    st->t = t1;
    st->iter = st->iter + 1;
    st->cur = 1 - st->cur;
    if (t1 + dtn > tf) dtn = tf - t1;  // set_dt!
    st->dt = dtn;
    st->finished = julia_isapprox(t1, tf);
}

// Final reduction of the per-workgroup partial maxima of J⁻¹(max λ1 + max λ2),
// one workgroup per instance over that instance's partials (rng[inst] = volume
// begin/end, surface begin/end; the volume range is skipped when use_vol is 0),
// then compute_dt / set_dt! / set_t_and_iter! / isfinished on the instance's
// state (lwfr.jl:76-103, 146-148, 176-190; Trixi max_dt).
constexpr int kFinalizeBlock = 1024;
// Per-step dt record and prescription of a single-instance device loop: the dt
// of step k (k = the iteration count since the last solve_init) is written to
// log[k] while k < log_cap, and taken from presc[k] instead of cfl · max_dt
// while k < presc_n (the truncation at the final time still applies).
struct DtLog {
    double* log;
    long long log_cap;
    const double* presc;
    long long presc_n;
};
// With pv_off > 0 the viscous maxima (pv_off entries further on) give the
// parabolic bound 1/((λa + λv)·2(2N+1)), λv = 2(2N+1)·max J⁻² max(4μ/(3ρ), μκ/(ρPr)),
// and max_dt is the smaller of the two (compute_dt for
// SemidiscretizationHyperbolicParabolic, lwfr.jl:192-210).
// The arguments of one context's finalize launch.
struct FinArgs {
    DevState* st_all;
    const double* partial;
    const int4* rng;
    int use_vol, reduce, pv_off;
    double* slots;
    DtLog dl;
};

// Instance `bid` of the context, with LDS s, sv [kFinalizeBlock].
template <int N>
__device__ __forceinline__ void finalize_body(const FinArgs& f, int bid, double* s, double* sv) {
    constexpr int n = N + 1;
    const double* partial = f.partial;
    const int reduce = f.reduce, pv_off = f.pv_off;
    double* slots = f.slots;
    const DtLog& dl = f.dl;
    DevState* st = f.st_all + bid;
    const int4 r = f.rng[bid];
    if (reduce == RED_STEP && st->finished) return;
    if (reduce == RED_STEP && st->adaptive) {
        // sum of the volume workgroups' temporal-error partials, then the controller
        double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int k = r.x + (int)threadIdx.x;
        for (; k + 7 * kFinalizeBlock < r.y; k += 8 * kFinalizeBlock) {
#pragma unroll
            for (int q = 0; q < 8; ++q) m[q] += partial[k + q * kFinalizeBlock];
        }
        for (; k < r.y; k += kFinalizeBlock) m[0] += partial[k];
        s[threadIdx.x] = ((m[0] + m[1]) + (m[2] + m[3])) + ((m[4] + m[5]) + (m[6] + m[7]));
        __syncthreads();
        for (int w = kFinalizeBlock / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) adaptive_control(st, s[0], (double)n);
        return;
    }
    // the instance's CFL (and viscous) slots, reset for the next reduction
    {
        const int t = (int)threadIdx.x;
        double* sl = slots + (size_t)bid * 2 * kSlots;
        s[t] = 0.0;
        sv[t] = 0.0;
        if (t < kSlots) {
            s[t] = sl[t];
            sl[t] = 0.0;
            if (pv_off > 0) {
                sv[t] = sl[kSlots + t];
                sl[kSlots + t] = 0.0;
            }
        }
    }
    __syncthreads();
    for (int w = kSlots / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s[threadIdx.x] = dmax(s[threadIdx.x], s[threadIdx.x + w]);
            if (pv_off > 0) sv[threadIdx.x] = dmax(sv[threadIdx.x], sv[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double tiny = 4.9406564584124654e-324;  // nextfloat(0.0), Trixi's start value
    const double M = dmax(tiny, s[0]);
    double max_dt = 2.0 / (n * M);
    if (pv_off > 0) {
        const double lam_v = dmax(tiny, sv[0]) * (2.0 * (2.0 * N + 1.0));
        const double dtv = 1.0 / ((M + lam_v) * (2 * (2.0 * N + 1.0)));
        if (dtv < max_dt) max_dt = dtv;
    }
    if (reduce == RED_QUERY) {
        st->query = max_dt;
        return;
    }
    double t = st->t;
    if (reduce == RED_STEP) {
        t = t + st->dt;  // set_t_and_iter!
        st->t = t;
        st->iter = st->iter + 1;
        st->cur = 1 - st->cur;
    }
    const double tf = st->t_final;
    double dt = st->cfl * max_dt;  // compute_dt
    const long long it = st->iter;
    if (dl.presc && it < dl.presc_n) dt = dl.presc[it];  // prescribed dt of step `it`
    if (t + dt > tf) dt = tf - t;  // set_dt!
    st->dt = dt;
    if (dl.log && it < dl.log_cap) dl.log[it] = dt;
    st->finished = julia_isapprox(t, tf);
}

template <int N>
__global__ __launch_bounds__(kFinalizeBlock) void k_finalize(FinArgs f) {
    __shared__ double s[kFinalizeBlock], sv[kFinalizeBlock];
    finalize_body<N>(f, blockIdx.x, s, sv);
}

// Two contexts advanced by the same launches (tlw_create_ensemble): one of
// polydeg 3 (workgroups [0, g3) of each launch) and one of polydeg 4, plain LW
// on periodic meshes; each half runs its own context's kernel body on its own
// tables, so every instance gets the bits of its own context's launches.
template <int EB3, int EB4>
__global__ __launch_bounds__((VGeo<4, EB4>::NT), kVolWaves<false>) void k_volume_pair(KArgs a3, KArgs a4, int g3) {
    static_assert(VGeo<3, EB3>::NT == VGeo<4, EB4>::NT, "one block size for both halves");
    constexpr int L = VGeo<3, EB3>::LDS > VGeo<4, EB4>::LDS ? VGeo<3, EB3>::LDS : VGeo<4, EB4>::LDS;
    __shared__ double lds[L];
    __shared__ double sDm[25], sDh[25], sD[25], s_alpha[1];
    if ((int)blockIdx.x < g3)
        volume_body<3, EB3, VOL_LW, false, false, false>(a3, blockIdx.x, lds, sDm, sDh, sD, s_alpha);
    else
        volume_body<4, EB4, VOL_LW, false, false, false>(a4, blockIdx.x - g3, lds, sDm, sDh, sD, s_alpha);
}

__global__ __launch_bounds__(kSurfBlock) void k_surface_pair(KArgs a3, KArgs a4, int g3) {
    if ((int)blockIdx.x < g3)
        surface_body<3, false, false, false>(a3, blockIdx.x);
    else
        surface_body<4, false, false, false>(a4, blockIdx.x - g3);
}

__global__ __launch_bounds__(kFinalizeBlock) void k_finalize_pair(FinArgs f3, FinArgs f4, int n3) {
    __shared__ double s[kFinalizeBlock], sv[kFinalizeBlock];
    if ((int)blockIdx.x < n3)
        finalize_body<3>(f3, blockIdx.x, s, sv);
    else
        finalize_body<4>(f4, blockIdx.x - n3, s, sv);
}

// Initialisation of the device integrator state of every instance.
// adaptive: dt_initial, tolerances and epsilon = (1, 1, 1) of the adaptive loop
// (solve_lwfr's set_dt!(integrator, dt_initial); LWIntegrator, lwfr.jl:162-168)
__global__ void k_init_state(DevState* st_all, int ninst, double t0, double tf, double cfl, const int* cur,
                             int adaptive, double dt0, double abstol, double reltol) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ninst) return;
    DevState* st = st_all + k;
    st->t = t0;
    st->dt = 0.0;
    st->adaptive = adaptive;
    st->nfail = 0;
    st->nreject = 0;
    st->eps[0] = st->eps[1] = st->eps[2] = 1.0;
    st->abstol = abstol;
    st->reltol = reltol;
    if (adaptive) st->dt = (t0 + dt0 > tf) ? tf - t0 : dt0;
    st->t_final = tf;
    st->cfl = cfl;
    st->query = 0.0;
    st->finished = julia_isapprox(t0, tf);
    st->cur = cur[k];
    st->bad_state = 0;
    st->iter = 0;
}

__global__ void k_clear_bad_state(DevState* st_all, int ninst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ninst) st_all[k].bad_state = 0;
}

}  // namespace tlw
