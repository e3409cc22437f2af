// tlw_oracle_parabolic.hpp -- CPU restatement of TrixiLW's hyperbolic-parabolic
// Lax-Wendroff / MDRK right-hand side on a periodic 2-D TreeMesh.
//
// TEST INFRASTRUCTURE ONLY (the parity oracle and CPU baseline of the
// parabolic path); the product never links or calls it.
//
// Reference (paths relative to the TrixiLW.jl tree):
//   rhs!                                   src/solvers/dgsem_tree/dg_2d_parabolic.jl:15-98
//   lw_volume_kernel_3! / 4!               dg_2d_parabolic.jl:725-1189, 1191-1860
//   prolong2interfaces_lw_parabolic!       dg_2d_parabolic.jl:1862-1909
//   calc_interface_flux_hyperbolic_parabolic!  dg_2d_parabolic.jl:1911-1948
//   rhs_mdrk1! / rhs_mdrk2!                src/solvers/dgsem_tree/multi_derivative_2d_parabolic.jl:59-219
//   mdrk_kernel_1! / 2!                    multi_derivative_2d_parabolic.jl:221-464, 506-711
//   compute_dt (hyperbolic-parabolic)      src/solvers/lwfr.jl:192-210
//   max_dt(..., LaplaceDiffusion2D)        src/equations/laplace_diffusion_2d.jl:4-43
//   max_dt(..., CompressibleNavierStokesDiffusion2D)  src/equations/compressible_navier_stokes_2d.jl:204-261
//   GradientVariablesConservative          compressible_navier_stokes_2d.jl:14-57
// and, from the un-vendored Trixi.jl (restated from their published
// definitions): calc_gradient! (BR1: weak-form D̂ volume term, central
// interface value ½(u_L + u_R), lifting with boundary_interpolation,
// apply_jacobian_parabolic! = ×J⁻¹), calc_viscous_fluxes!, the viscous
// flux of LaplaceDiffusion2D (κ∇u) and of CompressibleNavierStokesDiffusion2D
// (τ, Fourier heat flux with κ = γ/((γ-1) Pr), constant μ).
//
// How each piece is pinned (tests/test_oracle_parabolic_pins.py): the linear
// advection-diffusion step against a dense closed form (BR1 gradient, Taylor
// states of ∇u, F̃a/F̃v split, interface flux), its convergence to the exact
// decaying wave, the Navier-Stokes viscous flux against the chain rule
// through finite differences of (ρ, v1, v2, T), μ = 0 against the hyperbolic
// solver, free-stream preservation and conservation.
//
// Scope: periodic meshes (the parabolic boundary conditions are not
// restated), VolumeIntegralFR, no source terms (`nothing`).
#pragma once
#include "tlw_oracle.hpp"

namespace tlwo {

// Trixi LaplaceDiffusion2D: f_v = κ ∇u for every variable.
template <class T> struct LaplaceDiffusion2D {
    double diffusivity = 0.0;
    template <int NV> void fluxes(const T* u, const T* gx, const T* gy, T* fx, T* fy) const {
        for (int v = 0; v < NV; ++v) {
            fx[v] = diffusivity * gx[v];
            fy[v] = diffusivity * gy[v];
        }
    }
    // laplace_diffusion_2d.jl:29-33: J⁻² κ
    double lam_v(const T*, double inv_jac) const { return inv_jac * inv_jac * diffusivity; }
};

// Trixi CompressibleNavierStokesDiffusion2D with GradientVariablesConservative:
// the gradients are those of the conservative variables; primitive gradients
// follow from convert_derivative_to_primitive (compressible_navier_stokes_2d.jl:37-57).
template <class T> struct NavierStokesDiffusion2D {
    double gamma = 1.4, mu = 0.0, Pr = 0.72, kappa = 0.0;
    void set(double g, double mu_, double pr) {
        gamma = g;
        mu = mu_;
        Pr = pr;
        kappa = gamma * (1.0 / (gamma - 1.0)) / Pr;
    }
    // cons2prim of the parabolic equations: (ρ, v1, v2, T = p/ρ)
    void cons2prim(const T* u, T& rho, T& v1, T& v2, T& Tt) const {
        rho = u[0];
        v1 = u[1] / rho;
        v2 = u[2] / rho;
        const T p = (gamma - 1.0) * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]) / rho);
        Tt = p / rho;
    }
    void prim_derivative(const T* u, const T* W, T& v1d, T& v2d, T& Td) const {
        T rho, v1, v2, Tt;
        cons2prim(u, rho, v1, v2, Tt);
        const T rhod = W[0];
        v1d = (W[1] - rhod * v1) / rho;
        v2d = (W[2] - rhod * v2) / rho;
        const T pd = (W[3] - 0.5 * rhod * (v1 * v1 + v2 * v2) - rho * (v1 * v1d + v2 * v2d)) * (gamma - 1.0);
        Td = (pd - rhod * Tt) / rho;
    }
    template <int NV> void fluxes(const T* u, const T* gx, const T* gy, T* fx, T* fy) const {
        T rho, v1, v2, Tt;
        cons2prim(u, rho, v1, v2, Tt);
        T dv1dx, dv2dx, dTdx, dv1dy, dv2dy, dTdy;
        prim_derivative(u, gx, dv1dx, dv2dx, dTdx);
        prim_derivative(u, gy, dv1dy, dv2dy, dTdy);
        const T tau11 = 4.0 / 3.0 * dv1dx - 2.0 / 3.0 * dv2dy;
        const T tau12 = dv1dy + dv2dx;
        const T tau22 = 4.0 / 3.0 * dv2dy - 2.0 / 3.0 * dv1dx;
        const T q1 = kappa * dTdx, q2 = kappa * dTdy;
        fx[0] = T(0.0);
        fx[1] = tau11 * mu;
        fx[2] = tau12 * mu;
        fx[3] = (v1 * tau11 + v2 * tau12 + q1) * mu;
        fy[0] = T(0.0);
        fy[1] = tau12 * mu;
        fy[2] = tau22 * mu;
        fy[3] = (v1 * tau12 + v2 * tau22 + q2) * mu;
    }
    // compressible_navier_stokes_2d.jl:230-237: J⁻² max(4μ/(3ρ), μκ/(ρ Pr))
    double lam_v(const T* u, double inv_jac) const { return lam_v_scaled(u, inv_jac * inv_jac); }
    // s max(4μ/(3ρ), μκ/(ρ Pr)); s = J⁻¹ of the node on a curvilinear mesh
    double lam_v_scaled(const T* u, double s) const {
        const double rho = value(u[0]);
        return s * std::max(4.0 * mu / (3.0 * rho), mu * kappa / (rho * Pr));
    }
};

template <class T, class Eq, class Visc> struct ParabolicSolver : Solver<T, Eq> {
    using B = Solver<T, Eq>;
    static constexpr int NV = Eq::NV;
    using B::basis;
    using B::eq;
    using B::fidx;
    using B::idx;
    using B::mesh;
    using B::n;
    using B::sidx;
    Visc visc;
    std::vector<T> gx, gy;   // gradients [v, i, j, e] (viscous_container.gradients)
    std::vector<T> gsfv;     // surface values of the gradient computation [v, l, face, e]
    std::vector<T> Fv;       // cache_parabolic.Fv [v, d, i, j, e]
    std::vector<T> Fv2;      // cache_parabolic.mdrk_cache.Fv2

    // Curvilinear mesh (set by CurvedParabolicSolver): per node the contravariant
    // vectors [e][j][i][4] (Ja11, Ja12, Ja21, Ja22) and J⁻¹ [e][j][i]; null on a TreeMesh.
    const double* cJa = nullptr;
    const double* cijn = nullptr;
    size_t cnode(int i, int j, int e) const { return ((size_t)e * n + j) * n + i; }
    double node_ij(int i, int j, int e) const { return cijn ? cijn[cnode(i, j, e)] : mesh.inv_jac[e]; }

    // No-slip adiabatic walls on every boundary face (Euler with the reflecting wall
    // for the hyperbolic part): wall velocity per direction x_neg, x_pos, y_neg, y_pos
    int ns_walls = 0;
    double wall_v[4][2] = {};
    // heat condition per direction: 0 Adiabatic (zero normal heat flux), 1 Isothermal
    // with the wall temperature wall_T (T = p/ρ, the reference's R = 1)
    int wall_iso[4] = {0, 0, 0, 0};
    double wall_T[4] = {0, 0, 0, 0};

    void setup_parabolic() {
        const size_t nd = this->ndofs_total();
        gx.assign(nd, T(0));
        gy.assign(nd, T(0));
        gsfv.assign((size_t)mesh.E * 4 * n * NV, T(0));
        Fv.assign(2 * nd, T(0));
    }
    const std::vector<T>* extra_aux(int which) const override {
        switch (which) {
            case 9: return &gx;
            case 10: return &gy;
            case 11: return &Fv;
            case 12: return &Fv2;
        }
        return nullptr;
    }

    // Trixi calc_gradient! on a periodic TreeMesh (BR1): volume term with D̂,
    // interface value ½(u_L + u_R), surface lifting with
    // boundary_interpolation[1,1] and [n,2], then ×J⁻¹ (apply_jacobian_parabolic!).
    void calc_gradient(const T* u) {
        const auto& Dh = basis.Dhat;
        const long long nd = (long long)this->ndofs_total();
#pragma omp parallel for schedule(static)
        for (long long k = 0; k < nd; ++k) gx[k] = gy[k] = T(0);
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const T* un = &u[idx(0, i, j, e)];
                    for (int ii = 0; ii < n; ++ii)
                        for (int v = 0; v < NV; ++v) gx[idx(v, ii, j, e)] = gx[idx(v, ii, j, e)] + Dh[ii][i] * un[v];
                    for (int jj = 0; jj < n; ++jj)
                        for (int v = 0; v < NV; ++v) gy[idx(v, i, jj, e)] = gy[idx(v, i, jj, e)] + Dh[jj][j] * un[v];
                }
        const int nI = (int)mesh.if_left.size();
#pragma omp parallel for schedule(static)
        for (int I = 0; I < nI; ++I) {
            const int L = mesh.if_left[I], R = mesh.if_right[I], o = mesh.if_orient[I];
            const int ldir = 2 * o - 1, rdir = 2 * o - 2;
            for (int l = 0; l < n; ++l) {
                int li, lj, ri, rj;
                if (o == 1) { li = n - 1; lj = l; ri = 0; rj = l; }
                else { li = l; lj = n - 1; ri = l; rj = 0; }
                for (int v = 0; v < NV; ++v) {
                    const T fl = 0.5 * (u[idx(v, li, lj, L)] + u[idx(v, ri, rj, R)]);
                    gsfv[sidx(v, l, ldir, L)] = fl;
                    gsfv[sidx(v, l, rdir, R)] = fl;
                }
            }
        }
        if (ns_walls) {
            // Trixi calc_boundary_flux_by_direction_gradient!: the boundary value of the
            // gradient variables is the wall condition's u_outer; for
            // BoundaryConditionNavierStokesWall{NoSlip, Adiabatic} with conservative
            // gradient variables prim2cons(ρ, v_wall, p) of the inner trace
            // (compressible_navier_stokes_2d.jl:113-124)
            const int nB = (int)mesh.bd_elem.size();
#pragma omp parallel for schedule(static)
            for (int B = 0; B < nB; ++B) {
                const int e = mesh.bd_elem[B], o = mesh.bd_orient[B], side = mesh.bd_side[B], dir = mesh.bd_dir[B];
                for (int l = 0; l < n; ++l) {
                    int ii, jj;
                    if (o == 1) { ii = (side == 1) ? n - 1 : 0; jj = l; }
                    else { ii = l; jj = (side == 1) ? n - 1 : 0; }
                    T uo[NV];
                    wall_state(&u[idx(0, ii, jj, e)], wall_v[dir - 1], uo, wall_iso[dir - 1], wall_T[dir - 1]);
                    for (int v = 0; v < NV; ++v) gsfv[sidx(v, l, dir - 1, e)] = uo[v];
                }
            }
        }
        const double f1 = basis.bnd_lo, f2 = basis.bnd_hi;
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) {
            for (int l = 0; l < n; ++l)
                for (int v = 0; v < NV; ++v) {
                    gx[idx(v, 0, l, e)] = gx[idx(v, 0, l, e)] - gsfv[sidx(v, l, 0, e)] * f1;
                    gx[idx(v, n - 1, l, e)] = gx[idx(v, n - 1, l, e)] + gsfv[sidx(v, l, 1, e)] * f2;
                    gy[idx(v, l, 0, e)] = gy[idx(v, l, 0, e)] - gsfv[sidx(v, l, 2, e)] * f1;
                    gy[idx(v, l, n - 1, e)] = gy[idx(v, l, n - 1, e)] + gsfv[sidx(v, l, 3, e)] * f2;
                }
            const double factor = mesh.inv_jac[e];
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) {
                        gx[idx(v, i, j, e)] = gx[idx(v, i, j, e)] * factor;
                        gy[idx(v, i, j, e)] = gy[idx(v, i, j, e)] * factor;
                    }
        }
    }

    // Element-local gradient of a time derivative: wx = J⁻¹ D w (x), wy = J⁻¹ D w (y)
    // ("Compute ∇u_t", dg_2d_parabolic.jl:832-859); on a curvilinear mesh
    // compute_local_grad_u! (p4est/dg_2d_parabolic.jl:51-104): the reference-space
    // derivatives D w, turned with the transposed contravariant vectors,
    // wx = Ja11 w_ξ + Ja21 w_η, wy = Ja12 w_ξ + Ja22 w_η, then ×J⁻¹ of the node.
    void local_grad(const T (*w)[MAXN][NV], T (*wx)[MAXN][NV], T (*wy)[MAXN][NV], int e) const {
        const auto& D = basis.D;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) wx[j][i][v] = wy[j][i][v] = T(0);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) wx[j][ii][v] = wx[j][ii][v] + D[ii][i] * w[j][i][v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) wy[jj][i][v] = wy[jj][i][v] + D[jj][j] * w[j][i][v];
            }
        if (cJa) {
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const double* J = &cJa[4 * cnode(i, j, e)];
                    for (int v = 0; v < NV; ++v) {
                        const T g1 = wx[j][i][v], g2 = wy[j][i][v];
                        wx[j][i][v] = J[0] * g1 + J[2] * g2;
                        wy[j][i][v] = J[1] * g1 + J[3] * g2;
                    }
                }
        }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const double ij = node_ij(i, j, e);
                for (int v = 0; v < NV; ++v) {
                    wx[j][i][v] = wx[j][i][v] * ij;
                    wy[j][i][v] = wy[j][i][v] * ij;
                }
            }
    }
    // acc[ii][j] += c D[ii][i] f[j][i] (x) and c D[jj][j] g[j][i] (y); on a
    // curvilinear mesh of the contravariant pair (Ja¹·(f, g), Ja²·(f, g)) of the
    // node (contravariant_fluxes, p4est/dg_2d_parabolic.jl:4-38)
    void add_divergence(const T (*f)[MAXN][NV], const T (*g)[MAXN][NV], double c, T (*acc)[MAXN][NV], int e) const {
        const auto& D = basis.D;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                T fc[NV], gc[NV];
                if (cJa) {
                    const double* J = &cJa[4 * cnode(i, j, e)];
                    for (int v = 0; v < NV; ++v) {
                        fc[v] = J[0] * f[j][i][v] + J[1] * g[j][i][v];
                        gc[v] = J[2] * f[j][i][v] + J[3] * g[j][i][v];
                    }
                } else {
                    for (int v = 0; v < NV; ++v) {
                        fc[v] = f[j][i][v];
                        gc[v] = g[j][i][v];
                    }
                }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) acc[j][ii][v] = acc[j][ii][v] + (c * D[ii][i]) * fc[v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) acc[jj][i][v] = acc[jj][i][v] + (c * D[jj][j]) * gc[v];
            }
    }

    // Taylor states u±, u±± of u and of ∇u (upx, umx, ... in the reference)
    struct States {
        T s[4][MAXN][MAXN][NV];   // up, um, upp, umm
        T sx[4][MAXN][MAXN][NV];
        T sy[4][MAXN][MAXN][NV];
    };
    void add_states(States& S, const T (*w)[MAXN][NV], const T (*wx)[MAXN][NV], const T (*wy)[MAXN][NV],
                    const double c[4]) const {
        for (int q = 0; q < 4; ++q)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) {
                        S.s[q][j][i][v] = S.s[q][j][i][v] + c[q] * w[j][i][v];
                        S.sx[q][j][i][v] = S.sx[q][j][i][v] + c[q] * wx[j][i][v];
                        S.sy[q][j][i][v] = S.sy[q][j][i][v] + c[q] * wy[j][i][v];
                    }
    }
    // advective and viscous fluxes of state q at node (i, j)
    void state_fluxes(const States& S, int q, int i, int j, T* fa, T* ga, T* fv, T* gv) const {
        eq.fluxes(S.s[q][j][i], fa, ga);
        visc.template fluxes<NV>(S.s[q][j][i], S.sx[q][j][i], S.sy[q][j][i], fv, gv);
    }

    // Source terms of the hyperbolic-parabolic kernels (dg_2d_parabolic.jl:725-1860,
    // multi_derivative_2d_parabolic.jl:221-711; only sources that do not depend on u
    // are accepted with parabolic terms, but the stencil states are passed as the
    // reference passes them).  Level `order` of the Taylor procedure: the source's
    // time difference of that order at the states of the level before (s at uⁿ; s_t
    // of calc_source_t_N34; s_tt five-point, calc_source_tt_N4, in kernel 4 and
    // three-point, N23, in kernel 3; s_ttt of N34; s_tttt of N4), added as dt·s to the
    // level's time derivative w before its gradient is formed (w may be null), and
    // as wS·s to the element's S.
    using PNodeArr = T[MAXN][MAXN][NV];
    void src_level(int e, double dt, int order, const T* u, const States& St, T (*w)[MAXN][NV], PNodeArr& Sacc,
                   double wS, bool first) const {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double x, y;
                this->node_xy(i, j, e, x, y);
                const T* un = &u[idx(0, i, j, e)];
                const T *up = St.s[0][j][i], *um = St.s[1][j][i], *upp = St.s[2][j][i], *umm = St.s[3][j][i];
                T sv[NV];
                if (order == 0) {
                    this->src_at(un, x, y, 0.0, sv);
                } else if (order == 1) {
                    this->src_t_N34(up, upp, um, umm, x, y, dt, this->src_t, sv);
                } else if (order == 2 && basis.N == 4) {
                    T a[NV], b[NV], z[NV], c[NV], d[NV];
                    this->src_at(upp, x, y, 2.0 * dt, a);
                    this->src_at(up, x, y, dt, b);
                    this->src_at(un, x, y, 0.0, z);
                    this->src_at(um, x, y, -dt, c);
                    this->src_at(umm, x, y, -2.0 * dt, d);
                    for (int v = 0; v < NV; ++v)
                        sv[v] = (1.0 / 12.0) * (-a[v] + 16.0 * b[v] - 30.0 * z[v] + 16.0 * c[v] - d[v]);
                } else if (order == 2) {
                    this->src_tt_N23(un, up, um, x, y, dt, sv);
                } else if (order == 3) {
                    this->src_ttt_N34(up, upp, um, umm, x, y, dt, sv);
                } else {
                    this->src_tttt_N4(un, up, upp, um, umm, x, y, dt, sv);
                }
                for (int v = 0; v < NV; ++v) {
                    Sacc[j][i][v] = first ? wS * sv[v] : Sacc[j][i][v] + wS * sv[v];
                    if (w) w[j][i][v] = w[j][i][v] + dt * sv[v];
                }
            }
    }
    // du += -S/J⁻¹, which the Jacobian step turns into +S
    void src_to_du(T* du, int e, const PNodeArr& Sacc) const {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v)
                    du[idx(v, i, j, e)] = du[idx(v, i, j, e)] + (-1.0 / node_ij(i, j, e)) * Sacc[j][i][v];
    }

    // lw_volume_kernel_3!/4! (dg_2d_parabolic.jl:725-1189, 1191-1860), with the
    // source terms of src_level.  Kernel 4 on the TreeMesh also gives the element's
    // temporal error (1749-1775, 1841-1855): u_np1 = u - dt J⁻¹ D F̃ with the
    // total F̃ = F̃a - F̃v, u_np1_low the same before the f_tttt terms, compared
    // as in the Euler kernel; u_np1 - u_np1_low is formed as the update of the
    // f_tttt term itself.
    void lw_volume_parabolic(T* du, const T* u, double dt, int e) {
        const int N = basis.N;
        static thread_local States S;
        T fa[MAXN][MAXN][NV], ga[MAXN][MAXN][NV], fv[MAXN][MAXN][NV], gv[MAXN][MAXN][NV];
        T Fa[MAXN][MAXN][NV], Ga[MAXN][MAXN][NV], Fva[MAXN][MAXN][NV], Gva[MAXN][MAXN][NV];
        T Uc[MAXN][MAXN][NV], w[MAXN][MAXN][NV], wx[MAXN][MAXN][NV], wy[MAXN][MAXN][NV];
        T df[MAXN][MAXN][NV], dg[MAXN][MAXN][NV];
        const bool has_src = this->src.kind != SRC_NONE;
        PNodeArr Sacc;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) w[j][i][v] = T(0);
        // fluxes at uⁿ (advective from u, viscous from calc_viscous_fluxes!), ut
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const T* un = &u[idx(0, i, j, e)];
                const T* ux = &gx[idx(0, i, j, e)];
                const T* uy = &gy[idx(0, i, j, e)];
                eq.fluxes(un, fa[j][i], ga[j][i]);
                visc.template fluxes<NV>(un, ux, uy, fv[j][i], gv[j][i]);
                for (int v = 0; v < NV; ++v) {
                    df[j][i][v] = fa[j][i][v] - fv[j][i][v];
                    dg[j][i][v] = ga[j][i][v] - gv[j][i][v];
                    Fa[j][i][v] = fa[j][i][v];
                    Ga[j][i][v] = ga[j][i][v];
                    Fva[j][i][v] = fv[j][i][v];
                    Gva[j][i][v] = gv[j][i][v];
                    Uc[j][i][v] = un[v];
                    for (int q = 0; q < 4; ++q) {
                        S.s[q][j][i][v] = un[v];
                        S.sx[q][j][i][v] = ux[v];
                        S.sy[q][j][i][v] = uy[v];
                    }
                }
            }
        add_divergence(df, dg, -dt, w, e);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) w[j][i][v] = w[j][i][v] * node_ij(i, j, e);  // ut
        if (has_src) src_level(e, dt, 0, u, S, w, Sacc, 1.0, true);
        local_grad(w, wx, wy, e);
        {
            const double c[4] = {1.0, -1.0, 2.0, -2.0};
            add_states(S, w, wx, wy, c);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) Uc[j][i][v] = Uc[j][i][v] + 0.5 * w[j][i][v];
        }
        // stage 1: f_t = (-f(u++) + 8 f(u+) - 8 f(u-) + f(u--)) / 12; utt = -dt J⁻¹ D f_t
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                T fpa[NV], gpa[NV], fpv[NV], gpv[NV], fma[NV], gma[NV], fmv[NV], gmv[NV];
                T fppa[NV], gppa[NV], fppv[NV], gppv[NV], fmma[NV], gmma[NV], fmmv[NV], gmmv[NV];
                state_fluxes(S, 1, i, j, fma, gma, fmv, gmv);
                state_fluxes(S, 0, i, j, fpa, gpa, fpv, gpv);
                state_fluxes(S, 3, i, j, fmma, gmma, fmmv, gmmv);
                state_fluxes(S, 2, i, j, fppa, gppa, fppv, gppv);
                for (int v = 0; v < NV; ++v) {
                    const T fta = 1.0 / 12.0 * (-fppa[v] + 8.0 * fpa[v] - 8.0 * fma[v] + fmma[v]);
                    const T gta = 1.0 / 12.0 * (-gppa[v] + 8.0 * gpa[v] - 8.0 * gma[v] + gmma[v]);
                    const T ftv = 1.0 / 12.0 * (-fppv[v] + 8.0 * fpv[v] - 8.0 * fmv[v] + fmmv[v]);
                    const T gtv = 1.0 / 12.0 * (-gppv[v] + 8.0 * gpv[v] - 8.0 * gmv[v] + gmmv[v]);
                    Fa[j][i][v] = Fa[j][i][v] + 0.5 * fta;
                    Ga[j][i][v] = Ga[j][i][v] + 0.5 * gta;
                    Fva[j][i][v] = Fva[j][i][v] + 0.5 * ftv;
                    Gva[j][i][v] = Gva[j][i][v] + 0.5 * gtv;
                    df[j][i][v] = fta - ftv;
                    dg[j][i][v] = gta - gtv;
                }
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) w[j][i][v] = T(0);
        add_divergence(df, dg, 1.0, w, e);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) w[j][i][v] = w[j][i][v] * (-dt * node_ij(i, j, e));  // utt
        if (has_src) src_level(e, dt, 1, u, S, w, Sacc, 0.5, false);
        local_grad(w, wx, wy, e);
        {
            const double c[4] = {0.5, 0.5, 2.0, 2.0};
            add_states(S, w, wx, wy, c);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) Uc[j][i][v] = Uc[j][i][v] + (1.0 / 6.0) * w[j][i][v];
        }
        // stage 2: N=4 f_tt = (-f(u++) + 16 f(u+) - 30 f + 16 f(u-) - f(u--)) / 12,
        //          N=3 f_tt = f(u+) - 2 f + f(u-); uttt = -dt J⁻¹ D f_tt
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                T fpa[NV], gpa[NV], fpv[NV], gpv[NV], fma[NV], gma[NV], fmv[NV], gmv[NV];
                state_fluxes(S, 1, i, j, fma, gma, fmv, gmv);
                state_fluxes(S, 0, i, j, fpa, gpa, fpv, gpv);
                T ftta[NV], gtta[NV], fttv[NV], gttv[NV];
                if (N == 4) {
                    T fppa[NV], gppa[NV], fppv[NV], gppv[NV], fmma[NV], gmma[NV], fmmv[NV], gmmv[NV];
                    state_fluxes(S, 3, i, j, fmma, gmma, fmmv, gmmv);
                    state_fluxes(S, 2, i, j, fppa, gppa, fppv, gppv);
                    for (int v = 0; v < NV; ++v) {
                        ftta[v] = (1.0 / 12.0) * (-fppa[v] + 16.0 * fpa[v] - 30.0 * fa[j][i][v] + 16.0 * fma[v] - fmma[v]);
                        fttv[v] = (1.0 / 12.0) * (-fppv[v] + 16.0 * fpv[v] - 30.0 * fv[j][i][v] + 16.0 * fmv[v] - fmmv[v]);
                        gtta[v] = (1.0 / 12.0) * (-gppa[v] + 16.0 * gpa[v] - 30.0 * ga[j][i][v] + 16.0 * gma[v] - gmma[v]);
                        gttv[v] = (1.0 / 12.0) * (-gppv[v] + 16.0 * gpv[v] - 30.0 * gv[j][i][v] + 16.0 * gmv[v] - gmmv[v]);
                    }
                } else {
                    for (int v = 0; v < NV; ++v) {
                        ftta[v] = fpa[v] - 2.0 * fa[j][i][v] + fma[v];
                        gtta[v] = gpa[v] - 2.0 * ga[j][i][v] + gma[v];
                        fttv[v] = fpv[v] - 2.0 * fv[j][i][v] + fmv[v];
                        gttv[v] = gpv[v] - 2.0 * gv[j][i][v] + gmv[v];
                    }
                }
                for (int v = 0; v < NV; ++v) {
                    Fa[j][i][v] = Fa[j][i][v] + (1.0 / 6.0) * ftta[v];
                    Fva[j][i][v] = Fva[j][i][v] + (1.0 / 6.0) * fttv[v];
                    Ga[j][i][v] = Ga[j][i][v] + (1.0 / 6.0) * gtta[v];
                    Gva[j][i][v] = Gva[j][i][v] + (1.0 / 6.0) * gttv[v];
                    df[j][i][v] = ftta[v] - fttv[v];
                    dg[j][i][v] = gtta[v] - gttv[v];
                }
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) w[j][i][v] = T(0);
        add_divergence(df, dg, -dt, w, e);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) w[j][i][v] = w[j][i][v] * node_ij(i, j, e);  // uttt
        if (has_src) src_level(e, dt, 2, u, S, w, Sacc, 1.0 / 6.0, false);
        local_grad(w, wx, wy, e);
        {
            const double c[4] = {1.0 / 6.0, -1.0 / 6.0, 4.0 / 3.0, -4.0 / 3.0};
            add_states(S, w, wx, wy, c);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) Uc[j][i][v] = Uc[j][i][v] + (1.0 / 24.0) * w[j][i][v];
        }
        // stage 3: f_ttt = (f(u++) - 2 f(u+) + 2 f(u-) - f(u--)) / 2
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                T fpa[NV], gpa[NV], fpv[NV], gpv[NV], fma[NV], gma[NV], fmv[NV], gmv[NV];
                T fppa[NV], gppa[NV], fppv[NV], gppv[NV], fmma[NV], gmma[NV], fmmv[NV], gmmv[NV];
                state_fluxes(S, 1, i, j, fma, gma, fmv, gmv);
                state_fluxes(S, 0, i, j, fpa, gpa, fpv, gpv);
                state_fluxes(S, 3, i, j, fmma, gmma, fmmv, gmmv);
                state_fluxes(S, 2, i, j, fppa, gppa, fppv, gppv);
                for (int v = 0; v < NV; ++v) {
                    const T fttta = 0.5 * (fppa[v] - 2.0 * fpa[v] + 2.0 * fma[v] - fmma[v]);
                    const T ftttv = 0.5 * (fppv[v] - 2.0 * fpv[v] + 2.0 * fmv[v] - fmmv[v]);
                    const T gttta = 0.5 * (gppa[v] - 2.0 * gpa[v] + 2.0 * gma[v] - gmma[v]);
                    const T gtttv = 0.5 * (gppv[v] - 2.0 * gpv[v] + 2.0 * gmv[v] - gmmv[v]);
                    Fa[j][i][v] = Fa[j][i][v] + (1.0 / 24.0) * fttta;
                    Fva[j][i][v] = Fva[j][i][v] + (1.0 / 24.0) * ftttv;
                    Ga[j][i][v] = Ga[j][i][v] + (1.0 / 24.0) * gttta;
                    Gva[j][i][v] = Gva[j][i][v] + (1.0 / 24.0) * gtttv;
                    df[j][i][v] = fttta - ftttv;
                    dg[j][i][v] = gttta - gtttv;
                }
            }
        if (N == 4) {
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) w[j][i][v] = T(0);
            add_divergence(df, dg, -dt, w, e);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) w[j][i][v] = w[j][i][v] * node_ij(i, j, e);  // utttt
            if (has_src) src_level(e, dt, 3, u, S, w, Sacc, 1.0 / 24.0, false);
            local_grad(w, wx, wy, e);
            const double c[4] = {1.0 / 24.0, 1.0 / 24.0, 2.0 / 3.0, 2.0 / 3.0};
            add_states(S, w, wx, wy, c);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) Uc[j][i][v] = Uc[j][i][v] + (1.0 / 120.0) * w[j][i][v];
            // stage 4: f_tttt = 0.5 (f(u++) - 4 f(u+) + 6 f - 4 f(u-) + f(u--)), the
            // factor 0.5 as in the reference (dg_2d_parabolic.jl:1746-1747)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    T fpa[NV], gpa[NV], fpv[NV], gpv[NV], fma[NV], gma[NV], fmv[NV], gmv[NV];
                    T fppa[NV], gppa[NV], fppv[NV], gppv[NV], fmma[NV], gmma[NV], fmmv[NV], gmmv[NV];
                    state_fluxes(S, 1, i, j, fma, gma, fmv, gmv);
                    state_fluxes(S, 0, i, j, fpa, gpa, fpv, gpv);
                    state_fluxes(S, 3, i, j, fmma, gmma, fmmv, gmmv);
                    state_fluxes(S, 2, i, j, fppa, gppa, fppv, gppv);
                    for (int v = 0; v < NV; ++v) {
                        const T a4 = 0.5 * (fppa[v] - 4.0 * fpa[v] + 6.0 * fa[j][i][v] - 4.0 * fma[v] + fmma[v]);
                        const T v4 = 0.5 * (fppv[v] - 4.0 * fpv[v] + 6.0 * fv[j][i][v] - 4.0 * fmv[v] + fmmv[v]);
                        const T b4 = 0.5 * (gppa[v] - 4.0 * gpa[v] + 6.0 * ga[j][i][v] - 4.0 * gma[v] + gmma[v]);
                        const T w4 = 0.5 * (gppv[v] - 4.0 * gpv[v] + 6.0 * gv[j][i][v] - 4.0 * gmv[v] + gmmv[v]);
                        Fa[j][i][v] = Fa[j][i][v] + (1.0 / 120.0) * a4;
                        Fva[j][i][v] = Fva[j][i][v] + (1.0 / 120.0) * v4;
                        Ga[j][i][v] = Ga[j][i][v] + (1.0 / 120.0) * b4;
                        Gva[j][i][v] = Gva[j][i][v] + (1.0 / 120.0) * w4;
                        df[j][i][v] = (1.0 / 120.0) * (a4 - v4);  // the f_tttt terms of F̃
                        dg[j][i][v] = (1.0 / 120.0) * (b4 - w4);
                    }
                }
            if (has_src) src_level(e, dt, 4, u, S, nullptr, Sacc, 1.0 / 120.0, false);
            if (!cJa) temporal_error(u, dt, e, Fa, Ga, Fva, Gva, df, dg);
        } else if (has_src) {
            src_level(e, dt, 3, u, S, nullptr, Sacc, 1.0 / 24.0, false);  // kernel 3: S only
        }
        finish_element(du, e, Fa, Ga, Fva, Gva, Uc, this->F, Fv);
        if (has_src) src_to_du(du, e, Sacc);
    }

    // the element's temporal error of kernel 4 (dg_2d_parabolic.jl:1841-1855)
    void temporal_error(const T* u, double dt, int e, const T (*Fa)[MAXN][NV], const T (*Ga)[MAXN][NV],
                        const T (*Fva)[MAXN][NV], const T (*Gva)[MAXN][NV], const T (*d4x)[MAXN][NV],
                        const T (*d4y)[MAXN][NV]) {
        T Fx[MAXN][MAXN][NV], Fy[MAXN][MAXN][NV], hi[MAXN][MAXN][NV], d[MAXN][MAXN][NV];
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) {
                    Fx[j][i][v] = Fa[j][i][v] - Fva[j][i][v];
                    Fy[j][i][v] = Ga[j][i][v] - Gva[j][i][v];
                    hi[j][i][v] = d[j][i][v] = T(0);
                }
        add_divergence(Fx, Fy, -dt, hi, e);
        add_divergence(d4x, d4y, -dt, d, e);
        T err = T(0);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) {
                    const double ij = node_ij(i, j, e);
                    const T h = u[idx(v, i, j, e)] + ij * hi[j][i][v];
                    const T dd = ij * d[j][i][v];
                    const T lo = h - dd;
                    const T r = dd / (this->abstol + this->reltol * max_(abs_(h), abs_(lo)));
                    err = err + r * r;
                }
        this->temporal_errors[e] = err;
    }

    // du += D̂ (F̃a - F̃v), and the element caches U, F (advective) and Fv
    void finish_element(T* du, int e, const T (*Fa)[MAXN][NV], const T (*Ga)[MAXN][NV], const T (*Fva)[MAXN][NV],
                        const T (*Gva)[MAXN][NV], const T (*Uc)[MAXN][NV], std::vector<T>& Fout,
                        std::vector<T>& Fvout) {
        const auto& Dh = basis.Dhat;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                T Fn[NV], Gn[NV];
                for (int v = 0; v < NV; ++v) {
                    Fn[v] = Fa[j][i][v] - Fva[j][i][v];
                    Gn[v] = Ga[j][i][v] - Gva[j][i][v];
                }
                if (cJa) {  // D̂ on the contravariant F̃ (cv_F, cv_G of the reference)
                    const double* J = &cJa[4 * cnode(i, j, e)];
                    for (int v = 0; v < NV; ++v) {
                        const T a = Fn[v], b = Gn[v];
                        Fn[v] = J[0] * a + J[1] * b;
                        Gn[v] = J[2] * a + J[3] * b;
                    }
                }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) du[idx(v, ii, j, e)] = du[idx(v, ii, j, e)] + Dh[ii][i] * Fn[v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) du[idx(v, i, jj, e)] = du[idx(v, i, jj, e)] + Dh[jj][j] * Gn[v];
                for (int v = 0; v < NV; ++v) {
                    this->U[idx(v, i, j, e)] = Uc[j][i][v];
                    Fout[fidx(v, 0, i, j, e)] = Fa[j][i][v];
                    Fout[fidx(v, 1, i, j, e)] = Ga[j][i][v];
                    Fvout[fidx(v, 0, i, j, e)] = Fva[j][i][v];
                    Fvout[fidx(v, 1, i, j, e)] = Gva[j][i][v];
                }
            }
    }

    // f_t of the Taylor states w ± wt, w ± 2wt with their gradients (MDRK kernels)
    void mdrk_ft(const T* w, const T* wt, const T* wx, const T* wtx, const T* wy, const T* wty, T* fta, T* gta,
                 T* ftv, T* gtv) const {
        const double c[4] = {1.0, -1.0, 2.0, -2.0};
        T fa[4][NV], ga[4][NV], fv[4][NV], gv[4][NV];
        for (int q = 0; q < 4; ++q) {
            T s[NV], sx[NV], sy[NV];
            for (int v = 0; v < NV; ++v) {
                s[v] = w[v] + c[q] * wt[v];
                sx[v] = wx[v] + c[q] * wtx[v];
                sy[v] = wy[v] + c[q] * wty[v];
            }
            eq.fluxes(s, fa[q], ga[q]);
            visc.template fluxes<NV>(s, sx, sy, fv[q], gv[q]);
        }
        for (int v = 0; v < NV; ++v) {
            fta[v] = 1.0 / 12.0 * (-fa[2][v] + 8.0 * fa[0][v] - 8.0 * fa[1][v] + fa[3][v]);
            gta[v] = 1.0 / 12.0 * (-ga[2][v] + 8.0 * ga[0][v] - 8.0 * ga[1][v] + ga[3][v]);
            ftv[v] = 1.0 / 12.0 * (-fv[2][v] + 8.0 * fv[0][v] - 8.0 * fv[1][v] + fv[3][v]);
            gtv[v] = 1.0 / 12.0 * (-gv[2][v] + 8.0 * gv[0][v] - 8.0 * gv[1][v] + gv[3][v]);
        }
    }

    // the four stencil states w ± wt, w ± 2wt of a node (up, um, upp, umm)
    void stencil_states(const T* w, const T* wt, States& St, int i, int j) const {
        const double c[4] = {1.0, -1.0, 2.0, -2.0};
        for (int q = 0; q < 4; ++q)
            for (int v = 0; v < NV; ++v) St.s[q][j][i][v] = w[v] + c[q] * wt[v];
    }

    // mdrk_kernel_1! (multi_derivative_2d_parabolic.jl:221-464): F̃ = f/2 + f_t/8,
    // Ũ = u/2 + ut/8 for the advective and viscous parts; F2 = f + f_t/6,
    // U2 = u + ut/6.  The u_low update feeds only the admissibility machinery
    // (and names a field MDRKElementCache lacks) and is omitted.  Source terms
    // (290-300, 380-392, 441-446): ut += dt s(uⁿ, t), S = s/2 + s_t/8 into du,
    // S2 = s + s_t/6 to the stage cache.
    void mdrk_parabolic_1(T* du, const T* u, double dt, int e) {
        T Fa[MAXN][MAXN][NV], Ga[MAXN][MAXN][NV], Fva[MAXN][MAXN][NV], Gva[MAXN][MAXN][NV];
        T Fa2[MAXN][MAXN][NV], Ga2[MAXN][MAXN][NV], Fva2[MAXN][MAXN][NV], Gva2[MAXN][MAXN][NV];
        T Uc[MAXN][MAXN][NV], U2c[MAXN][MAXN][NV], ut[MAXN][MAXN][NV], utx[MAXN][MAXN][NV], uty[MAXN][MAXN][NV];
        T df[MAXN][MAXN][NV], dg[MAXN][MAXN][NV];
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) ut[j][i][v] = T(0);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const T* un = &u[idx(0, i, j, e)];
                T fa[NV], ga[NV], fv[NV], gv[NV];
                eq.fluxes(un, fa, ga);
                visc.template fluxes<NV>(un, &gx[idx(0, i, j, e)], &gy[idx(0, i, j, e)], fv, gv);
                for (int v = 0; v < NV; ++v) {
                    df[j][i][v] = fa[v] - fv[v];
                    dg[j][i][v] = ga[v] - gv[v];
                    Fa[j][i][v] = 0.5 * fa[v];
                    Ga[j][i][v] = 0.5 * ga[v];
                    Fva[j][i][v] = 0.5 * fv[v];
                    Gva[j][i][v] = 0.5 * gv[v];
                    Fa2[j][i][v] = fa[v];
                    Ga2[j][i][v] = ga[v];
                    Fva2[j][i][v] = fv[v];
                    Gva2[j][i][v] = gv[v];
                    Uc[j][i][v] = 0.5 * un[v];
                    U2c[j][i][v] = un[v];
                }
            }
        add_divergence(df, dg, -dt, ut, e);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) ut[j][i][v] = ut[j][i][v] * node_ij(i, j, e);
        const bool has_src = this->src.kind != SRC_NONE;
        static thread_local States St;
        PNodeArr S1, S2c;
        if (has_src) src_level(e, dt, 0, u, St, ut, S2c, 1.0, true);  // S2 = s, ut += dt s
        local_grad(ut, utx, uty, e);
        if (has_src) {
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    stencil_states(&u[idx(0, i, j, e)], ut[j][i], St, i, j);
                    for (int v = 0; v < NV; ++v) S1[j][i][v] = 0.5 * S2c[j][i][v];
                }
            src_level(e, dt, 1, u, St, nullptr, S1, 0.125, false);
            src_level(e, dt, 1, u, St, nullptr, S2c, 1.0 / 6.0, false);
        }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                T fta[NV], gta[NV], ftv[NV], gtv[NV];
                mdrk_ft(&u[idx(0, i, j, e)], ut[j][i], &gx[idx(0, i, j, e)], utx[j][i], &gy[idx(0, i, j, e)],
                        uty[j][i], fta, gta, ftv, gtv);
                for (int v = 0; v < NV; ++v) {
                    Fa[j][i][v] = Fa[j][i][v] + 0.125 * fta[v];
                    Fa2[j][i][v] = Fa2[j][i][v] + (1.0 / 6.0) * fta[v];
                    Ga[j][i][v] = Ga[j][i][v] + 0.125 * gta[v];
                    Ga2[j][i][v] = Ga2[j][i][v] + (1.0 / 6.0) * gta[v];
                    Fva[j][i][v] = Fva[j][i][v] + 0.125 * ftv[v];
                    Fva2[j][i][v] = Fva2[j][i][v] + (1.0 / 6.0) * ftv[v];
                    Gva[j][i][v] = Gva[j][i][v] + 0.125 * gtv[v];
                    Gva2[j][i][v] = Gva2[j][i][v] + (1.0 / 6.0) * gtv[v];
                    Uc[j][i][v] = Uc[j][i][v] + 0.125 * ut[j][i][v];
                    U2c[j][i][v] = U2c[j][i][v] + (1.0 / 6.0) * ut[j][i][v];
                }
            }
        finish_element(du, e, Fa, Ga, Fva, Gva, Uc, this->F, Fv);
        if (has_src) src_to_du(du, e, S1);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) {
                    if (has_src) this->S2[idx(v, i, j, e)] = S2c[j][i][v];
                    this->U2[idx(v, i, j, e)] = U2c[j][i][v];
                    this->F2[fidx(v, 0, i, j, e)] = Fa2[j][i][v];
                    this->F2[fidx(v, 1, i, j, e)] = Ga2[j][i][v];
                    Fv2[fidx(v, 0, i, j, e)] = Fva2[j][i][v];
                    Fv2[fidx(v, 1, i, j, e)] = Gva2[j][i][v];
                }
    }

    // mdrk_kernel_2! (multi_derivative_2d_parabolic.jl:506-711): with the
    // gradient of us, F̃ = F2 + f_t(us)/3 and Ũ = U2 + ust/3 (both parts).
    // Source terms (576-601, 668-674): ust += dt s(us, t) and S = S2 + s_t/3 with
    // s_t of the us states, both at the call's t (where the Euler stage 2 takes
    // t + dt/2).
    void mdrk_parabolic_2(T* du, double dt, int e) {
        const T* us = this->us.data();
        T Fa[MAXN][MAXN][NV], Ga[MAXN][MAXN][NV], Fva[MAXN][MAXN][NV], Gva[MAXN][MAXN][NV];
        T Uc[MAXN][MAXN][NV], ust[MAXN][MAXN][NV], ustx[MAXN][MAXN][NV], usty[MAXN][MAXN][NV];
        T df[MAXN][MAXN][NV], dg[MAXN][MAXN][NV];
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) {
                    Fa[j][i][v] = this->F2[fidx(v, 0, i, j, e)];
                    Ga[j][i][v] = this->F2[fidx(v, 1, i, j, e)];
                    Fva[j][i][v] = Fv2[fidx(v, 0, i, j, e)];
                    Gva[j][i][v] = Fv2[fidx(v, 1, i, j, e)];
                    Uc[j][i][v] = this->U2[idx(v, i, j, e)];
                    ust[j][i][v] = T(0);
                }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const T* un = &us[idx(0, i, j, e)];
                T fa[NV], ga[NV], fv[NV], gv[NV];
                eq.fluxes(un, fa, ga);
                visc.template fluxes<NV>(un, &gx[idx(0, i, j, e)], &gy[idx(0, i, j, e)], fv, gv);
                for (int v = 0; v < NV; ++v) {
                    df[j][i][v] = fa[v] - fv[v];
                    dg[j][i][v] = ga[v] - gv[v];
                }
            }
        add_divergence(df, dg, -dt, ust, e);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) ust[j][i][v] = ust[j][i][v] * node_ij(i, j, e);
        const bool has_src = this->src.kind != SRC_NONE;
        static thread_local States St;
        PNodeArr Sn, Sdummy;
        if (has_src) src_level(e, dt, 0, us, St, ust, Sdummy, 1.0, true);
        local_grad(ust, ustx, usty, e);
        if (has_src) {
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    stencil_states(&us[idx(0, i, j, e)], ust[j][i], St, i, j);
                    for (int v = 0; v < NV; ++v) Sn[j][i][v] = this->S2[idx(v, i, j, e)];
                }
            src_level(e, dt, 1, us, St, nullptr, Sn, 1.0 / 3.0, false);
        }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                T fta[NV], gta[NV], ftv[NV], gtv[NV];
                mdrk_ft(&us[idx(0, i, j, e)], ust[j][i], &gx[idx(0, i, j, e)], ustx[j][i], &gy[idx(0, i, j, e)],
                        usty[j][i], fta, gta, ftv, gtv);
                for (int v = 0; v < NV; ++v) {
                    Fa[j][i][v] = Fa[j][i][v] + (1.0 / 3.0) * fta[v];
                    Ga[j][i][v] = Ga[j][i][v] + (1.0 / 3.0) * gta[v];
                    Fva[j][i][v] = Fva[j][i][v] + (1.0 / 3.0) * ftv[v];
                    Gva[j][i][v] = Gva[j][i][v] + (1.0 / 3.0) * gtv[v];
                    Uc[j][i][v] = Uc[j][i][v] + (1.0 / 3.0) * ust[j][i][v];
                }
            }
        finish_element(du, e, Fa, Ga, Fva, Gva, Uc, this->F, Fv);
        if (has_src) src_to_du(du, e, Sn);
    }

    // prolong2interfaces_lw_parabolic! + calc_interface_flux_hyperbolic_parabolic!
    // (dg_2d_parabolic.jl:1862-1948): F* = LW Rusanov(F̃a, Ũ, uⁿ) - ½(F̃v_L + F̃v_R)
    void calc_interface_flux_hp(const T* u) {
        const int nI = (int)mesh.if_left.size();
#pragma omp parallel for schedule(static)
        for (int I = 0; I < nI; ++I) {
            const int L = mesh.if_left[I], R = mesh.if_right[I], o = mesh.if_orient[I];
            const int ldir = 2 * o - 1, rdir = 2 * o - 2;
            for (int l = 0; l < n; ++l) {
                int li, lj, ri, rj;
                if (o == 1) { li = n - 1; lj = l; ri = 0; rj = l; }
                else { li = l; lj = n - 1; ri = l; rj = 0; }
                T Fh[NV];
                this->lw_numflux(&this->F[fidx(0, o - 1, li, lj, L)], &this->F[fidx(0, o - 1, ri, rj, R)],
                                 &this->U[idx(0, li, lj, L)], &this->U[idx(0, ri, rj, R)], &u[idx(0, li, lj, L)],
                                 &u[idx(0, ri, rj, R)], o, Fh);
                for (int v = 0; v < NV; ++v) {
                    const T fp = 0.5 * (Fv[fidx(v, o - 1, li, lj, L)] + Fv[fidx(v, o - 1, ri, rj, R)]);
                    this->sfv[sidx(v, l, ldir, L)] = Fh[v] - fp;
                    this->sfv[sidx(v, l, rdir, R)] = Fh[v] - fp;
                }
            }
        }
    }

    // the wall condition's u_outer of an inner state (Euler): prim2cons(ρ, v_wall, p) with
    // the inner pressure (Adiabatic, compressible_navier_stokes_2d.jl:113-124) or
    // p = ρ T_wall (Isothermal, 93-101)
    void wall_state(const T* ui, const double* vw, T* uo, int iso = 0, double Tw = 0.0) const {
        if constexpr (NV == 4) {
            const double g = this->eq.gamma;
            const T rho = ui[0];
            const T p = iso ? rho * Tw : (g - 1.0) * (ui[3] - 0.5 * (ui[1] * ui[1] + ui[2] * ui[2]) / rho);
            uo[0] = rho;
            uo[1] = rho * vw[0];
            uo[2] = rho * vw[1];
            uo[3] = p / (g - 1.0) + 0.5 * rho * (vw[0] * vw[0] + vw[1] * vw[1]);
        } else {
            for (int v = 0; v < NV; ++v) uo[v] = ui[v];
        }
    }

    // prolong2boundaries_visc_lw! + calc_boundary_flux_divergence_lw!
    // (dg_2d_parabolic.jl:1951-2085) with the NoSlip/Adiabatic wall of
    // compressible_navier_stokes_2d.jl:81-91: the boundary flux (the hyperbolic wall
    // flux already in place) minus (F̃v1, F̃v2, F̃v3, v_wall · (F̃v2, F̃v3) + 0) of the
    // inner element's time-averaged viscous flux in the face's orientation
    void wall_divergence_flux() {
        const int nB = (int)mesh.bd_elem.size();
#pragma omp parallel for schedule(static)
        for (int B = 0; B < nB; ++B) {
            const int e = mesh.bd_elem[B], o = mesh.bd_orient[B], side = mesh.bd_side[B], dir = mesh.bd_dir[B];
            const double* vw = wall_v[dir - 1];
            for (int l = 0; l < n; ++l) {
                int ii, jj;
                if (o == 1) { ii = (side == 1) ? n - 1 : 0; jj = l; }
                else { ii = l; jj = (side == 1) ? n - 1 : 0; }
                T fv[NV];
                for (int v = 0; v < NV; ++v) fv[v] = Fv[fidx(v, o - 1, ii, jj, e)];
                // Adiabatic: the energy flux v_wall · (τ1n, τ2n) + 0; Isothermal: flux_inner
                if constexpr (NV == 4)
                    if (!wall_iso[dir - 1]) fv[3] = vw[0] * fv[1] + vw[1] * fv[2] + 0.0;
                for (int v = 0; v < NV; ++v) this->sfv[sidx(v, l, dir - 1, e)] = this->sfv[sidx(v, l, dir - 1, e)] - fv[v];
            }
        }
    }

    void rhs_lw(T* du, const T* u, double t, double dt) override {
        this->src_t = t;
        this->reset_du(du);
        calc_gradient(u);
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) lw_volume_parabolic(du, u, dt, e);
        calc_interface_flux_hp(u);
        if (ns_walls && !mesh.bd_elem.empty()) {
            // the hyperbolic wall flux from F̃a (element cache F), then the viscous part
            this->calc_boundary_flux(u, t, dt, 1.0);
            wall_divergence_flux();
        }
        this->surface_integral_and_jacobian(du);
    }
    void rhs_mdrk1(T* du, const T* u, double t, double dt) override {
        this->src_t = t;
        this->ensure_mdrk();
        if (Fv2.size() != Fv.size()) Fv2.assign(Fv.size(), T(0));
        this->reset_du(du);
        calc_gradient(u);
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) mdrk_parabolic_1(du, u, dt, e);
        calc_interface_flux_hp(u);
        if (ns_walls && !mesh.bd_elem.empty()) {
            // multi_derivative_2d_parabolic.jl:110-126: the hyperbolic wall flux with
            // the stage's scaling 0.5 (which the reflecting wall does not read), then
            // the wall's viscous flux from the stage's F̃v (f_v/2 + f_v,t/8)
            this->calc_boundary_flux(u, t, dt, 0.5);
            wall_divergence_flux();
        }
        this->surface_integral_and_jacobian(du);
    }
    void rhs_mdrk2(T* du, const T* u, double t, double dt) override {
        this->src_t = t;
        this->ensure_mdrk();
        if (Fv2.size() != Fv.size()) Fv2.assign(Fv.size(), T(0));
        this->reset_du(du);
        calc_gradient(this->us.data());
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) mdrk_parabolic_2(du, dt, e);
        calc_interface_flux_hp(u);
        if (ns_walls && !mesh.bd_elem.empty()) {
            // multi_derivative_2d_parabolic.jl:192-207: scaling 1, the stage's F̃v
            // (f_v + f_v,t/6 + f_v,t(us)/3) at the wall
            this->calc_boundary_flux(u, t, dt, 1.0);
            wall_divergence_flux();
        }
        this->surface_integral_and_jacobian(du);
    }

    // compute_dt for SemidiscretizationHyperbolicParabolic (lwfr.jl:192-210):
    // min of the advective Trixi max_dt and the parabolic max_dt
    //   1 / ((λa + λv) · 2(2N+1)), λa = max_e J⁻¹(max λ1 + max λ2),
    //   λv = 2(2N+1) max_{e,nodes} lam_v
    // (the CFL number multiplies both, so it multiplies their min).
    double max_dt_viscous(const T* u) const {
        const double tiny = std::nextafter(0.0, 1.0);
        double lam_a = tiny, lam_v = tiny;
        for (int e = 0; e < mesh.E; ++e) {
            double m1 = 0, m2 = 0;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    T l1, l2;
                    eq.max_abs_speeds(&u[idx(0, i, j, e)], l1, l2);
                    m1 = max_(m1, value(l1));
                    m2 = max_(m2, value(l2));
                }
            lam_a = max_(lam_a, mesh.inv_jac[e] * (m1 + m2));
        }
        for (int e = 0; e < mesh.E; ++e)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) lam_v = max_(lam_v, visc.lam_v(&u[idx(0, i, j, e)], mesh.inv_jac[e]));
        const int N = basis.N;
        lam_v *= 2.0 * (2.0 * N + 1.0);
        return 1.0 / ((lam_a + lam_v) * (2 * (2.0 * N + 1.0)));
    }
    double max_dt(const T* u) const override { return std::min(B::max_dt(u), max_dt_viscous(u)); }
};

}  // namespace tlwo
