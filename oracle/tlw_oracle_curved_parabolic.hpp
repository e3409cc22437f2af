// tlw_oracle_curved_parabolic.hpp -- CPU restatement of TrixiLW's LW and MDRK
// right-hand sides of the compressible Navier-Stokes equations (conservative
// gradient variables, constant viscosity) on a periodic curvilinear 2-D mesh.
//
// TEST INFRASTRUCTURE ONLY (parity oracle of the curvilinear hyperbolic-parabolic
// path); the product never links or calls it.
//
// Reference (paths relative to the TrixiLW.jl tree), the P4est kernels of
// src/solvers/dgsem_p4est/dg_2d_parabolic.jl:
//   contravariant_fluxes              4-38 (Ja·(fa, ga) - Ja·(fv, gv))
//   scale_ut!                         40-49 (×J⁻¹ of the node)
//   compute_local_grad_u!             51-104 (D in ξ and η, transposed Ja, ×J⁻¹)
//   lw_volume_kernel_3! / 4!          641-1189, 1190-1787
//   prolong2interfaces_lw_parabolic!  1788-1876
//   calc_interface_flux_hyperbolic_parabolic! 1878-1960 (LW Rusanov with the
//                                     normal, minus ½(F̃v_L·n + F̃v_R·n))
// and, from the un-vendored Trixi.jl (restated from its published definition),
// calc_gradient! for P4estMesh (BR1): the weak-form volume term D̂ u in reference
// coordinates turned with the transposed contravariant vectors, the interface
// value ½(u_L + u_R) times boundary_interpolation and the element's own outward
// normal (±Ja of its face node; the secondary side's value is not negated),
// then ×J⁻¹ of the node; calc_viscous_fluxes! as on the TreeMesh.
//
// Decisions by specification (the reference has no StructuredMesh parabolic
// path, no curvilinear MDRK parabolic kernel, and its P4est max_dt reads
// inverse_jacobian[element] of a per-node array):
//   * interfaces follow the curvilinear hyperbolic path: one normal per face
//     node, the right element's contravariant vector, the LW Rusanov part in the
//     Structured order (λ from uⁿ, jump on Ũ);
//   * the flux time derivatives are formed for the Cartesian fluxes and turned
//     into contravariant form once per level (the same linear maps as the
//     reference's contravariant fluxes of each Taylor state);
//   * the MDRK stages are the TreeMesh parabolic MDRK kernels with these
//     curvilinear operators;
//   * dt = min(the curvilinear advective max_dt, 1/((λa + λv)·2(2N+1))) with λa
//     the largest element sum of the curvilinear scaled speeds and
//     λv = 2(2N+1) max over nodes of J⁻¹ max(4μ/(3ρ), μκ/(ρPr)); on an affine
//     mesh this is the TreeMesh formula (J⁻¹ of the node = (2/h)²).
//
// Pins (tests/test_oracle_curved_parabolic_pins.py): an affine mapping
// reproduces the TreeMesh Navier-Stokes solver (gradients, LW and both MDRK
// stages, dt); the gradient of a linear field is exact on the warped mesh away
// from the periodic seam; a constant state stays at rest; conservation.
#pragma once
#include "tlw_oracle_curved.hpp"
#include "tlw_oracle_parabolic.hpp"

namespace tlwo {

template <class T, class Eq, class Visc>
struct CurvedParabolicSolver : CurvedSolver<T, Eq, ParabolicSolver<T, Eq, Visc>> {
    using P = ParabolicSolver<T, Eq, Visc>;
    using C = CurvedSolver<T, Eq, P>;
    static constexpr int NV = Eq::NV;
    using C::basis;
    using C::eq;
    using C::fidx;
    using C::idx;
    using C::ijn;
    using C::Ja;
    using C::mesh;
    using C::n;
    using C::nidx;
    using C::sidx;

    void setup_curved_parabolic(int kind, const double* p) {
        this->setup_curved(kind, p);
        this->setup_parabolic();
        this->cJa = Ja.data();
        this->cijn = ijn.data();
    }

    // calc_gradient! for P4estMesh (BR1) on a periodic curvilinear mesh
    void calc_gradient_curved(const T* u) {
        auto& gx = this->gx;
        auto& gy = this->gy;
        auto& gsfv = this->gsfv;
        const auto& Dh = basis.Dhat;
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) {
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) gx[idx(v, i, j, e)] = gy[idx(v, i, j, e)] = T(0);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const T* un = &u[idx(0, i, j, e)];
                    for (int ii = 0; ii < n; ++ii)
                        for (int v = 0; v < NV; ++v) gx[idx(v, ii, j, e)] = gx[idx(v, ii, j, e)] + Dh[ii][i] * un[v];
                    for (int jj = 0; jj < n; ++jj)
                        for (int v = 0; v < NV; ++v) gy[idx(v, i, jj, e)] = gy[idx(v, i, jj, e)] + Dh[jj][j] * un[v];
                }
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const double* J = &Ja[4 * nidx(i, j, e)];
                    for (int v = 0; v < NV; ++v) {
                        const T g1 = gx[idx(v, i, j, e)], g2 = gy[idx(v, i, j, e)];
                        gx[idx(v, i, j, e)] = J[0] * g1 + J[2] * g2;
                        gy[idx(v, i, j, e)] = J[1] * g1 + J[3] * g2;
                    }
                }
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
        const double f1 = basis.bnd_lo, f2 = basis.bnd_hi;
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) {
            for (int l = 0; l < n; ++l) {
                // outward normals: -Ja¹ at (0, l), +Ja¹ at (n-1, l), -Ja² at (l, 0), +Ja² at (l, n-1)
                const double* Jw = &Ja[4 * nidx(0, l, e)];
                const double* Je = &Ja[4 * nidx(n - 1, l, e)];
                const double* Js = &Ja[4 * nidx(l, 0, e)];
                const double* Jn = &Ja[4 * nidx(l, n - 1, e)];
                for (int v = 0; v < NV; ++v) {
                    gx[idx(v, 0, l, e)] = gx[idx(v, 0, l, e)] + gsfv[sidx(v, l, 0, e)] * f1 * (-Jw[0]);
                    gx[idx(v, n - 1, l, e)] = gx[idx(v, n - 1, l, e)] + gsfv[sidx(v, l, 1, e)] * f2 * Je[0];
                    gx[idx(v, l, 0, e)] = gx[idx(v, l, 0, e)] + gsfv[sidx(v, l, 2, e)] * f1 * (-Js[2]);
                    gx[idx(v, l, n - 1, e)] = gx[idx(v, l, n - 1, e)] + gsfv[sidx(v, l, 3, e)] * f2 * Jn[2];
                    gy[idx(v, 0, l, e)] = gy[idx(v, 0, l, e)] + gsfv[sidx(v, l, 0, e)] * f1 * (-Jw[1]);
                    gy[idx(v, n - 1, l, e)] = gy[idx(v, n - 1, l, e)] + gsfv[sidx(v, l, 1, e)] * f2 * Je[1];
                    gy[idx(v, l, 0, e)] = gy[idx(v, l, 0, e)] + gsfv[sidx(v, l, 2, e)] * f1 * (-Js[3]);
                    gy[idx(v, l, n - 1, e)] = gy[idx(v, l, n - 1, e)] + gsfv[sidx(v, l, 3, e)] * f2 * Jn[3];
                }
            }
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const double factor = ijn[nidx(i, j, e)];
                    for (int v = 0; v < NV; ++v) {
                        gx[idx(v, i, j, e)] = gx[idx(v, i, j, e)] * factor;
                        gy[idx(v, i, j, e)] = gy[idx(v, i, j, e)] * factor;
                    }
                }
        }
    }

    // calc_interface_flux_hyperbolic_parabolic! with the interface normal of the
    // curvilinear hyperbolic path: F* = ½(F̃a_L + F̃a_R)·n - ½ λ(uⁿ, n)(Ũ_R - Ũ_L)
    // - ½(F̃v_L + F̃v_R)·n, stored for both elements (periodic mesh)
    void calc_interface_flux_hp_curved(const T* u) {
        const auto& Fv = this->Fv;
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e)
            for (int o = 1; o <= 2; ++o) {
                const int left = mesh.nbr[4 * e + (o == 1 ? 0 : 2)];
                if (left < 0) continue;
                for (int l = 0; l < n; ++l) {
                    int li, lj, ri, rj;
                    if (o == 1) { li = n - 1; lj = l; ri = 0; rj = l; }
                    else { li = l; lj = n - 1; ri = l; rj = 0; }
                    double nrm[2], sgn;
                    this->face_normal(e, o, l, nrm, sgn);
                    T fll[NV], frr[NV], vll[NV], vrr[NV];
                    for (int v = 0; v < NV; ++v) {
                        fll[v] = this->F[fidx(v, 0, li, lj, left)] * nrm[0] + this->F[fidx(v, 1, li, lj, left)] * nrm[1];
                        frr[v] = this->F[fidx(v, 0, ri, rj, e)] * nrm[0] + this->F[fidx(v, 1, ri, rj, e)] * nrm[1];
                        vll[v] = Fv[fidx(v, 0, li, lj, left)] * nrm[0] + Fv[fidx(v, 1, li, lj, left)] * nrm[1];
                        vrr[v] = Fv[fidx(v, 0, ri, rj, e)] * nrm[0] + Fv[fidx(v, 1, ri, rj, e)] * nrm[1];
                    }
                    const T* ull = &u[idx(0, li, lj, left)];
                    const T* urr = &u[idx(0, ri, rj, e)];
                    const T lam = eq.max_abs_speed_naive_normal(ull, urr, nrm);
                    for (int v = 0; v < NV; ++v) {
                        const T Fa = 0.5 * (fll[v] + frr[v]) -
                                     0.5 * lam * (this->U[idx(v, ri, rj, e)] - this->U[idx(v, li, lj, left)]);
                        const T Fn = sgn * (Fa - 0.5 * (vll[v] + vrr[v]));
                        this->sfv[sidx(v, l, 2 * o - 1, left)] = Fn;
                        this->sfv[sidx(v, l, 2 * o - 2, e)] = Fn;
                    }
                }
            }
    }

    void rhs_lw(T* du, const T* u, double t, double dt) override {
        (void)t;
        this->reset_du(du);
        calc_gradient_curved(u);
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) this->lw_volume_parabolic(du, u, dt, e);
        calc_interface_flux_hp_curved(u);
        this->surface_integral_and_jacobian_curved(du);
    }
    void rhs_mdrk1(T* du, const T* u, double t, double dt) override {
        (void)t;
        this->ensure_mdrk();
        if (this->Fv2.size() != this->Fv.size()) this->Fv2.assign(this->Fv.size(), T(0));
        this->reset_du(du);
        calc_gradient_curved(u);
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) this->mdrk_parabolic_1(du, u, dt, e);
        calc_interface_flux_hp_curved(u);
        this->surface_integral_and_jacobian_curved(du);
    }
    void rhs_mdrk2(T* du, const T* u, double t, double dt) override {
        (void)t;
        this->ensure_mdrk();
        if (this->Fv2.size() != this->Fv.size()) this->Fv2.assign(this->Fv.size(), T(0));
        this->reset_du(du);
        calc_gradient_curved(this->us.data());
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) this->mdrk_parabolic_2(du, dt, e);
        calc_interface_flux_hp_curved(u);
        this->surface_integral_and_jacobian_curved(du);
    }

    // min(curvilinear advective max_dt, 1 / ((λa + λv)·2(2N+1)))
    double max_dt(const T* u) const override {
        const double tiny = std::nextafter(0.0, 1.0);
        double lam_a = tiny, lam_v = tiny;
        for (int e = 0; e < mesh.E; ++e) {
            double m1 = 0, m2 = 0;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    T l1, l2;
                    eq.max_abs_speeds(&u[idx(0, i, j, e)], l1, l2);
                    const double* J = &Ja[4 * nidx(i, j, e)];
                    const double ij = std::fabs(ijn[nidx(i, j, e)]);
                    m1 = max_(m1, std::fabs(J[0] * value(l1) + J[1] * value(l2)) * ij);
                    m2 = max_(m2, std::fabs(J[2] * value(l1) + J[3] * value(l2)) * ij);
                    lam_v = max_(lam_v, this->visc.lam_v_scaled(&u[idx(0, i, j, e)], ijn[nidx(i, j, e)]));
                }
            lam_a = max_(lam_a, m1 + m2);
        }
        const int N = basis.N;
        const double dta = 2.0 / (n * lam_a);
        lam_v *= 2.0 * (2.0 * N + 1.0);
        const double dtv = 1.0 / ((lam_a + lam_v) * (2 * (2.0 * N + 1.0)));
        return std::min(dta, dtv);
    }
};

}  // namespace tlwo
