// tlw_oracle_curved.hpp -- CPU restatement of TrixiLW's Lax-Wendroff right-hand
// side on a curvilinear 2-D StructuredMesh (periodic), compressible Euler.
//
// TEST INFRASTRUCTURE ONLY (parity oracle and CPU baseline of the curvilinear
// path); the product never links or calls it.
//
// Reference (paths relative to the TrixiLW.jl tree):
//   rhs!                          src/solvers/dgsem_structured/dg_2d.jl:17-53
//   contravariant_flux            dg_2d.jl:541-556
//   lw_volume_kernel_3! / 4!      dg_2d.jl:1379-1685, 1687-2105
//   calc_interface_flux!          dg_2d.jl:2107-2290 (λ from uⁿ, jump on Ũ, F̃ in the
//                                 right element's contravariant direction)
//   rhs_mdrk1! / rhs_mdrk2!       structured/multi_derivative_2d.jl:1-80
//   mdrk_kernel_1! / 2!           structured/multi_derivative_2d.jl:82-243, 245-445
//   normal_product                src/solvers/fr.jl:146-150
// and, from the un-vendored Trixi.jl (restated from their published
// definitions): StructuredMesh node coordinates (the element's GLL nodes in the
// global reference square, mapped), calc_jacobian_matrix! (derivatives of the
// interpolated coordinates with D), the 2-D cross-product contravariant vectors
// Ja¹ = (y_η, -x_η), Ja² = (-y_ξ, x_ξ), inverse_jacobian = 1/(x_ξ y_η - x_η y_ξ),
// max_abs_speed_naive with a normal direction, the weak-form surface integral of
// StructuredMesh (the TreeMesh one), apply_jacobian! with the node's J⁻¹, and
// max_dt of curvilinear meshes (|Ja^k · (λ1, λ2)| |J⁻¹| per node).
//
// Pins (tests/test_oracle_curved_pins.py): an affine mapping reproduces the
// TreeMesh solver with the Structured argument order; free-stream
// preservation and discrete conservation on a warped mesh; convergence of the
// isentropic vortex on it.
#pragma once
#include <type_traits>

#include "tlw_oracle.hpp"

namespace tlwo {

// Mappings of the global reference square [-1, 1]² (StructuredMesh(cells, mapping)):
//   MAP_AFFINE: the box [p0, p0 + p2] × [p1, p1 + p3];
//   MAP_WARP:   [0, L]² (p0 = L) with x = L (s + A sin 2πs sin 2πt),
//               y = L (t + A sin 2πs sin 2πt), s, t ∈ [0, 1], A = p1; the warp
//               vanishes on the domain boundary, so the mesh is periodic;
//   MAP_PARALLELOGRAM: x = ξ + η, y = η (the mapping of the reference's
//               elixir_euler_source_terms_parallelogram.jl; p unused).
enum MappingKind { MAP_AFFINE = 0, MAP_WARP = 1, MAP_PARALLELOGRAM = 2 };

inline void apply_mapping(int kind, const double* p, double xi, double eta, double& x, double& y) {
    if (kind == MAP_AFFINE) {
        x = p[0] + 0.5 * (xi + 1.0) * p[2];
        y = p[1] + 0.5 * (eta + 1.0) * p[3];
    } else if (kind == MAP_PARALLELOGRAM) {
        x = xi + eta;
        y = eta;
    } else {
        const double pi = 3.141592653589793;
        const double s = 0.5 * (xi + 1.0), t = 0.5 * (eta + 1.0);
        const double w = p[1] * std::sin(2.0 * pi * s) * std::sin(2.0 * pi * t);
        x = p[0] * (s + w);
        y = p[0] * (t + w);
    }
}

// Base: Solver<T, Eq>, or ParabolicSolver for the hyperbolic-parabolic
// system on the curvilinear mesh (tlw_oracle_curved_parabolic.hpp).
template <class T, class Eq, class Base = Solver<T, Eq>> struct CurvedSolver : Base {
    using B = Base;
    static constexpr int NV = Eq::NV;
    using B::basis;
    using B::eq;
    using B::fidx;
    using B::idx;
    using B::mesh;
    using B::n;
    using B::sidx;
    std::vector<double> xn, yn;  // node coordinates [e][j][i]
    std::vector<double> Ja;      // [e][j][i][4]: Ja11, Ja12, Ja21, Ja22
    std::vector<double> ijn;     // inverse Jacobian [e][j][i]
    // P4estMesh{2} of trees_per_dimension cells and a mapping (p4est = true): the
    // cells of the StructuredMesh, each element's local frame turned by rot[e]·90°
    // against its cell (the orientations an unstructured quadrilateral mesh has);
    // Trixi's P4est interface container (primary element and face, secondary
    // element and face, flipped node order) and boundary faces, found from the
    // cell adjacency.  Faces are 0-based -ξ, +ξ, -η, +η of the element's frame.
    bool p4est = false;
    std::vector<int> rot;
    struct P4If { int prim, pface, sec, sface, flip; };
    std::vector<P4If> p4if;
    std::vector<int> p4bd_e, p4bd_f;

    size_t nidx(int i, int j, int e) const { return ((size_t)e * n + j) * n + i; }
    void node_xy(int i, int j, int e, double& x, double& y) const override {
        x = xn[nidx(i, j, e)];
        y = yn[nidx(i, j, e)];
    }
    const std::vector<T>* extra_aux(int which) const override {
        if constexpr (std::is_same<T, double>::value) {
            if (which == 13) return &Ja;
            if (which == 14) return &ijn;
        }
        return B::extra_aux(which);
    }

    // Trixi StructuredMesh cache: node coordinates, Jacobian matrix from the
    // derivative matrix, contravariant vectors, inverse Jacobian.
    // local (ξ, η) of an element turned by r·90° -> its cell's (a, b)
    static void turn(int r, double xi, double eta, double& a, double& b) {
        switch (r & 3) {
            case 0: a = xi; b = eta; break;
            case 1: a = -eta; b = xi; break;
            case 2: a = -xi; b = -eta; break;
            default: a = eta; b = -xi; break;
        }
    }
    void setup_curved(int kind, const double* p, const int* rotations = nullptr) {
        const int nx = mesh.nx, ny = mesh.ny, E = mesh.E;
        xn.assign((size_t)E * n * n, 0.0);
        yn.assign(xn.size(), 0.0);
        Ja.assign(4 * xn.size(), 0.0);
        ijn.assign(xn.size(), 0.0);
        rot.assign(E, 0);
        if (rotations)
            for (int e = 0; e < E; ++e) rot[e] = rotations[e] & 3;
        const double dx = 2.0 / nx, dy = 2.0 / ny;
        for (int e = 0; e < E; ++e) {
            const int cx = e % nx, cy = e / nx;
            const double ox = -1.0 + cx * dx + dx / 2, oy = -1.0 + cy * dy + dy / 2;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double a, b;
                    turn(rot[e], basis.nodes[i], basis.nodes[j], a, b);
                    apply_mapping(kind, p, ox + dx / 2 * a, oy + dy / 2 * b, xn[nidx(i, j, e)], yn[nidx(i, j, e)]);
                }
            const auto& D = basis.D;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double x_xi = 0, y_xi = 0, x_eta = 0, y_eta = 0;
                    for (int k = 0; k < n; ++k) {
                        x_xi += D[i][k] * xn[nidx(k, j, e)];
                        y_xi += D[i][k] * yn[nidx(k, j, e)];
                        x_eta += xn[nidx(i, k, e)] * D[j][k];
                        y_eta += yn[nidx(i, k, e)] * D[j][k];
                    }
                    double* J = &Ja[4 * nidx(i, j, e)];
                    J[0] = y_eta;
                    J[1] = -x_eta;
                    J[2] = -y_xi;
                    J[3] = x_xi;
                    ijn[nidx(i, j, e)] = 1.0 / (x_xi * y_eta - x_eta * y_xi);
                }
        }
    }

    // contravariant_flux (dg_2d.jl:550-556)
    void cv_fluxes(const T* u, const double* J, T* f, T* g, T* fc, T* gc) const {
        eq.fluxes(u, f, g);
        for (int v = 0; v < NV; ++v) {
            fc[v] = J[0] * f[v] + J[1] * g[v];
            gc[v] = J[2] * f[v] + J[3] * g[v];
        }
    }

    // lw_volume_kernel_3!/4! for StructuredMesh (dg_2d.jl:1379-2105): the
    // TreeMesh procedure with contravariant fluxes in the derivative
    // contractions and the node's J⁻¹; the Cartesian time average F̃ is kept
    // for the interfaces.  The temporal-error estimate of adaptive steps is
    // computed as the reference does (without source terms).  Source terms
    // (1440-1448, 1505-1513, 1564-1570, 1646-1666 in kernel 3; 1750-1756,
    // 1812-1819, 1884-1889, 1946-1953, 2039-2058 in kernel 4): each level's time
    // derivative gets dt times the source's time difference of its order at the
    // states of the level before (s; s_t of N34; s_tt five-point, tt_N4, in
    // kernel 4, three-point, N23, in kernel 3; s_ttt of N34 -- in kernel 3 for S
    // only; s_tttt of N4 for S only), S the F̃ weights, du += -S/J⁻¹ per node.
    void lw_volume_curved(T* du, const T* u, double dt, int e, double alpha_dg = 1.0) {
        const auto& D = basis.D;
        const auto& Dh = basis.Dhat;
        const int N = basis.N;
        T f[MAXN][MAXN][NV], g[MAXN][MAXN][NV], fc[MAXN][MAXN][NV], gc[MAXN][MAXN][NV];
        T F[MAXN][MAXN][NV], G[MAXN][MAXN][NV], Fc[MAXN][MAXN][NV], Gc[MAXN][MAXN][NV];
        T Uc[MAXN][MAXN][NV], w[MAXN][MAXN][NV], dfc[MAXN][MAXN][NV], dgc[MAXN][MAXN][NV];
        T st[4][MAXN][MAXN][NV];  // up, um, upp, umm
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const T* un = &u[idx(0, i, j, e)];
                cv_fluxes(un, &Ja[4 * nidx(i, j, e)], f[j][i], g[j][i], fc[j][i], gc[j][i]);
                for (int v = 0; v < NV; ++v) {
                    F[j][i][v] = f[j][i][v];
                    G[j][i][v] = g[j][i][v];
                    Fc[j][i][v] = fc[j][i][v];
                    Gc[j][i][v] = gc[j][i][v];
                    Uc[j][i][v] = un[v];
                    for (int q = 0; q < 4; ++q) st[q][j][i][v] = un[v];
                }
            }
        // w = J⁻¹ (-dt D) (a, b) for the contravariant pair (a, b)
        auto advance = [&](const T (*a)[MAXN][NV], const T (*b)[MAXN][NV]) {
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) w[j][i][v] = T(0);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    for (int ii = 0; ii < n; ++ii)
                        for (int v = 0; v < NV; ++v) w[j][ii][v] = w[j][ii][v] + (-dt * D[ii][i]) * a[j][i][v];
                    for (int jj = 0; jj < n; ++jj)
                        for (int v = 0; v < NV; ++v) w[jj][i][v] = w[jj][i][v] + (-dt * D[jj][j]) * b[j][i][v];
                }
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) w[j][i][v] = w[j][i][v] * ijn[nidx(i, j, e)];
        };
        auto add_states = [&](double cU, double c0, double c1, double c2, double c3) {
            const double c[4] = {c0, c1, c2, c3};
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) {
                        Uc[j][i][v] = Uc[j][i][v] + cU * w[j][i][v];
                        for (int q = 0; q < 4; ++q) st[q][j][i][v] = st[q][j][i][v] + c[q] * w[j][i][v];
                    }
        };
        // stencil weights on (u++, u+, u, u-, u--), scale and the F̃ weight of a stage
        auto stage = [&](const double* wts, double scale, double fw) {
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const double* J = &Ja[4 * nidx(i, j, e)];
                    T sf[5][NV], sg[5][NV], sfc[5][NV], sgc[5][NV];
                    const int q_of[5] = {2, 0, -1, 1, 3};
                    for (int s = 0; s < 5; ++s) {
                        if (wts[s] == 0.0) continue;
                        if (q_of[s] < 0) {
                            for (int v = 0; v < NV; ++v) {
                                sf[s][v] = f[j][i][v];
                                sg[s][v] = g[j][i][v];
                                sfc[s][v] = fc[j][i][v];
                                sgc[s][v] = gc[j][i][v];
                            }
                        } else {
                            cv_fluxes(st[q_of[s]][j][i], J, sf[s], sg[s], sfc[s], sgc[s]);
                        }
                    }
                    for (int v = 0; v < NV; ++v) {
                        T a = T(0), b = T(0), ac = T(0), bc = T(0);
                        for (int s = 0; s < 5; ++s) {
                            if (wts[s] == 0.0) continue;
                            a = a + wts[s] * sf[s][v];
                            b = b + wts[s] * sg[s][v];
                            ac = ac + wts[s] * sfc[s][v];
                            bc = bc + wts[s] * sgc[s][v];
                        }
                        a = scale * a;
                        b = scale * b;
                        ac = scale * ac;
                        bc = scale * bc;
                        F[j][i][v] = F[j][i][v] + fw * a;
                        G[j][i][v] = G[j][i][v] + fw * b;
                        Fc[j][i][v] = Fc[j][i][v] + fw * ac;
                        Gc[j][i][v] = Gc[j][i][v] + fw * bc;
                        dfc[j][i][v] = ac;
                        dgc[j][i][v] = bc;
                    }
                }
        };
        const bool has_src = this->src.kind != SRC_NONE;
        T Sacc[MAXN][MAXN][NV];
        // source level `order` at the current states: w += dt s^(k), S += wS s^(k)
        auto src_level = [&](int order, double wS, bool into_w) {
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double x, y;
                    node_xy(i, j, e, x, y);
                    const T* un = &u[idx(0, i, j, e)];
                    const T *up = st[0][j][i], *um = st[1][j][i], *upp = st[2][j][i], *umm = st[3][j][i];
                    T sv[NV];
                    if (order == 0) {
                        this->src_at(un, x, y, 0.0, sv);
                    } else if (order == 1) {
                        this->src_t_N34(up, upp, um, umm, x, y, dt, this->src_t, sv);
                    } else if (order == 2 && N == 4) {
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
                        Sacc[j][i][v] = order == 0 ? wS * sv[v] : Sacc[j][i][v] + wS * sv[v];
                        if (into_w) w[j][i][v] = w[j][i][v] + dt * sv[v];
                    }
                }
        };
        advance(fc, gc);  // ut
        if (has_src) src_level(0, 1.0, true);
        add_states(0.5, 1.0, -1.0, 2.0, -2.0);
        {
            const double wt[5] = {-1.0, 8.0, 0.0, -8.0, 1.0};  // f_t
            stage(wt, 1.0 / 12.0, 0.5);
        }
        advance(dfc, dgc);  // utt
        if (has_src) src_level(1, 0.5, true);
        add_states(1.0 / 6.0, 0.5, 0.5, 2.0, 2.0);
        if (N == 4) {
            const double wt[5] = {-1.0, 16.0, -30.0, 16.0, -1.0};  // f_tt
            stage(wt, 1.0 / 12.0, 1.0 / 6.0);
        } else {
            const double wt[5] = {0.0, 1.0, -2.0, 1.0, 0.0};
            stage(wt, 1.0, 1.0 / 6.0);
        }
        advance(dfc, dgc);  // uttt
        if (has_src) src_level(2, 1.0 / 6.0, true);
        add_states(1.0 / 24.0, 1.0 / 6.0, -1.0 / 6.0, 4.0 / 3.0, -4.0 / 3.0);
        {
            const double wt[5] = {1.0, -2.0, 0.0, 2.0, -1.0};  // f_ttt
            stage(wt, 0.5, 1.0 / 24.0);
        }
        if (N == 4) {
            advance(dfc, dgc);  // utttt
            if (has_src) src_level(3, 1.0 / 24.0, true);
            add_states(1.0 / 120.0, 1.0 / 24.0, 1.0 / 24.0, 2.0 / 3.0, 2.0 / 3.0);
            const double wt[5] = {1.0, -4.0, 6.0, -4.0, 1.0};  // f_tttt, with the 0.5 of dg_2d.jl:2018-2021
            stage(wt, 0.5, 1.0 / 120.0);
            if (has_src) src_level(4, 1.0 / 120.0, false);
        } else if (has_src) {
            src_level(3, 1.0 / 24.0, false);
        }
        // temporal error estimate of the adaptive controller (dg_2d.jl:1595-1619, 1664-1681 at N=3;
        // 1983-2007, 2058-2077 at N=4): u_np1 = u + J⁻¹(node) (-dt D)·F̃ with the contravariant F̃,
        // u_np1_low the same without the last stage's term (f_ttt/24 at N=3, f_tttt/120 at N=4);
        // the numerator is that term's update itself, as in the TreeMesh kernel
        {
            const double wlast = (N == 4) ? 1.0 / 120.0 : 1.0 / 24.0;
            advance(dfc, dgc);
            T dlo[MAXN][MAXN][NV];
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) dlo[j][i][v] = wlast * w[j][i][v];
            advance(Fc, Gc);
            T err = T(0);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) {
                        const T hi = u[idx(v, i, j, e)] + w[j][i][v];
                        const T lo = hi - dlo[j][i][v];
                        const T r = dlo[j][i][v] / (this->abstol + this->reltol * max_(abs_(hi), abs_(lo)));
                        err = err + r * r;
                    }
            this->temporal_errors[e] = err;
        }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, ii, j, e)] = du[idx(v, ii, j, e)] + (alpha_dg * Dh[ii][i]) * Fc[j][i][v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, i, jj, e)] = du[idx(v, i, jj, e)] + (alpha_dg * Dh[jj][j]) * Gc[j][i][v];
                for (int v = 0; v < NV; ++v) {
                    this->U[idx(v, i, j, e)] = Uc[j][i][v];
                    this->F[fidx(v, 0, i, j, e)] = F[j][i][v];
                    this->F[fidx(v, 1, i, j, e)] = G[j][i][v];
                    if (has_src)
                        du[idx(v, i, j, e)] = du[idx(v, i, j, e)] + (-1.0 / ijn[nidx(i, j, e)]) * Sacc[j][i][v];
                }
            }
    }

    // ---------------- shock capturing on curvilinear meshes ----------------
    // Subcell normal of direction d (1: ξ, 2: η) at subcell face k (0-based,
    // 1..n-1) of row r (Trixi calcflux_fv! for curvilinear meshes): the
    // contravariant vector Ja^d of the row's first node plus, for every subcell
    // before the face, w_q Σ_m D[q][m] Ja^d(m), summed in that order.  With
    // `quirk`, the face n-1 normal of the reference's calc_fn_low!
    // (structured/dg_2d.jl:111-173), whose `for i in (2, nnodes(dg))` adds the
    // q = 0 and q = n-2 terms only.
    void subcell_normal(int d, int e, int r, int k, double* nrm, bool quirk = false) const {
        const auto& D = basis.D;
        const auto& w = basis.weights;
        auto ja = [&](int m) {
            const size_t node = d == 1 ? nidx(m, r, e) : nidx(r, m, e);
            return &Ja[4 * node + (d == 1 ? 0 : 2)];
        };
        nrm[0] = ja(0)[0];
        nrm[1] = ja(0)[1];
        for (int q = 0; q < k; ++q) {
            if (quirk && q != 0 && q != n - 2) continue;
            for (int m = 0; m < n; ++m) {
                const double c = w[q] * D[q][m];
                nrm[0] = nrm[0] + c * ja(m)[0];
                nrm[1] = nrm[1] + c * ja(m)[1];
            }
        }
    }
    // flux_lax_friedrichs with a normal direction (Trixi FluxPlusDissipation of
    // flux_central and DissipationLocalLaxFriedrichs)
    void rusanov_normal(const T* ul, const T* ur, const double* nrm, T* out) const {
        T fl[NV], fr[NV];
        eq.flux_normal(ul, nrm, fl);
        eq.flux_normal(ur, nrm, fr);
        const T lam = eq.max_abs_speed_naive_normal(ul, ur, nrm);
        for (int v = 0; v < NV; ++v) out[v] = 0.5 * (fl[v] + fr[v]) + (-0.5 * lam) * (ur[v] - ul[v]);
    }
    // fv_kernel! with FirstOrderReconstruction on a curvilinear mesh (dg_2d.jl:251-282
    // with Trixi calcflux_fv!): Rusanov fluxes with the subcell normals between
    // neighbouring nodes, zero at the element faces, the α-scaled FV term and
    // fn_low (load_fn_low!).  With add = false only fn_low, from the quirky
    // normals of calc_fn_low! (the reference's pure-DG elements).
    void fv_kernel_fo_curved(T* du, const T* u, int e, double fv_scale, bool add) {
        T fx[MAXN][MAXN + 1][NV], gy[MAXN + 1][MAXN][NV];
        for (int a = 0; a < n; ++a)
            for (int v = 0; v < NV; ++v) fx[a][0][v] = fx[a][n][v] = gy[0][a][v] = gy[n][a][v] = T(0);
        for (int r = 0; r < n; ++r)
            for (int k = 1; k < n; ++k) {
                if (!add && k != 1 && k != n - 1) continue;
                double nrm[2];
                subcell_normal(1, e, r, k, nrm, !add);
                rusanov_normal(&u[idx(0, k - 1, r, e)], &u[idx(0, k, r, e)], nrm, fx[r][k]);
                subcell_normal(2, e, r, k, nrm, !add);
                rusanov_normal(&u[idx(0, r, k - 1, e)], &u[idx(0, r, k, e)], nrm, gy[k][r]);
            }
        if (add) {
            const auto& iw = basis.inv_w;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, i, j, e)] = du[idx(v, i, j, e)] + fv_scale * (iw[i] * (fx[j][i + 1][v] - fx[j][i][v]) +
                                                                                 iw[j] * (gy[j + 1][i][v] - gy[j][i][v]));
        }
        for (int k = 0; k < n; ++k)
            for (int v = 0; v < NV; ++v) {
                this->fn_low[this->sidx(v, k, 0, e)] = fx[k][1][v];
                this->fn_low[this->sidx(v, k, 1, e)] = fx[k][n - 1][v];
                this->fn_low[this->sidx(v, k, 2, e)] = gy[1][k][v];
                this->fn_low[this->sidx(v, k, 3, e)] = gy[n - 1][k][v];
            }
    }
    void sc_element_curved(T* du, const T* u, int e) {
        const double a_e = this->alpha[e];
        fv_kernel_fo_curved(du, u, e, a_e, !(std::fabs(a_e) <= 1e-12));
    }
    // compute_alp for StructuredMesh (structured/dg_2d.jl:2169-2192): the nodes'
    // own inverse Jacobians
    double compute_alp_lr(const T* ul, const T* ur, int L, int R, double Jl, double Jr, double dt, const T* fn,
                          const T* Fn_, const T* fin_ll, const T* fin_rr) const {
        const double alp = 0.5 * (this->alpha[L] + this->alpha[R]);
        T Fn[NV], low[NV];
        for (int v = 0; v < NV; ++v) Fn[v] = (1.0 - alp) * Fn_[v] + alp * fn[v];
        const double lx = 0.5;
        const double cl = dt * Jl / (basis.weights[n - 1] * lx);
        for (int v = 0; v < NV; ++v) low[v] = ul[v] - cl * (Fn[v] - fin_ll[v]);
        if (!eq.is_admissible(low)) return 1.0;
        const double cr = dt * Jr / (basis.weights[0] * lx);
        for (int v = 0; v < NV; ++v) low[v] = ur[v] - cr * (fin_rr[v] - Fn[v]);
        if (!eq.is_admissible(low)) return 1.0;
        return alp;
    }

    // Normal of the face node shared by left and right elements: the right
    // element's contravariant vector of the interface orientation (dg_2d.jl:2229-2264)
    void face_normal(int right, int orientation, int l, double* nrm, double& sgn) const {
        const size_t k = orientation == 1 ? nidx(0, l, right) : nidx(l, 0, right);
        sgn = ijn[k] > 0 ? 1.0 : (ijn[k] < 0 ? -1.0 : 0.0);
        const double* J = &Ja[4 * k];
        nrm[0] = sgn * (orientation == 1 ? J[0] : J[2]);
        nrm[1] = sgn * (orientation == 1 ? J[1] : J[3]);
    }

    // calc_interface_flux! for StructuredMesh (dg_2d.jl:2107-2290), plain FR:
    // Fn = sign · [½(F̃_ll·n + F̃_rr·n) - ½ λ(uⁿ, n) (Ũ_rr - Ũ_ll)] (periodic mesh)
    void calc_interface_flux_curved(const T* u) {
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
                    face_normal(e, o, l, nrm, sgn);
                    T fll[NV], frr[NV];
                    for (int v = 0; v < NV; ++v) {
                        fll[v] = this->F[fidx(v, 0, li, lj, left)] * nrm[0] + this->F[fidx(v, 1, li, lj, left)] * nrm[1];
                        frr[v] = this->F[fidx(v, 0, ri, rj, e)] * nrm[0] + this->F[fidx(v, 1, ri, rj, e)] * nrm[1];
                    }
                    const T* ull = &u[idx(0, li, lj, left)];
                    const T* urr = &u[idx(0, ri, rj, e)];
                    const T lam = eq.max_abs_speed_naive_normal(ull, urr, nrm);
                    T Fn[NV];
                    for (int v = 0; v < NV; ++v)
                        Fn[v] = sgn * (0.5 * (fll[v] + frr[v]) -
                                       0.5 * lam * (this->U[idx(v, ri, rj, e)] - this->U[idx(v, li, lj, left)]));
                    if (this->sc) {
                        // compute_low_order_flux and compute_alp (structured/dg_2d.jl:2276-2285)
                        T fn[NV];
                        rusanov_normal(ull, urr, nrm, fn);
                        for (int v = 0; v < NV; ++v) fn[v] = sgn * fn[v];
                        const T* fin_ll = &this->fn_low[sidx(0, l, 2 * o - 1, left)];
                        const T* fin_rr = &this->fn_low[sidx(0, l, 2 * o - 2, e)];
                        const double alp = compute_alp_lr(ull, urr, left, e, ijn[nidx(li, lj, left)],
                                                          ijn[nidx(ri, rj, e)], this->sc_dt, fn, Fn, fin_ll, fin_rr);
                        for (int v = 0; v < NV; ++v) Fn[v] = alp * fn[v] + (1.0 - alp) * Fn[v];
                    }
                    for (int v = 0; v < NV; ++v) {
                        this->sfv[sidx(v, l, 2 * o - 1, left)] = Fn[v];
                        this->sfv[sidx(v, l, 2 * o - 2, e)] = Fn[v];
                    }
                }
            }
    }

    // calc_boundary_flux! for StructuredMesh (structured/dg_2d.jl:2297-2391) with a
    // time-averaged Dirichlet value (equations.jl:6-37): the element's own
    // contravariant vector at the boundary node as normal, F̃·n inside, the GLL
    // time average of the boundary value and of its normal flux outside, and the
    // LW Rusanov flux in the TreeMesh argument order (λ from Ũ, jump on uⁿ).
    // scaling_factor is 1 in both MDRK stages on this mesh (the call passes none).
    void calc_boundary_flux_curved(const T* u, double t, double dt) {
        if constexpr (std::is_same<Eq, Euler2D<T>>::value) {
            const auto& nodes = basis.nodes;
            const auto& w = basis.weights;
            const int nB = (int)mesh.bd_elem.size();
#pragma omp parallel for schedule(static)
            for (int B = 0; B < nB; ++B) {
                const int e = mesh.bd_elem[B], o = mesh.bd_orient[B], side = mesh.bd_side[B], dir = mesh.bd_dir[B];
                for (int l = 0; l < n; ++l) {
                    int ii, jj;
                    if (o == 1) { ii = (side == 1) ? n - 1 : 0; jj = l; }
                    else { ii = l; jj = (side == 1) ? n - 1 : 0; }
                    const size_t k = nidx(ii, jj, e);
                    const double* J = &Ja[4 * k];
                    const double sgn = ijn[k] > 0 ? 1.0 : (ijn[k] < 0 ? -1.0 : 0.0);
                    const double nrm[2] = {sgn * (o == 1 ? J[0] : J[2]), sgn * (o == 1 ? J[1] : J[3])};
                    const T* u_in = &u[idx(0, ii, jj, e)];
                    const T* U_in = &this->U[idx(0, ii, jj, e)];
                    if (this->bval.kind == BV_OUTFLOW || this->bval.kind == BV_SLIP_WALL_APPROX) {
                        // conditions on uⁿ with the face's normal (their P4est call,
                        // p4est/dg_2d.jl:361-393, with the outward normal): outflow is the
                        // normal flux of uⁿ; the approximate slip wall the Rusanov flux between
                        // uⁿ and its reflection (get_reflection, equations.jl:175-183).  In the
                        // +ξ/+η direction of the element's contravariant vector both give the
                        // flux below: the outer state on the side of the boundary.
                        T fb[NV];
                        if (this->bval.kind == BV_OUTFLOW) {
                            eq.flux_normal(u_in, nrm, fb);
                        } else {
                            T rho, v1, v2, p, ur[NV];
                            eq.cons2prim(u_in, rho, v1, v2, p);
                            const double ns = nrm[0] * nrm[0] + nrm[1] * nrm[1];
                            const T vn = (v1 * nrm[0] + v2 * nrm[1]) / ns;
                            const T w1 = v1 - 2.0 * vn * nrm[0], w2 = v2 - 2.0 * vn * nrm[1];
                            ur[0] = rho;
                            ur[1] = rho * w1;
                            ur[2] = rho * w2;
                            ur[3] = p / (eq.gamma - 1.0) + 0.5 * rho * (w1 * w1 + w2 * w2);
                            const T* ul = (dir % 2 == 0) ? u_in : ur;
                            const T* uq = (dir % 2 == 0) ? ur : u_in;
                            T fl[NV], fr[NV];
                            eq.flux_normal(ul, nrm, fl);
                            eq.flux_normal(uq, nrm, fr);
                            const T lam = eq.max_abs_speed_naive_normal(ul, uq, nrm);
                            for (int v = 0; v < NV; ++v) fb[v] = 0.5 * (fl[v] + fr[v]) + (-0.5 * lam) * (uq[v] - ul[v]);
                        }
                        for (int v = 0; v < NV; ++v) this->sfv[sidx(v, l, dir - 1, e)] = sgn * fb[v];
                        continue;
                    }
                    T f_in[NV], Uo[NV], Fo[NV], uo[NV];
                    for (int v = 0; v < NV; ++v) {
                        f_in[v] = this->F[fidx(v, 0, ii, jj, e)] * nrm[0] + this->F[fidx(v, 1, ii, jj, e)] * nrm[1];
                        Uo[v] = T(0);
                        Fo[v] = T(0);
                    }
                    const double x = xn[k], y = yn[k];
                    for (int q = 0; q < n; ++q) {
                        const double ts = t + 0.5 * dt * (nodes[q] + 1.0);
                        double ubd[4];
                        this->bval.eval(x, y, ts, ubd);
                        T ub[NV], fb[NV];
                        for (int v = 0; v < NV; ++v) ub[v] = T(ubd[v]);
                        eq.flux_normal(ub, nrm, fb);
                        for (int v = 0; v < NV; ++v) {
                            Uo[v] = Uo[v] + 0.5 * ub[v] * w[q];
                            Fo[v] = Fo[v] + 0.5 * fb[v] * w[q];
                        }
                    }
                    double uod[4];
                    this->bval.eval(x, y, t, uod);
                    for (int v = 0; v < NV; ++v) uo[v] = T(uod[v]);
                    // surface_flux(F_l, F_r, Ũ_l, Ũ_r, uⁿ_l, uⁿ_r, n): λ from Ũ, jump on uⁿ
                    const T* Ul = (dir % 2 == 0) ? U_in : Uo;
                    const T* Ur = (dir % 2 == 0) ? Uo : U_in;
                    const T* fl = (dir % 2 == 0) ? f_in : Fo;
                    const T* fr = (dir % 2 == 0) ? Fo : f_in;
                    const T* jl = (dir % 2 == 0) ? u_in : uo;
                    const T* jr = (dir % 2 == 0) ? uo : u_in;
                    const T lam = eq.max_abs_speed_naive_normal(Ul, Ur, nrm);
                    for (int v = 0; v < NV; ++v)
                        this->sfv[sidx(v, l, dir - 1, e)] =
                            sgn * (0.5 * (fl[v] + fr[v]) + (-0.5 * lam) * (jr[v] - jl[v]));
                }
            }
        }
    }

    // ---- P4estMesh{2} -------------------------------------------------------
    static void face_node(int n_, int f, int l, int& i, int& j) {
        if (f < 2) { i = (f == 0) ? 0 : n_ - 1; j = l; }
        else { i = l; j = (f == 2) ? 0 : n_ - 1; }
    }
    // Interfaces and boundary faces from the cell adjacency: local face f of
    // element e points along its turned outward vector; the neighbour cell's
    // element meets it with the face whose turned outward vector is opposite,
    // and its node order is flipped when the turned tangents are opposite.
    void setup_p4est() {
        p4est = true;
        const int nx = mesh.nx, ny = mesh.ny, E = mesh.E;
        const int out[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
        const int tan[4][2] = {{0, 1}, {0, 1}, {1, 0}, {1, 0}};
        auto turned = [&](int r, const int* v, int* w) {
            int a = v[0], b = v[1];
            for (int k = 0; k < (r & 3); ++k) { const int t = a; a = -b; b = t; }
            w[0] = a; w[1] = b;
        };
        p4if.clear(); p4bd_e.clear(); p4bd_f.clear();
        for (int e = 0; e < E; ++e)
            for (int f = 0; f < 4; ++f) {
                int d[2], t[2];
                turned(rot[e], out[f], d);
                turned(rot[e], tan[f], t);
                int cx = e % nx + d[0], cy = e / nx + d[1];
                if (cx < 0 || cx >= nx) {
                    if (!mesh.per_x) { p4bd_e.push_back(e); p4bd_f.push_back(f); continue; }
                    cx = (cx + nx) % nx;
                }
                if (cy < 0 || cy >= ny) {
                    if (!mesh.per_y) { p4bd_e.push_back(e); p4bd_f.push_back(f); continue; }
                    cy = (cy + ny) % ny;
                }
                const int e2 = cx + nx * cy;
                int f2 = -1, flip = 0;
                for (int g = 0; g < 4; ++g) {
                    int d2[2], t2[2];
                    turned(rot[e2], out[g], d2);
                    if (d2[0] == -d[0] && d2[1] == -d[1]) {
                        f2 = g;
                        turned(rot[e2], tan[g], t2);
                        flip = (t2[0] != t[0] || t2[1] != t[1]) ? 1 : 0;
                    }
                }
                // each interface once, primary = the lower (element, face)
                if (e < e2 || (e == e2 && f < f2)) p4if.push_back({e, f, e2, f2, flip});
            }
    }
    // outward normal of local face f at its node (i, j) (Trixi get_normal_direction)
    void outward(int e, int f, int i, int j, double* nrm) const {
        const double* J = &Ja[4 * nidx(i, j, e)];
        const double s = (f & 1) ? 1.0 : -1.0;
        nrm[0] = s * (f < 2 ? J[0] : J[2]);
        nrm[1] = s * (f < 2 ? J[1] : J[3]);
    }
    // prolong2interfaces! + calc_interface_flux! for P4estMesh (p4est/dg_2d.jl:8-192),
    // plain FR: each side's F̃ in the primary's outward direction from its own
    // contravariant vector (the secondary's outward normal negated), λ from uⁿ with
    // the primary's normal, jump on Ũ; the primary keeps F*, the secondary -F* at its
    // own (possibly reversed) node.  Stored in the +ξ/+η convention of the surface
    // integral: an outward flux of face f is (+1 | -1 for f odd | even) times it.
    void calc_interface_flux_p4est(const T* u) {
        const int nI = (int)p4if.size();
#pragma omp parallel for schedule(static)
        for (int I = 0; I < nI; ++I) {
            const P4If& q = p4if[I];
            for (int l = 0; l < n; ++l) {
                const int ls = q.flip ? n - 1 - l : l;
                int ip, jp, is, js;
                face_node(n, q.pface, l, ip, jp);
                face_node(n, q.sface, ls, is, js);
                double np_[2], ns_[2];
                outward(q.prim, q.pface, ip, jp, np_);
                outward(q.sec, q.sface, is, js, ns_);
                T fll[NV], frr[NV];
                for (int v = 0; v < NV; ++v) {
                    fll[v] = this->F[fidx(v, 0, ip, jp, q.prim)] * np_[0] + this->F[fidx(v, 1, ip, jp, q.prim)] * np_[1];
                    frr[v] = this->F[fidx(v, 0, is, js, q.sec)] * (-ns_[0]) + this->F[fidx(v, 1, is, js, q.sec)] * (-ns_[1]);
                }
                const T* ull = &u[idx(0, ip, jp, q.prim)];
                const T* urr = &u[idx(0, is, js, q.sec)];
                const T lam = eq.max_abs_speed_naive_normal(ull, urr, np_);
                const double sp = (q.pface & 1) ? 1.0 : -1.0, ss = (q.sface & 1) ? 1.0 : -1.0;
                for (int v = 0; v < NV; ++v) {
                    const T Fn = 0.5 * (fll[v] + frr[v]) -
                                 0.5 * lam * (this->U[idx(v, is, js, q.sec)] - this->U[idx(v, ip, jp, q.prim)]);
                    this->sfv[sidx(v, l, q.pface, q.prim)] = sp * Fn;
                    this->sfv[sidx(v, ls, q.sface, q.sec)] = -ss * Fn;
                }
            }
        }
    }
    // prolong2boundaries! + calc_boundary_flux! for P4estMesh (p4est/dg_2d.jl:283-393)
    // with the conditions' P4est methods: the outward normal, uⁿ, Ũ and F̃·n inside;
    // the time-averaged Dirichlet condition (equations.jl:66-93) takes
    // surface_flux(F̃·n, F_outer, uⁿ, U_outer, Ũ, U_outer, n): λ from uⁿ and the time
    // average U_outer, jump on Ũ; outflow and the approximate slip wall
    // (equations.jl:185-197, 251-260) use uⁿ alone, the inner state first.
    void calc_boundary_flux_p4est(const T* u, double t, double dt, double scaling_factor) {
        if constexpr (std::is_same<Eq, Euler2D<T>>::value) {
            const auto& nodes = basis.nodes;
            const auto& w = basis.weights;
            const int nB = (int)p4bd_e.size();
#pragma omp parallel for schedule(static)
            for (int B = 0; B < nB; ++B) {
                const int e = p4bd_e[B], f = p4bd_f[B];
                const double sf = (f & 1) ? 1.0 : -1.0;
                for (int l = 0; l < n; ++l) {
                    int i, j;
                    face_node(n, f, l, i, j);
                    double nrm[2];
                    outward(e, f, i, j, nrm);
                    const T* u_in = &u[idx(0, i, j, e)];
                    const T* U_in = &this->U[idx(0, i, j, e)];
                    T fb[NV];
                    if (this->bval.kind == BV_OUTFLOW) {
                        eq.flux_normal(u_in, nrm, fb);
                    } else if (this->bval.kind == BV_SLIP_WALL_APPROX) {
                        T rho, v1, v2, p, ur[NV];
                        eq.cons2prim(u_in, rho, v1, v2, p);
                        const double ns = nrm[0] * nrm[0] + nrm[1] * nrm[1];
                        const T vn = (v1 * nrm[0] + v2 * nrm[1]) / ns;
                        const T w1 = v1 - 2.0 * vn * nrm[0], w2 = v2 - 2.0 * vn * nrm[1];
                        ur[0] = rho;
                        ur[1] = rho * w1;
                        ur[2] = rho * w2;
                        ur[3] = p / (eq.gamma - 1.0) + 0.5 * rho * (w1 * w1 + w2 * w2);
                        T fl[NV], fr[NV];
                        eq.flux_normal(u_in, nrm, fl);
                        eq.flux_normal(ur, nrm, fr);
                        const T lam = eq.max_abs_speed_naive_normal(u_in, ur, nrm);
                        for (int v = 0; v < NV; ++v) fb[v] = 0.5 * (fl[v] + fr[v]) + (-0.5 * lam) * (ur[v] - u_in[v]);
                    } else {
                        T f_in[NV], Uo[NV], Fo[NV];
                        for (int v = 0; v < NV; ++v) {
                            f_in[v] = this->F[fidx(v, 0, i, j, e)] * nrm[0] + this->F[fidx(v, 1, i, j, e)] * nrm[1];
                            Uo[v] = T(0);
                            Fo[v] = T(0);
                        }
                        const double x = xn[nidx(i, j, e)], y = yn[nidx(i, j, e)];
                        const double dts = scaling_factor * dt;
                        for (int q = 0; q < n; ++q) {
                            const double ts = t + 0.5 * dts * (nodes[q] + 1.0);
                            double ubd[4];
                            this->bval.eval(x, y, ts, ubd);
                            T ub[NV], fbq[NV];
                            for (int v = 0; v < NV; ++v) ub[v] = T(ubd[v]);
                            eq.flux_normal(ub, nrm, fbq);
                            for (int v = 0; v < NV; ++v) {
                                Uo[v] = Uo[v] + ub[v] * scaling_factor * w[q] / 2.0;
                                Fo[v] = Fo[v] + fbq[v] * scaling_factor * w[q] / 2.0;
                            }
                        }
                        const T lam = eq.max_abs_speed_naive_normal(u_in, Uo, nrm);
                        for (int v = 0; v < NV; ++v) fb[v] = 0.5 * (f_in[v] + Fo[v]) - 0.5 * lam * (Uo[v] - U_in[v]);
                    }
                    for (int v = 0; v < NV; ++v) this->sfv[sidx(v, l, f, e)] = sf * fb[v];
                }
            }
        }
    }
    // the interface and boundary phases of one rhs call (scaling_factor: MDRK stage 1 of
    // the P4est mesh passes 0.5, tree/multi_derivative_2d.jl:34)
    void surface_fluxes(const T* u, double t, double dt, double p4est_scaling) {
        if (p4est) {
            calc_interface_flux_p4est(u);
            if (!p4bd_e.empty()) calc_boundary_flux_p4est(u, t, dt, p4est_scaling);
        } else {
            calc_interface_flux_curved(u);
            if (!mesh.bd_elem.empty()) calc_boundary_flux_curved(u, t, dt);
        }
    }

    // Trixi calc_surface_integral! (StructuredMesh uses the TreeMesh form) and
    // apply_jacobian! with the node's J⁻¹
    void surface_integral_and_jacobian_curved(T* du) {
        const double f1 = basis.bnd_lo, f2 = basis.bnd_hi;
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) {
            for (int l = 0; l < n; ++l)
                for (int v = 0; v < NV; ++v) {
                    du[idx(v, 0, l, e)] = du[idx(v, 0, l, e)] - this->sfv[sidx(v, l, 0, e)] * f1;
                    du[idx(v, n - 1, l, e)] = du[idx(v, n - 1, l, e)] + this->sfv[sidx(v, l, 1, e)] * f2;
                    du[idx(v, l, 0, e)] = du[idx(v, l, 0, e)] - this->sfv[sidx(v, l, 2, e)] * f1;
                    du[idx(v, l, n - 1, e)] = du[idx(v, l, n - 1, e)] + this->sfv[sidx(v, l, 3, e)] * f2;
                }
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const double factor = -ijn[nidx(i, j, e)];
                    for (int v = 0; v < NV; ++v) du[idx(v, i, j, e)] = du[idx(v, i, j, e)] * factor;
                }
        }
    }

    void rhs_lw(T* du, const T* u, double t, double dt) override {
        this->src_t = t;
        this->reset_du(du);
        if (this->sc) this->sc_prepare(u, dt);
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) {
            lw_volume_curved(du, u, dt, e, this->sc ? 1.0 - this->alpha[e] : 1.0);
            if (this->sc) sc_element_curved(du, u, e);
        }
        surface_fluxes(u, t, dt, 1.0);
        surface_integral_and_jacobian_curved(du);
    }

    // mdrk_kernel_1! for StructuredMesh (structured/multi_derivative_2d.jl:82-243):
    // Cartesian F = f/2 + f_t/8 (element cache) and F2 = f + f_t/6 (stage cache),
    // contravariant F̃ = cv_f/2 + cv_f_t/8 accumulated separately from the
    // contravariant fluxes of each Taylor state; U = u/2 + ut/8, U2 = u + ut/6.
    // The low-order update u_np1_low (an adaptive-mode output) is left out.
    void mdrk_curved_1(T* du, const T* u, double dt, int e, double alpha_dg = 1.0) {
        const auto& D = basis.D;
        const auto& Dh = basis.Dhat;
        T ut[MAXN][MAXN][NV], Ft[MAXN][MAXN][NV], Gt[MAXN][MAXN][NV];
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) ut[j][i][v] = T(0);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const T* un = &u[idx(0, i, j, e)];
                T f[NV], g[NV], fc[NV], gc[NV];
                cv_fluxes(un, &Ja[4 * nidx(i, j, e)], f, g, fc, gc);
                for (int v = 0; v < NV; ++v) {
                    this->F[fidx(v, 0, i, j, e)] = 0.5 * f[v];
                    this->F[fidx(v, 1, i, j, e)] = 0.5 * g[v];
                    this->F2[fidx(v, 0, i, j, e)] = f[v];
                    this->F2[fidx(v, 1, i, j, e)] = g[v];
                    Ft[j][i][v] = 0.5 * fc[v];
                    Gt[j][i][v] = 0.5 * gc[v];
                    this->U[idx(v, i, j, e)] = 0.5 * un[v];
                    this->U2[idx(v, i, j, e)] = un[v];
                }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) ut[j][ii][v] = ut[j][ii][v] + (-dt * D[ii][i]) * fc[v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) ut[jj][i][v] = ut[jj][i][v] + (-dt * D[jj][j]) * gc[v];
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) ut[j][i][v] = ut[j][i][v] * ijn[nidx(i, j, e)];
        // source terms (multi_derivative_2d.jl:146-155, 179-200): ut += dt s, S = s/2 + s_t/8
        // into du, S2 = s + s_t/6 to the stage cache (s_t at the stencil states about t)
        const bool has_src = this->src.kind != SRC_NONE;
        T S1[MAXN][MAXN][NV];
        if (has_src)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double x, y;
                    node_xy(i, j, e, x, y);
                    T sv[NV];
                    this->src_at(&u[idx(0, i, j, e)], x, y, 0.0, sv);
                    for (int v = 0; v < NV; ++v) {
                        S1[j][i][v] = 0.5 * sv[v];
                        this->S2[idx(v, i, j, e)] = sv[v];
                        ut[j][i][v] = ut[j][i][v] + dt * sv[v];
                    }
                }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                T ft[NV], gt[NV], fct[NV], gct[NV];
                mdrk_curved_ft(&u[idx(0, i, j, e)], ut[j][i], &Ja[4 * nidx(i, j, e)], ft, gt, fct, gct);
                if (has_src) {
                    double x, y;
                    node_xy(i, j, e, x, y);
                    T up[NV], um[NV], upp[NV], umm[NV], stv[NV];
                    for (int v = 0; v < NV; ++v) {
                        const T w = u[idx(v, i, j, e)], d = ut[j][i][v];
                        up[v] = w + d;
                        um[v] = w - d;
                        upp[v] = w + 2.0 * d;
                        umm[v] = w - 2.0 * d;
                    }
                    this->src_t_N34(up, upp, um, umm, x, y, dt, this->src_t, stv);
                    for (int v = 0; v < NV; ++v) {
                        S1[j][i][v] = S1[j][i][v] + 0.125 * stv[v];
                        this->S2[idx(v, i, j, e)] = this->S2[idx(v, i, j, e)] + (1.0 / 6.0) * stv[v];
                        du[idx(v, i, j, e)] = du[idx(v, i, j, e)] + (-1.0 / ijn[nidx(i, j, e)]) * S1[j][i][v];
                    }
                }
                for (int v = 0; v < NV; ++v) {
                    this->F[fidx(v, 0, i, j, e)] = this->F[fidx(v, 0, i, j, e)] + 0.125 * ft[v];
                    this->F[fidx(v, 1, i, j, e)] = this->F[fidx(v, 1, i, j, e)] + 0.125 * gt[v];
                    Ft[j][i][v] = Ft[j][i][v] + 0.125 * fct[v];
                    Gt[j][i][v] = Gt[j][i][v] + 0.125 * gct[v];
                    this->U[idx(v, i, j, e)] = this->U[idx(v, i, j, e)] + 0.125 * ut[j][i][v];
                    this->F2[fidx(v, 0, i, j, e)] = this->F2[fidx(v, 0, i, j, e)] + (1.0 / 6.0) * ft[v];
                    this->F2[fidx(v, 1, i, j, e)] = this->F2[fidx(v, 1, i, j, e)] + (1.0 / 6.0) * gt[v];
                    this->U2[idx(v, i, j, e)] = this->U2[idx(v, i, j, e)] + (1.0 / 6.0) * ut[j][i][v];
                }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, ii, j, e)] = du[idx(v, ii, j, e)] + (alpha_dg * Dh[ii][i]) * Ft[j][i][v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, i, jj, e)] = du[idx(v, i, jj, e)] + (alpha_dg * Dh[jj][j]) * Gt[j][i][v];
            }
    }

    // f_t of the MDRK stages, Cartesian and contravariant, from the contravariant
    // fluxes of the four Taylor states u ± ut, u ± 2ut (multi_derivative_2d.jl:161-176)
    void mdrk_curved_ft(const T* un, const T* ut, const double* J, T* ft, T* gt, T* fct, T* gct) const {
        T um[NV], up[NV], umm[NV], upp[NV];
        for (int v = 0; v < NV; ++v) {
            um[v] = un[v] - ut[v];
            up[v] = un[v] + ut[v];
            umm[v] = un[v] - 2.0 * ut[v];
            upp[v] = un[v] + 2.0 * ut[v];
        }
        T fm[NV], gm[NV], cfm[NV], cgm[NV], fp[NV], gp[NV], cfp[NV], cgp[NV];
        T fmm[NV], gmm[NV], cfmm[NV], cgmm[NV], fpp[NV], gpp[NV], cfpp[NV], cgpp[NV];
        cv_fluxes(um, J, fm, gm, cfm, cgm);
        cv_fluxes(up, J, fp, gp, cfp, cgp);
        cv_fluxes(umm, J, fmm, gmm, cfmm, cgmm);
        cv_fluxes(upp, J, fpp, gpp, cfpp, cgpp);
        for (int v = 0; v < NV; ++v) {
            ft[v] = 1.0 / 12.0 * (-fpp[v] + 8.0 * fp[v] - 8.0 * fm[v] + fmm[v]);
            gt[v] = 1.0 / 12.0 * (-gpp[v] + 8.0 * gp[v] - 8.0 * gm[v] + gmm[v]);
            fct[v] = 1.0 / 12.0 * (-cfpp[v] + 8.0 * cfp[v] - 8.0 * cfm[v] + cfmm[v]);
            gct[v] = 1.0 / 12.0 * (-cgpp[v] + 8.0 * cgp[v] - 8.0 * cgm[v] + cgmm[v]);
        }
    }

    // mdrk_kernel_2! for StructuredMesh (structured/multi_derivative_2d.jl:245-445):
    // F = F2 + f_t(us)/3 (Cartesian), F̃ = Ja·F2 + cv_f_t(us)/3, U = U2 + ust/3.
    void mdrk_curved_2(T* du, double dt, int e, double alpha_dg = 1.0) {
        const auto& D = basis.D;
        const auto& Dh = basis.Dhat;
        T ust[MAXN][MAXN][NV], Ft[MAXN][MAXN][NV], Gt[MAXN][MAXN][NV];
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const double* J = &Ja[4 * nidx(i, j, e)];
                for (int v = 0; v < NV; ++v) {
                    const T f2 = this->F2[fidx(v, 0, i, j, e)], g2 = this->F2[fidx(v, 1, i, j, e)];
                    this->F[fidx(v, 0, i, j, e)] = f2;
                    this->F[fidx(v, 1, i, j, e)] = g2;
                    Ft[j][i][v] = J[0] * f2 + J[1] * g2;
                    Gt[j][i][v] = J[2] * f2 + J[3] * g2;
                    this->U[idx(v, i, j, e)] = this->U2[idx(v, i, j, e)];
                    ust[j][i][v] = T(0);
                }
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                T f[NV], g[NV], fc[NV], gc[NV];
                cv_fluxes(&this->us[idx(0, i, j, e)], &Ja[4 * nidx(i, j, e)], f, g, fc, gc);
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) ust[j][ii][v] = ust[j][ii][v] + (-dt * D[ii][i]) * fc[v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) ust[jj][i][v] = ust[jj][i][v] + (-dt * D[jj][j]) * gc[v];
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) ust[j][i][v] = ust[j][i][v] * ijn[nidx(i, j, e)];
        // source terms (multi_derivative_2d.jl:326-339, 358-385): ust += dt s(us, t) and S =
        // s(us, t) + s_t/3 about t -- the reference loads S2 into S and then overwrites S
        // with s, so stage 1's s + s_t/6 does not reach the step (restated as is)
        const bool has_src = this->src.kind != SRC_NONE;
        T Sn[MAXN][MAXN][NV];
        if (has_src)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double x, y;
                    node_xy(i, j, e, x, y);
                    T sv[NV];
                    this->src_at(&this->us[idx(0, i, j, e)], x, y, 0.0, sv);
                    for (int v = 0; v < NV; ++v) {
                        Sn[j][i][v] = sv[v];
                        ust[j][i][v] = ust[j][i][v] + dt * sv[v];
                    }
                }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                T ft[NV], gt[NV], fct[NV], gct[NV];
                mdrk_curved_ft(&this->us[idx(0, i, j, e)], ust[j][i], &Ja[4 * nidx(i, j, e)], ft, gt, fct, gct);
                if (has_src) {
                    double x, y;
                    node_xy(i, j, e, x, y);
                    T up[NV], um[NV], upp[NV], umm[NV], stv[NV];
                    for (int v = 0; v < NV; ++v) {
                        const T w = this->us[idx(v, i, j, e)], d = ust[j][i][v];
                        up[v] = w + d;
                        um[v] = w - d;
                        upp[v] = w + 2.0 * d;
                        umm[v] = w - 2.0 * d;
                    }
                    this->src_t_N34(up, upp, um, umm, x, y, dt, this->src_t, stv);
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, i, j, e)] = du[idx(v, i, j, e)] +
                                              (-1.0 / ijn[nidx(i, j, e)]) * (Sn[j][i][v] + (1.0 / 3.0) * stv[v]);
                }
                for (int v = 0; v < NV; ++v) {
                    this->F[fidx(v, 0, i, j, e)] = this->F[fidx(v, 0, i, j, e)] + (1.0 / 3.0) * ft[v];
                    this->F[fidx(v, 1, i, j, e)] = this->F[fidx(v, 1, i, j, e)] + (1.0 / 3.0) * gt[v];
                    Ft[j][i][v] = Ft[j][i][v] + (1.0 / 3.0) * fct[v];
                    Gt[j][i][v] = Gt[j][i][v] + (1.0 / 3.0) * gct[v];
                    this->U[idx(v, i, j, e)] = this->U[idx(v, i, j, e)] + (1.0 / 3.0) * ust[j][i][v];
                }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, ii, j, e)] = du[idx(v, ii, j, e)] + (alpha_dg * Dh[ii][i]) * Ft[j][i][v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, i, jj, e)] = du[idx(v, i, jj, e)] + (alpha_dg * Dh[jj][j]) * Gt[j][i][v];
            }
    }

    // rhs_mdrk1! / rhs_mdrk2! for StructuredMesh (multi_derivative_2d.jl:1-80):
    // both stages evaluate the interface flux with λ of uⁿ
    void rhs_mdrk1(T* du, const T* u, double t, double dt) override {
        this->src_t = t;
        this->ensure_mdrk();
        this->reset_du(du);
        if (this->sc) this->sc_prepare(u, dt);
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) {
            mdrk_curved_1(du, u, dt, e, this->sc ? 1.0 - this->alpha[e] : 1.0);
            if (this->sc) sc_element_curved(du, u, e);
        }
        surface_fluxes(u, t, dt, 0.5);
        surface_integral_and_jacobian_curved(du);
    }
    void rhs_mdrk2(T* du, const T* u, double t, double dt) override {
        this->src_t = t;
        this->ensure_mdrk();
        this->reset_du(du);
        if (this->sc) this->sc_prepare(u, dt);
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) {
            mdrk_curved_2(du, dt, e, this->sc ? 1.0 - this->alpha[e] : 1.0);
            if (this->sc) sc_element_curved(du, u, e);
        }
        surface_fluxes(u, t, dt, 1.0);
        surface_integral_and_jacobian_curved(du);
    }

    // Trixi max_dt for curvilinear meshes: per node the speeds transformed to the
    // reference element, |Ja^k · (λ1, λ2)| |J⁻¹|, element maxima summed,
    // 2 / (nnodes · max)
    double max_dt(const T* u) const override {
        double mss = std::nextafter(0.0, 1.0);
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
                }
            mss = max_(mss, m1 + m2);
        }
        return 2.0 / (n * mss);
    }
};

}  // namespace tlwo
