// tlw_oracle_capi.cpp -- C entry points of the CPU oracle for ctypes.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker.  Never part of the product.
#include <omp.h>

#include <cstdio>
#include <memory>

#include "tlw_oracle_curved.hpp"
#include "tlw_oracle_curved_parabolic.hpp"
#include "tlw_oracle_parabolic.hpp"

using namespace tlwo;

struct tlwo_desc {
    int polydeg;
    int nx, ny;
    double xmin, ymin, h;  // lower-left corner and element edge length
    int periodic_x, periodic_y;
    int equation;  // 0 compressible Euler, 1 linear scalar advection
    double gamma;
    double adv_a1, adv_a2;
    int flux_order;  // 0 TreeMesh, 1 Structured/P4est argument order
    int bc_kind;     // BoundaryValueKind
    double bc_par[24];
    double abstol, reltol;
    int nthreads;    // 0: leave OpenMP's default
    int shock_capturing;
    double alpha_max, alpha_min;
    int alpha_smooth;
    int positivity;              // Zhang-Shu limiter on (density, pressure) in solve
    double pos_threshold[2];
    // SemidiscretizationHyperbolicParabolic: 0 none, 1 LaplaceDiffusion2D (with
    // linear advection), 2 CompressibleNavierStokesDiffusion2D with
    // GradientVariablesConservative (with Euler); periodic meshes only
    int parabolic;
    double visc_mu;  // diffusivity (Laplace) or dynamic viscosity μ (Navier-Stokes)
    double prandtl;
    // StructuredMesh(cells, mapping): 0 TreeMesh, 1 curvilinear with map_kind
    // (MappingKind) and map_par; plain FR, LW or MDRK, periodic or with the
    // time-averaged Dirichlet boundary
    int curved;
    int map_kind;
    double map_par[4];
    // shock-capturing reconstruction: 0 FirstOrderReconstruction,
    // 1 MUSCLHancockReconstruction (Euler only)
    int reconstruction;
    // Navier-Stokes with no-slip adiabatic walls on every boundary face (LW; the
    // hyperbolic part takes bc_kind = BV_SLIP_WALL): wall velocity (v1, v2) per
    // direction x_neg, x_pos, y_neg, y_pos
    int ns_walls;
    double wall_v[8];
    int wall_iso[4];     // 1: Isothermal with wall_T, 0: Adiabatic
    double wall_T[4];
    // bc_kind = BV_CALLBACK: the boundary value function u(x, y, t)
    void (*bc_fn)(double, double, double, double*);
    // source terms (SourceKind; Euler on the TreeMesh, or SRC_NS_CONVERGENCE with
    // the Navier-Stokes TreeMesh kernels): SRC_CALLBACK calls src_fn(u, x, y, t, s)
    int source_kind;
    void (*src_fn)(const double*, double, double, double, double*);
    // a TreeMesh with refined cells instead of the uniform nx x ny grid (n_tree > 0):
    // element centres, edge lengths, face neighbours (-1 boundary, -2 on a mortar)
    // and n_mortars records (lower, upper, large, orientation, large_side); plain FR,
    // hyperbolic only
    int n_tree;
    const double* tree_cx;
    const double* tree_cy;
    const double* tree_len;
    const int* tree_nbr;
    int n_mortars;
    const int* mortars;
    // P4estMesh{2} of the curvilinear mapping's nx x ny cells (curved = 1): each
    // element's frame turned by rot[e]·90° (rot may be null), P4est interfaces
    // and boundaries (Euler or linear advection, plain FR, hyperbolic only)
    int p4est;
    const int* rot;
};

struct tlwo_ctx {
    int equation;
    std::unique_ptr<Solver<double, Euler2D<double>>> euler;
    std::unique_ptr<Solver<double, LinAdv2D<double>>> linadv;
};

template <class S> static void configure(S& s, const tlwo_desc* d) {
    s.basis.init(d->polydeg);
    if (d->n_tree > 0)
        s.mesh.init_general(d->n_tree, d->tree_cx, d->tree_cy, d->tree_len, d->tree_nbr, d->n_mortars, d->mortars);
    else
        s.mesh.init(d->nx, d->ny, d->xmin, d->ymin, d->h, d->periodic_x != 0, d->periodic_y != 0);
    s.flux_order = d->flux_order;
    s.bval.kind = d->bc_kind;
    for (int k = 0; k < 24; ++k) s.bval.par[k] = d->bc_par[k];
    s.bval.fn = d->bc_fn;
    s.abstol = d->abstol;
    s.reltol = d->reltol;
    s.sc = d->shock_capturing != 0;
    s.alpha_max = d->alpha_max;
    s.alpha_min = d->alpha_min;
    s.alpha_smooth = d->alpha_smooth != 0;
    s.muscl_hancock = d->reconstruction == 1;
    s.positivity = d->positivity != 0;
    s.pos_threshold[0] = d->pos_threshold[0];
    s.pos_threshold[1] = d->pos_threshold[1];
    s.src.kind = d->source_kind;
    s.src.gamma = d->gamma;
    s.src.fn = d->src_fn;
    s.src.mu = d->visc_mu;
    s.src.prandtl = d->prandtl;
    s.setup();
}

extern "C" {

void* tlwo_create(const tlwo_desc* d) {
    if (d->polydeg != 3 && d->polydeg != 4) return nullptr;
    if (d->nthreads > 0) omp_set_num_threads(d->nthreads);
    if (d->parabolic) {
        const bool walls = d->ns_walls && d->parabolic == 2 && d->bc_kind == BV_SLIP_WALL && !d->curved;
        if ((!walls && (!d->periodic_x || !d->periodic_y)) || d->shock_capturing) return nullptr;
        if ((d->parabolic == 1) != (d->equation == 1)) return nullptr;
    }
    // curvilinear shock capturing: first order; Trixi's alpha smoothing on a
    // StructuredMesh asserts a periodic mesh
    const bool periodic = d->periodic_x && d->periodic_y;
    if (d->curved && d->parabolic == 1) return nullptr;
    if (d->p4est && (!d->curved || d->parabolic || d->shock_capturing || d->bc_kind == BV_SLIP_WALL)) return nullptr;
    if (d->curved && d->shock_capturing && (d->reconstruction != 0 || (d->alpha_smooth && !periodic))) return nullptr;
    if (d->reconstruction == 1 && (d->equation != 0 || !d->shock_capturing)) return nullptr;
    // the reflecting wall is the Euler TreeMesh one (x faces vertical, y faces horizontal)
    if (d->bc_kind == BV_SLIP_WALL && (d->equation != 0 || d->curved)) return nullptr;
    // the outflow and approximate slip-wall conditions take a normal direction: curvilinear Euler
    if ((d->bc_kind == BV_OUTFLOW || d->bc_kind == BV_SLIP_WALL_APPROX) &&
        (d->equation != 0 || !d->curved || d->parabolic))
        return nullptr;
    if (d->bc_kind == BV_CALLBACK && !d->bc_fn) return nullptr;
    // refined TreeMesh: the hyperbolic plain-FR kernels (compute_alp and the indicator
    // smoothing read the uniform mesh's Jacobian and neighbours)
    if (d->n_tree > 0 && (d->curved || d->parabolic || d->shock_capturing || !d->tree_cx || !d->tree_cy ||
                          !d->tree_len || !d->tree_nbr || d->n_mortars < 0 || (d->n_mortars > 0 && !d->mortars)))
        return nullptr;
    // source terms: the Euler kernels of the TreeMesh (LW N = 3, 4 and MDRK); the
    // Navier-Stokes convergence source with the TreeMesh Navier-Stokes kernels
    if (d->source_kind < SRC_NONE || d->source_kind > SRC_NS_CONVERGENCE) return nullptr;
    // (a callback source may also go with the TreeMesh Navier-Stokes kernels, where the
    // oracle's pins give it sources of t alone)
    if (d->source_kind == SRC_NS_CONVERGENCE && (d->equation != 0 || d->parabolic != 2 || d->curved)) return nullptr;
    // (the curvilinear and P4est LW and MDRK kernels take them too, without shock capturing)
    const bool curved_src_ok = !d->curved || (!d->parabolic && !d->shock_capturing);
    if (d->source_kind == SRC_CONVERGENCE_TEST && (d->equation != 0 || d->parabolic || !curved_src_ok))
        return nullptr;
    if (d->source_kind == SRC_CALLBACK &&
        (d->equation != 0 || (d->parabolic && d->parabolic != 2) || !curved_src_ok || !d->src_fn))
        return nullptr;
    auto* c = new tlwo_ctx;
    c->equation = d->equation;
    if (d->curved && d->parabolic == 2) {
        auto* p = new CurvedParabolicSolver<double, Euler2D<double>, NavierStokesDiffusion2D<double>>;
        p->eq.gamma = d->gamma;
        p->visc.set(d->gamma, d->visc_mu, d->prandtl);
        configure(*p, d);
        p->setup_curved_parabolic(d->map_kind, d->map_par);
        c->euler.reset(p);
    } else if (d->curved) {
        if (d->equation == 0) {
            auto* p = new CurvedSolver<double, Euler2D<double>>;
            p->eq.gamma = d->gamma;
            configure(*p, d);
            p->setup_curved(d->map_kind, d->map_par, d->p4est ? d->rot : nullptr);
            if (d->p4est) p->setup_p4est();
            c->euler.reset(p);
        } else {
            auto* p = new CurvedSolver<double, LinAdv2D<double>>;
            p->eq.a1 = d->adv_a1;
            p->eq.a2 = d->adv_a2;
            configure(*p, d);
            p->setup_curved(d->map_kind, d->map_par, d->p4est ? d->rot : nullptr);
            if (d->p4est) p->setup_p4est();
            c->linadv.reset(p);
        }
    } else if (d->parabolic == 2) {
        auto* p = new ParabolicSolver<double, Euler2D<double>, NavierStokesDiffusion2D<double>>;
        p->eq.gamma = d->gamma;
        p->visc.set(d->gamma, d->visc_mu, d->prandtl);
        configure(*p, d);
        p->setup_parabolic();
        p->ns_walls = d->ns_walls;
        for (int k = 0; k < 4; ++k) {
            p->wall_v[k][0] = d->wall_v[2 * k];
            p->wall_v[k][1] = d->wall_v[2 * k + 1];
            p->wall_iso[k] = d->wall_iso[k];
            p->wall_T[k] = d->wall_T[k];
        }
        c->euler.reset(p);
    } else if (d->parabolic == 1) {
        auto* p = new ParabolicSolver<double, LinAdv2D<double>, LaplaceDiffusion2D<double>>;
        p->eq.a1 = d->adv_a1;
        p->eq.a2 = d->adv_a2;
        p->visc.diffusivity = d->visc_mu;
        configure(*p, d);
        p->setup_parabolic();
        c->linadv.reset(p);
    } else if (d->equation == 0) {
        c->euler.reset(new Solver<double, Euler2D<double>>);
        c->euler->eq.gamma = d->gamma;
        configure(*c->euler, d);
    } else {
        c->linadv.reset(new Solver<double, LinAdv2D<double>>);
        c->linadv->eq.a1 = d->adv_a1;
        c->linadv->eq.a2 = d->adv_a2;
        configure(*c->linadv, d);
    }
    return c;
}

void tlwo_destroy(void* p) { delete static_cast<tlwo_ctx*>(p); }

void tlwo_set_threads(int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
}

int tlwo_max_threads(void) { return omp_get_max_threads(); }

#define DISPATCH(ctx, call)                               \
    do {                                                  \
        auto* c_ = static_cast<tlwo_ctx*>(ctx);           \
        if (c_->equation == 0) { auto& s = *c_->euler; call; } \
        else { auto& s = *c_->linadv; call; }             \
    } while (0)

long long tlwo_ndofs(void* ctx) {
    long long r = 0;
    DISPATCH(ctx, r = (long long)s.ndofs_total());
    return r;
}

// Basis export: nodes, weights (n), D, Dhat (n*n, row-major [i][k])
void tlwo_basis(void* ctx, double* nodes, double* weights, double* D, double* Dhat) {
    DISPATCH(ctx, {
        const int n = s.n;
        for (int i = 0; i < n; ++i) {
            nodes[i] = s.basis.nodes[i];
            weights[i] = s.basis.weights[i];
            for (int k = 0; k < n; ++k) {
                D[i * n + k] = s.basis.D[i][k];
                Dhat[i * n + k] = s.basis.Dhat[i][k];
            }
        }
    });
}

// Trixi LobattoLegendreMortarL2 operators (n*n, row-major [output node][input node])
void tlwo_mortar_matrices(void* ctx, double* fwd_lo, double* fwd_up, double* rev_lo, double* rev_up) {
    DISPATCH(ctx, {
        const int n = s.n;
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) {
                fwd_lo[i * n + k] = s.basis.fwd_lo[i][k];
                fwd_up[i * n + k] = s.basis.fwd_up[i][k];
                rev_lo[i * n + k] = s.basis.rev_lo[i][k];
                rev_up[i * n + k] = s.basis.rev_up[i][k];
            }
    });
}

// P4est interfaces {primary, primary face, secondary, secondary face, flipped} and
// boundary faces {element, face} of a P4est context; returns the interface count
// (boundary count in *nb); rows are written when the buffers are given
int tlwo_p4est_connectivity(void* ctx, int* interfaces, int* boundaries, int* nb) {
    int r = -1;
    if (!ctx) return r;
    auto* c = static_cast<tlwo_ctx*>(ctx);
    auto fill = [&](auto* s) {
        if (!s || !s->p4est) return;
        r = (int)s->p4if.size();
        if (nb) *nb = (int)s->p4bd_e.size();
        for (size_t k = 0; interfaces && k < s->p4if.size(); ++k) {
            const auto& q = s->p4if[k];
            const int row[5] = {q.prim, q.pface, q.sec, q.sface, q.flip};
            for (int m = 0; m < 5; ++m) interfaces[5 * k + m] = row[m];
        }
        for (size_t k = 0; boundaries && k < s->p4bd_e.size(); ++k) {
            boundaries[2 * k] = s->p4bd_e[k];
            boundaries[2 * k + 1] = s->p4bd_f[k];
        }
    };
    if (c->euler) fill(dynamic_cast<CurvedSolver<double, Euler2D<double>>*>(c->euler.get()));
    if (c->linadv) fill(dynamic_cast<CurvedSolver<double, LinAdv2D<double>>*>(c->linadv.get()));
    return r;
}

// Node coordinates in [i, j, e] order
void tlwo_node_coords(void* ctx, double* x, double* y) {
    DISPATCH(ctx, {
        const int n = s.n;
        for (int e = 0; e < s.mesh.E; ++e)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const size_t k = ((size_t)e * n + j) * n + i;
                    s.node_xy(i, j, e, x[k], y[k]);
                }
    });
}

void tlwo_neighbors(void* ctx, int* nbr) {
    DISPATCH(ctx, {
        for (size_t k = 0; k < s.mesh.nbr.size(); ++k) nbr[k] = s.mesh.nbr[k];
    });
}

double tlwo_max_dt(void* ctx, const double* u) {
    double r = 0;
    DISPATCH(ctx, r = s.max_dt(u));
    return r;
}

void tlwo_rhs(void* ctx, const double* u, double* du, double t, double dt) {
    DISPATCH(ctx, s.rhs_lw(du, u, t, dt));
}

void tlwo_rhs_mdrk1(void* ctx, const double* u, double* du, double t, double dt) {
    DISPATCH(ctx, s.rhs_mdrk1(du, u, t, dt));
}

void tlwo_rhs_mdrk2(void* ctx, const double* u, double* du, double t, double dt) {
    DISPATCH(ctx, s.rhs_mdrk2(du, u, t, dt));
}

void tlwo_set_us(void* ctx, const double* us) {
    DISPATCH(ctx, {
        s.ensure_mdrk();
        std::copy(us, us + s.ndofs_total(), s.us.begin());
    });
}

// which: 0 = Ũ [v,i,j,e], 1 = F̃ [v,d,i,j,e], 2 = temporal_errors [e],
//        3 = surface_flux_values [v,l,face,e], 4 = us, 5 = U2, 6 = F2,
//        7 = alpha [e], 8 = fn_low [v,l,face,e],
//        hyperbolic-parabolic only: 9 = gradients_x, 10 = gradients_y [v,i,j,e],
//        11 = Fv [v,d,i,j,e], 12 = mdrk Fv2,
//        curvilinear only: 13 = contravariant vectors [4,i,j,e] (Ja11, Ja12, Ja21, Ja22),
//        14 = inverse Jacobian [i,j,e]
long long tlwo_get_aux(void* ctx, int which, double* out) {
    long long len = -1;
    DISPATCH(ctx, {
        const std::vector<double>* src = nullptr;
        switch (which) {
            case 0: src = &s.U; break;
            case 1: src = &s.F; break;
            case 2: src = &s.temporal_errors; break;
            case 3: src = &s.sfv; break;
            case 4: src = &s.us; break;
            case 5: src = &s.U2; break;
            case 6: src = &s.F2; break;
            case 7: src = &s.alpha; break;
            case 8: src = &s.fn_low; break;
            default: src = s.extra_aux(which); break;
        }
        if (src) {
            len = (long long)src->size();
            if (out) std::copy(src->begin(), src->end(), out);
        }
    });
    return len;
}

int tlwo_solve(void* ctx, double* u, double t0, double tf, double cfl, int max_steps, int mdrk, double* t_out,
               double* dts) {
    int r = 0;
    DISPATCH(ctx, r = solve(s, u, t0, tf, cfl, max_steps, mdrk, t_out, dts));
    return r;
}

// Adaptive LW loop (tolerances from the descriptor's abstol/reltol); returns accepted steps.
int tlwo_solve_adaptive(void* ctx, double* u, double t0, double tf, double dt_initial, int max_attempts,
                        double* t_out, double* dts, long long* nreject) {
    int r = 0;
    DISPATCH(ctx, r = solve_adaptive(s, u, t0, tf, dt_initial, max_attempts, t_out, dts, nreject));
    return r;
}

// Prescribed blending factors (len E) used instead of the indicator; NULL clears.
void tlwo_positivity_limiter(void* ctx, double* u) {
    auto* c = static_cast<tlwo_ctx*>(ctx);
    if (c->euler) c->euler->positivity_limiter(u);
    else c->linadv->positivity_limiter(u);
}

void tlwo_set_alpha(void* ctx, const double* alpha) {
    DISPATCH(ctx, {
        if (alpha) s.alpha_override.assign(alpha, alpha + s.mesh.E);
        else s.alpha_override.clear();
    });
}

// Inverse Legendre Vandermonde of the basis (n*n, row-major [i][m]).
void tlwo_inverse_vandermonde(void* ctx, double* out) {
    DISPATCH(ctx, {
        for (int i = 0; i < s.n; ++i)
            for (int m = 0; m < s.n; ++m) out[i * s.n + m] = s.basis.Vinv[i][m];
    });
}

// Pointwise Euler pieces, for closed-form checks.
void tlwo_euler_fluxes(const double* u, double gamma, double* f, double* g) {
    Euler2D<double> eq;
    eq.gamma = gamma;
    eq.fluxes(u, f, g);
}

double tlwo_max_abs_speed_naive(const double* ul, const double* ur, int orientation, double gamma) {
    Euler2D<double> eq;
    eq.gamma = gamma;
    return eq.max_abs_speed_naive(ul, ur, orientation);
}


void tlwo_solve_fixed_dt(void* ctx, double* u, double t0, const double* dts, int nsteps) {
    auto* c = static_cast<tlwo_ctx*>(ctx);
    if (c->euler) solve_fixed_dt(*c->euler, u, t0, dts, nsteps);
    else solve_fixed_dt(*c->linadv, u, t0, dts, nsteps);
}

// Tangent of the fixed-dt LW loop: u0 + ε du0 -> u_K + ε du_K with dual numbers
// (the forward-mode gradient check; Euler only).
// Forward-mode tangent of the fixed-dt loop (LW, or MDRK with mdrk != 0);
// alpha (E values, or NULL) freezes the shock-capturing blending factors.
int tlwo_tangent(const tlwo_desc* d, const double* u0, const double* du0, const double* dts, int nsteps, int mdrk,
                 const double* alpha, double* u_out, double* du_out) {
    if (d->polydeg != 3 && d->polydeg != 4) return 1;
    if (d->source_kind == SRC_CALLBACK) return 2;  // a callback sees values only: no tangent
    if (d->nthreads > 0) omp_set_num_threads(d->nthreads);
    Solver<Dual, Euler2D<Dual>> s;
    s.eq.gamma = d->gamma;
    configure(s, d);
    if (alpha) s.alpha_override.assign(alpha, alpha + s.mesh.E);
    const size_t nd = s.ndofs_total();
    std::vector<Dual> u(nd);
    for (size_t k = 0; k < nd; ++k) u[k] = Dual(u0[k], du0[k]);
    solve_fixed_dt(s, u.data(), 0.0, dts, nsteps, mdrk != 0);
    for (size_t k = 0; k < nd; ++k) {
        u_out[k] = u[k].v;
        du_out[k] = u[k].d;
    }
    return 0;
}

int tlwo_tangent_lw(const tlwo_desc* d, const double* u0, const double* du0, const double* dts, int nsteps,
                    double* u_out, double* du_out) {
    if (d->polydeg != 3 && d->polydeg != 4) return 1;
    if (d->nthreads > 0) omp_set_num_threads(d->nthreads);
    Solver<Dual, Euler2D<Dual>> s;
    s.eq.gamma = d->gamma;
    configure(s, d);
    const size_t nd = s.ndofs_total();
    std::vector<Dual> u(nd);
    for (size_t k = 0; k < nd; ++k) u[k] = Dual(u0[k], du0[k]);
    solve_fixed_dt(s, u.data(), 0.0, dts, nsteps);
    for (size_t k = 0; k < nd; ++k) {
        u_out[k] = u[k].v;
        du_out[k] = u[k].d;
    }
    return 0;
}
}  // extern "C"

extern "C" {
// Pointwise viscous flux of CompressibleNavierStokesDiffusion2D with conservative
// gradient variables, for the chain-rule check.
// the Navier-Stokes manufactured solution and its source at n points
void tlwo_ns_convergence(int n, const double* x, const double* y, double t, double gamma, double mu, double prandtl,
                         double* u, double* s) {
    for (int k = 0; k < n; ++k) {
        ns_convergence_state(x[k], y[k], t, gamma, u + 4 * k);
        ns_convergence_source(x[k], y[k], t, gamma, mu, prandtl, s + 4 * k);
    }
}

void tlwo_ns_viscous_fluxes(const double* u, const double* gx, const double* gy, double gamma, double mu,
                            double prandtl, double* fx, double* fy) {
    NavierStokesDiffusion2D<double> v;
    v.set(gamma, mu, prandtl);
    v.fluxes<4>(u, gx, gy, fx, fy);
}
}  // extern "C"
