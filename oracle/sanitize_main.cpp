// sanitize_main.cpp -- host-only driver that runs every entry point of the CPU
// oracle on small meshes, for AddressSanitizer / UndefinedBehaviorSanitizer
// builds (`make -C oracle sanitize`).
//
// TEST INFRASTRUCTURE ONLY: it exercises the checker, never the product.  It
// includes the C API's translation unit so that one executable holds the whole
// oracle; every scenario checks that its outputs are finite, and the
// sanitizers report any out-of-bounds access, use after free, leak, signed
// overflow or misaligned load on the way.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "tlw_oracle_capi.cpp"

namespace {

int g_failures = 0;

void check_finite(const char* what, const std::vector<double>& v) {
    for (double x : v)
        if (!std::isfinite(x)) {
            std::printf("FAIL %s: non-finite value\n", what);
            ++g_failures;
            return;
        }
    std::printf("ok   %s (%zu values)\n", what, v.size());
}

tlwo_desc base(int N, int nx, double xmin, double h) {
    tlwo_desc d;
    std::memset(&d, 0, sizeof d);
    d.polydeg = N;
    d.nx = d.ny = nx;
    d.xmin = d.ymin = xmin;
    d.h = h;
    d.periodic_x = d.periodic_y = 1;
    d.gamma = 1.4;
    d.adv_a1 = 1.0;
    d.adv_a2 = 0.5;
    d.nthreads = 2;
    d.alpha_max = 0.5;
    d.alpha_min = 0.001;
    d.alpha_smooth = 1;
    d.prandtl = 0.72;
    return d;
}

// conservative state of (ρ, v1, v2, p)
void prim(double rho, double v1, double v2, double p, double* u) {
    u[0] = rho;
    u[1] = rho * v1;
    u[2] = rho * v2;
    u[3] = p / 0.4 + 0.5 * rho * (v1 * v1 + v2 * v2);
}

// node states of a context from a pointwise function of (x, y)
template <class F> std::vector<double> nodal(void* ctx, F f) {
    const long long nd = tlwo_ndofs(ctx);
    const size_t nn = (size_t)nd / 4;
    std::vector<double> x(nn), y(nn), u(nd);
    tlwo_node_coords(ctx, x.data(), y.data());
    for (size_t k = 0; k < nn; ++k) f(x[k], y[k], &u[4 * k]);
    return u;
}

void smooth(double x, double y, double* u) {
    prim(1.0 + 0.2 * std::sin(0.3 * x) * std::cos(0.2 * y), 0.3 + 0.1 * std::sin(0.2 * y),
         -0.2 + 0.1 * std::cos(0.3 * x), 1.0 + 0.1 * std::cos(0.25 * (x + y)), u);
}

void riemann(double x, double y, double* u) {
    if (x >= 0.5) {
        if (y >= 0.5) prim(1.5, 0.0, 0.0, 1.5, u);
        else prim(0.5323, 0.0, 1.206, 0.3, u);
    } else {
        if (y >= 0.5) prim(0.5323, 1.206, 0.0, 0.3, u);
        else prim(0.138, 1.206, 1.206, 0.029, u);
    }
}

void bc_callback(double x, double y, double t, double* u) { smooth(x + 0.1 * t, y, u); }

// rhs of every kind, max_dt, a few CFL steps and the aux arrays
void exercise(const char* name, tlwo_desc d, void (*ic)(double, double, double*), double cfl, bool mdrk,
              int steps = 3) {
    void* c = tlwo_create(&d);
    if (!c) {
        std::printf("FAIL %s: tlwo_create refused\n", name);
        ++g_failures;
        return;
    }
    auto u = nodal(c, ic);
    std::vector<double> du(u.size()), du2(u.size());
    const double dt = cfl * tlwo_max_dt(c, u.data());
    char what[160];
    if (mdrk) {
        tlwo_rhs_mdrk1(c, u.data(), du.data(), 0.0, dt);
        std::vector<double> us(u);
        for (size_t k = 0; k < u.size(); ++k) us[k] += dt * du[k];
        tlwo_set_us(c, us.data());
        tlwo_rhs_mdrk2(c, u.data(), du2.data(), 0.0, dt);
        std::snprintf(what, sizeof what, "%s rhs_mdrk2", name);
        check_finite(what, du2);
    } else {
        tlwo_rhs(c, u.data(), du.data(), 0.0, dt);
    }
    std::snprintf(what, sizeof what, "%s rhs", name);
    check_finite(what, du);
    for (int which = 0; which < 12; ++which) {
        const long long len = tlwo_get_aux(c, which, nullptr);
        if (len <= 0) continue;
        std::vector<double> a((size_t)len);
        tlwo_get_aux(c, which, a.data());
    }
    std::vector<double> dts((size_t)steps);
    double tout = 0.0;
    const int k = tlwo_solve(c, u.data(), 0.0, 1e9, cfl, steps, mdrk ? 1 : 0, &tout, dts.data());
    std::snprintf(what, sizeof what, "%s %d steps", name, k);
    check_finite(what, u);
    tlwo_destroy(c);
}

void tangents(int N) {
    tlwo_desc d = base(N, 4, -10.0, 5.0);
    void* c = tlwo_create(&d);
    auto u = nodal(c, smooth);
    const double dt = 0.3 * tlwo_max_dt(c, u.data());
    tlwo_destroy(c);
    std::vector<double> du0(u.size()), uo(u.size()), duo(u.size());
    for (size_t k = 0; k < u.size(); ++k) du0[k] = std::sin(0.37 * (double)k);
    const double dts[3] = {dt, dt, dt};
    tlwo_tangent_lw(&d, u.data(), du0.data(), dts, 3, uo.data(), duo.data());
    check_finite("tangent LW", duo);
    tlwo_tangent(&d, u.data(), du0.data(), dts, 3, 1, nullptr, uo.data(), duo.data());
    check_finite("tangent MDRK", duo);
    d.shock_capturing = 1;
    std::vector<double> alpha((size_t)d.nx * d.ny);
    for (size_t e = 0; e < alpha.size(); ++e) alpha[e] = 0.1 * (double)(e % 5);
    tlwo_tangent(&d, u.data(), du0.data(), dts, 3, 0, alpha.data(), uo.data(), duo.data());
    check_finite("tangent LW, frozen blend", duo);
}

// A periodic TreeMesh of [-10, 10]² with 4 x 4 cells whose central 2 x 2 are
// refined once (28 leaves, 8 mortars), connected by brute-force geometry.
struct Refined {
    std::vector<double> cx, cy, len;
    std::vector<int> nbr, mort;
};
Refined refined_mesh() {
    Refined m;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) {
            const double x = -10.0 + 5.0 * (i + 0.5), y = -10.0 + 5.0 * (j + 0.5);
            if (i >= 1 && i <= 2 && j >= 1 && j <= 2) {
                for (int b = 0; b < 2; ++b)
                    for (int a = 0; a < 2; ++a) {
                        m.cx.push_back(x + 2.5 * (a - 0.5));
                        m.cy.push_back(y + 2.5 * (b - 0.5));
                        m.len.push_back(2.5);
                    }
            } else {
                m.cx.push_back(x);
                m.cy.push_back(y);
                m.len.push_back(5.0);
            }
        }
    const int E = (int)m.cx.size();
    m.nbr.assign(4 * (size_t)E, -3);
    for (int e = 0; e < E; ++e)
        for (int f = 0; f < 4; ++f) {
            const int ax = f / 2;
            const double sgn = (f % 2) ? 1.0 : -1.0;
            const double* cn = ax ? m.cy.data() : m.cx.data();
            const double* ct = ax ? m.cx.data() : m.cy.data();
            double pos = cn[e] + sgn * 0.5 * m.len[e];
            if (pos > 10.0 - 1e-9) pos -= 20.0;
            if (pos < -10.0 + 1e-9) pos += 20.0;
            std::vector<int> cand;
            for (int k = 0; k < E; ++k) {
                double opp = cn[k] - sgn * 0.5 * m.len[k];
                if (opp > 10.0 - 1e-9) opp -= 20.0;
                if (opp < -10.0 + 1e-9) opp += 20.0;
                const double ov = std::min(ct[k] + 0.5 * m.len[k], ct[e] + 0.5 * m.len[e]) -
                                  std::max(ct[k] - 0.5 * m.len[k], ct[e] - 0.5 * m.len[e]);
                if (std::fabs(opp - pos) < 1e-9 && ov > 1e-9) cand.push_back(k);
            }
            if (cand.size() == 1 && m.len[cand[0]] == m.len[e]) {
                m.nbr[4 * e + f] = cand[0];
            } else {
                m.nbr[4 * e + f] = -2;
                if (cand.size() == 2) {
                    const int lo = ct[cand[0]] < ct[cand[1]] ? cand[0] : cand[1];
                    const int up = lo == cand[0] ? cand[1] : cand[0];
                    const int rec[5] = {lo, up, e, ax + 1, sgn > 0 ? 1 : 2};
                    m.mort.insert(m.mort.end(), rec, rec + 5);
                }
            }
        }
    return m;
}

}  // namespace

int main() {
    for (int N = 3; N <= 4; ++N) {
        std::printf("-- polydeg %d\n", N);
        tlwo_desc d = base(N, 4, -10.0, 5.0);
        exercise("euler LW", d, smooth, 0.3, false);
        exercise("euler MDRK", d, smooth, 0.15, true);
        d.flux_order = 1;
        exercise("euler LW structured order", d, smooth, 0.3, false);

        // four-quadrant Riemann problem with the Dirichlet boundary, shock
        // capturing (first order and MUSCL-Hancock) and the positivity limiter
        tlwo_desc r = base(N, 8, 0.0, 1.0 / 8);
        r.periodic_x = r.periodic_y = 0;
        r.bc_kind = BV_RIEMANN4;
        const double rp[19] = {0.5, 0.5, 1.4, 1.5, 0.0, 0.0, 1.5, 0.5323, 1.206, 0.0, 0.3,
                               0.138, 1.206, 1.206, 0.029, 0.5323, 0.0, 1.206, 0.3};
        std::memcpy(r.bc_par, rp, sizeof rp);
        r.shock_capturing = 1;
        r.alpha_smooth = 0;
        r.positivity = 1;
        r.pos_threshold[0] = r.pos_threshold[1] = 1e-10;
        exercise("riemann FO blend + limiter", r, riemann, 0.3, false, 5);
        exercise("riemann FO blend MDRK", r, riemann, 0.15, true, 5);
        r.reconstruction = 1;
        exercise("riemann MUSCL-Hancock", r, riemann, 0.1, false, 5);
        exercise("riemann MUSCL-Hancock MDRK", r, riemann, 0.05, true, 5);

        // reflecting walls and a caller-defined Dirichlet function
        tlwo_desc w = base(N, 4, -1.0, 0.5);
        w.periodic_x = w.periodic_y = 0;
        w.bc_kind = BV_SLIP_WALL;
        exercise("slip walls", w, smooth, 0.3, false);
        w.bc_kind = BV_CALLBACK;
        w.bc_fn = bc_callback;
        exercise("callback Dirichlet", w, smooth, 0.3, false);
        exercise("callback Dirichlet MDRK", w, smooth, 0.15, true);

        // Navier-Stokes: periodic, no-slip adiabatic and isothermal walls
        tlwo_desc ns = base(N, 4, -1.0, 0.5);
        ns.parabolic = 2;
        ns.visc_mu = 0.01;
        exercise("navier-stokes LW", ns, smooth, 0.5, false);
        exercise("navier-stokes MDRK", ns, smooth, 0.25, true);
        ns.periodic_x = ns.periodic_y = 0;
        ns.bc_kind = BV_SLIP_WALL;
        ns.ns_walls = 1;
        ns.wall_v[6] = 1.0;
        exercise("navier-stokes adiabatic walls", ns, smooth, 0.5, false);
        for (int k = 0; k < 4; ++k) {
            ns.wall_iso[k] = 1;
            ns.wall_T[k] = 1.0;
        }
        exercise("navier-stokes isothermal walls", ns, smooth, 0.5, false);

        // linear advection and advection-diffusion
        tlwo_desc la = base(N, 4, -1.0, 0.5);
        la.equation = 1;
        auto scalar = [](double x, double y, double* u) {
            u[0] = std::sin(3.0 * x) * std::cos(2.0 * y);
            u[1] = u[2] = u[3] = 0.0;
        };
        {
            void* c = tlwo_create(&la);
            const long long nd = tlwo_ndofs(c);
            std::vector<double> x((size_t)nd), y((size_t)nd), u((size_t)nd), du((size_t)nd);
            tlwo_node_coords(c, x.data(), y.data());
            for (long long k = 0; k < nd; ++k) {
                double q[4];
                scalar(x[k], y[k], q);
                u[k] = q[0];
            }
            tlwo_rhs(c, u.data(), du.data(), 0.0, 1e-3);
            check_finite("linear advection rhs", du);
            tlwo_destroy(c);
            la.parabolic = 1;
            la.visc_mu = 0.01;
            c = tlwo_create(&la);
            tlwo_rhs(c, u.data(), du.data(), 0.0, 1e-3);
            check_finite("advection-diffusion rhs", du);
            tlwo_destroy(c);
        }

        // curvilinear mesh: periodic warp (LW, MDRK, blend, Navier-Stokes) and
        // the Dirichlet boundary on a non-periodic warp
        tlwo_desc cv = base(N, 4, 0.0, 5.0);
        cv.curved = 1;
        cv.map_kind = MAP_WARP;
        cv.map_par[0] = 20.0;
        cv.map_par[1] = 0.08;
        exercise("curved LW", cv, smooth, 0.3, false);
        exercise("curved MDRK", cv, smooth, 0.15, true);
        cv.shock_capturing = 1;
        exercise("curved FO blend", cv, smooth, 0.3, false);
        cv.shock_capturing = 0;
        cv.parabolic = 2;
        cv.visc_mu = 0.01;
        exercise("curved navier-stokes LW", cv, smooth, 0.5, false);
        exercise("curved navier-stokes MDRK", cv, smooth, 0.25, true);
        cv.parabolic = 0;
        cv.periodic_x = cv.periodic_y = 0;
        cv.bc_kind = BV_CALLBACK;
        cv.bc_fn = bc_callback;
        exercise("curved callback Dirichlet", cv, smooth, 0.3, false);

        // adaptive steps (TreeMesh LW N=4, curvilinear LW N=3/4) and the limiter
        {
            tlwo_desc a = N == 4 ? base(N, 4, -10.0, 5.0) : cv;
            if (N != 4) {
                a.periodic_x = a.periodic_y = 1;
                a.bc_kind = BV_NONE;
                a.bc_fn = nullptr;
            }
            a.abstol = a.reltol = 1e-6;
            void* c = tlwo_create(&a);
            auto u = nodal(c, smooth);
            std::vector<double> dts(50);
            double tout = 0.0;
            long long nrej = 0;
            tlwo_solve_adaptive(c, u.data(), 0.0, 0.5, 5.0, 50, &tout, dts.data(), &nrej);
            check_finite("adaptive loop", u);
            for (size_t k = 0; k < u.size(); k += 7) u[k] *= -0.01;
            tlwo_positivity_limiter(c, u.data());
            tlwo_destroy(c);
        }
        tangents(N);

        // TreeMesh with refined cells: L2 mortars in LW and MDRK
        {
            static Refined m;
            m = refined_mesh();
            tlwo_desc t = base(N, 1, 0.0, 1.0);
            t.n_tree = (int)m.cx.size();
            t.tree_cx = m.cx.data();
            t.tree_cy = m.cy.data();
            t.tree_len = m.len.data();
            t.tree_nbr = m.nbr.data();
            t.n_mortars = (int)m.mort.size() / 5;
            t.mortars = m.mort.data();
            if (t.n_mortars != 8) {
                std::printf("FAIL refined mesh: %d mortars\n", t.n_mortars);
                ++g_failures;
            }
            exercise("refined LW", t, smooth, 0.3, false);
            exercise("refined MDRK", t, smooth, 0.15, true);
        }
    }
    std::printf("%s: %d failure(s)\n", g_failures ? "FAILED" : "PASSED", g_failures);
    return g_failures ? 1 : 0;
}
