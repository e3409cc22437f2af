// tlw_oracle.hpp -- CPU restatement of TrixiLW's TreeMesh Lax-Wendroff / MDRK
// right-hand side.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the timed CPU
// baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it, and only as the checker.  The product (libtrixilw_hip.so)
// never links or calls it.
//
// Every function cites the reference file:line it restates (paths relative to
// the TrixiLW.jl source tree).  Functions that live in the un-vendored Trixi.jl
// dependency (basis, Euler flux, wave speeds, surface integral, Jacobian,
// max_dt) are restated from their published definitions; how each is pinned
// is listed next to it.
//
// Layouts follow Trixi's column-major arrays, first index fastest:
//   u, du        [v, i, j, e]        -> ((e*n + j)*n + i)*NV + v
//   U (Ũ)        [v, i, j, e]
//   F (F̃)        [v, d, i, j, e]     -> (((e*n + j)*n + i)*2 + d)*NV + v
//   sfv          [v, l, face, e]     -> ((e*4 + face)*n + l)*NV + v
// Faces are numbered as Trixi's `direction`: 0 = -x, 1 = +x, 2 = -y, 3 = +y.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace tlwo {

constexpr int MAXN = 8;

// ---------------------------------------------------------------------------
// Scalar helpers.  A forward-mode dual type provides the same overloads.
inline double sqrt_(double x) { return std::sqrt(x); }
inline double abs_(double x) { return std::fabs(x); }
inline double value(double x) { return x; }
// Julia Base.max for non-NaN floats: y if y > x, else x.
template <class T> inline T max_(const T& x, const T& y) { return (value(y) > value(x)) ? y : x; }

// Forward-mode dual number (value, one directional derivative), the role
// ForwardDiff.Dual plays for the reference: Solver<Dual, Euler2D<Dual>> runs
// the same code and carries d/dε of every quantity along a seed direction.
struct Dual {
    double v = 0.0, d = 0.0;
    Dual() = default;
    Dual(double x) : v(x), d(0.0) {}
    Dual(double x, double dx) : v(x), d(dx) {}
};
inline double value(const Dual& x) { return x.v; }
inline Dual operator+(const Dual& a, const Dual& b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(const Dual& a, const Dual& b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(const Dual& a) { return {-a.v, -a.d}; }
inline Dual operator*(const Dual& a, const Dual& b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(const Dual& a, const Dual& b) {
    const double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
inline Dual operator+(const Dual& a, double b) { return {a.v + b, a.d}; }
inline Dual operator+(double a, const Dual& b) { return {a + b.v, b.d}; }
inline Dual operator-(const Dual& a, double b) { return {a.v - b, a.d}; }
inline Dual operator-(double a, const Dual& b) { return {a - b.v, -b.d}; }
inline Dual operator*(const Dual& a, double b) { return {a.v * b, a.d * b}; }
inline Dual operator*(double a, const Dual& b) { return {a * b.v, a * b.d}; }
inline Dual operator/(const Dual& a, double b) { return {a.v / b, a.d / b}; }
inline Dual operator/(double a, const Dual& b) {
    const double q = a / b.v;
    return {q, -q * b.d / b.v};
}
inline Dual sqrt_(const Dual& x) {
    const double r = std::sqrt(x.v);
    return {r, x.d / (2.0 * r)};
}
inline Dual abs_(const Dual& x) { return x.v < 0.0 ? -x : x; }

// ---------------------------------------------------------------------------
// Gauss-Lobatto-Legendre basis.  Restates Trixi's LobattoLegendreBasis
// (gauss_lobatto_nodes_weights, barycentric_weights,
// polynomial_derivative_matrix, calc_dhat, calc_lhat), which TrixiLW reads as
// `derivative_matrix`, `derivative_dhat`, `weights`, `inverse_weights` and
// `boundary_interpolation` (src/solvers/dgsem_tree/dg_2d.jl:932, 1211-1226).
// Pinned by closed forms (N=3: ±1, ±sqrt(1/5), w = 1/6, 5/6; N=4: 0,
// ±sqrt(3/7), ±1, w = 32/45, 49/90, 1/10), exact differentiation of
// polynomials of degree <= N, and the SBP identity W D + D^T W = B.
struct Basis {
    int N = 0, n = 0;
    double nodes[MAXN] = {}, weights[MAXN] = {}, wbary[MAXN] = {}, inv_w[MAXN] = {};
    double D[MAXN][MAXN] = {}, Dhat[MAXN][MAXN] = {};
    double bnd_lo = 0, bnd_hi = 0;  // boundary_interpolation[1,1], [n,2]

    void init(int polydeg) {
        N = polydeg;
        n = N + 1;
        // Newton iteration on q = L_{N+1} - L_{N-1} in extended precision
        // (Kopriva, "Implementing Spectral Methods", Alg. 24/25).
        long double x[MAXN], w[MAXN];
        auto qL = [&](long double xx, long double& q, long double& dq, long double& LN) {
            long double Lm2 = 1, Lm1 = xx, dLm2 = 0, dLm1 = 1, Lk = 0, dLk = 0;
            for (int k = 2; k <= N; ++k) {
                Lk = (2.0L * k - 1) / k * xx * Lm1 - (k - 1.0L) / k * Lm2;
                dLk = dLm2 + (2.0L * k - 1) * Lm1;
                Lm2 = Lm1; Lm1 = Lk; dLm2 = dLm1; dLm1 = dLk;
            }
            if (N == 1) { Lk = xx; dLk = 1; Lm2 = 1; dLm2 = 0; }
            const int k = N + 1;
            long double Lk1 = (2.0L * k - 1) / k * xx * Lk - (k - 1.0L) / k * Lm2;
            long double dLk1 = dLm2 + (2.0L * k - 1) * Lk;
            q = Lk1 - Lm2;
            dq = dLk1 - dLm2;
            LN = Lk;
        };
        const long double pi = 3.141592653589793238462643383279502884L;
        x[0] = -1; x[N] = 1;
        w[0] = w[N] = 2.0L / (N * (N + 1.0L));
        for (int j = 1; j <= (N + 1) / 2 - 1; ++j) {
            long double xx = -cosl((j + 0.25L) * pi / N - 3.0L / (8.0L * N * pi * (j + 0.25L)));
            for (int it = 0; it < 100; ++it) {
                long double q, dq, LN;
                qL(xx, q, dq, LN);
                long double d = -q / dq;
                xx += d;
                if (fabsl(d) <= 1e-19L * fabsl(xx) + 1e-30L) break;
            }
            long double q, dq, LN;
            qL(xx, q, dq, LN);
            x[j] = xx;
            x[N - j] = -xx;
            w[j] = w[N - j] = 2.0L / (N * (N + 1.0L) * LN * LN);
        }
        if (N % 2 == 0) {
            long double q, dq, LN;
            qL(0.0L, q, dq, LN);
            x[N / 2] = 0;
            w[N / 2] = 2.0L / (N * (N + 1.0L) * LN * LN);
        }
        for (int i = 0; i < n; ++i) {
            nodes[i] = (double)x[i];
            weights[i] = (double)w[i];
            inv_w[i] = 1.0 / weights[i];
        }
        // barycentric weights, in double as Trixi computes them
        for (int i = 0; i < n; ++i) wbary[i] = 1.0;
        for (int j = 1; j < n; ++j)
            for (int k = 0; k < j; ++k) {
                wbary[k] *= nodes[k] - nodes[j];
                wbary[j] *= nodes[j] - nodes[k];
            }
        for (int j = 0; j < n; ++j) wbary[j] = 1.0 / wbary[j];
        // polynomial_derivative_matrix: negative-sum diagonal
        for (int i = 0; i < n; ++i) {
            D[i][i] = 0;
            for (int j = 0; j < n; ++j)
                if (j != i) {
                    D[i][j] = wbary[j] / wbary[i] * 1.0 / (nodes[i] - nodes[j]);
                    D[i][i] -= D[i][j];
                }
        }
        // calc_dhat: dhat = D', dhat[j,k] *= -w[k]/w[j]
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) Dhat[j][k] = D[k][j] * (-weights[k] / weights[j]);
        // calc_lhat at -1 and +1 (the interpolant is exact at the end nodes)
        bnd_lo = 1.0 / weights[0];
        bnd_hi = 1.0 / weights[n - 1];
        // Trixi vandermonde_legendre: V[i][m] = sqrt(m + 1/2) L_m(x_i); its inverse
        // maps nodal values to modal coefficients (IndicatorHennemannGassner)
        long double V[MAXN][2 * MAXN];
        for (int i = 0; i < n; ++i) {
            long double Lm2 = 1, Lm1 = x[i];
            for (int m = 0; m < n; ++m) {
                long double L;
                if (m == 0) L = 1;
                else if (m == 1) L = x[i];
                else {
                    L = ((2.0L * m - 1) * x[i] * Lm1 - (m - 1.0L) * Lm2) / m;
                    Lm2 = Lm1;
                    Lm1 = L;
                }
                V[i][m] = L * sqrtl(m + 0.5L);
                V[i][n + m] = (i == m) ? 1.0L : 0.0L;
            }
        }
        for (int col = 0; col < n; ++col) {  // Gauss-Jordan with partial pivoting
            int piv = col;
            for (int r = col + 1; r < n; ++r)
                if (fabsl(V[r][col]) > fabsl(V[piv][col])) piv = r;
            for (int q = 0; q < 2 * n; ++q) std::swap(V[col][q], V[piv][q]);
            const long double d = V[col][col];
            for (int q = 0; q < 2 * n; ++q) V[col][q] /= d;
            for (int r = 0; r < n; ++r)
                if (r != col) {
                    const long double f = V[r][col];
                    for (int q = 0; q < 2 * n; ++q) V[r][q] -= f * V[col][q];
                }
        }
        for (int i = 0; i < n; ++i)
            for (int m = 0; m < n; ++m) Vinv[i][m] = (double)V[i][n + m];
        init_mortar(x);
    }
    double Vinv[MAXN][MAXN] = {};

    // Trixi's LobattoLegendreMortarL2 (Trixi.jl, src/solvers/dgsem/l2projection.jl;
    // Trixi is an unpinned dependency of the reference's Project.toml), restated
    // from its definition: forward_lower/upper interpolate the large face's
    // polynomial to the GLL nodes of its lower half (ξ-1)/2 and upper half (ξ+1)/2;
    // reverse_lower/upper are the L2 projection of the two halves back onto the
    // large face, evaluated by (N+1)-point Gauss quadrature (exact for these
    // degree-2N integrands) between interpolations GLL -> Gauss -> GLL.  Here the
    // Lagrange polynomials are plain products in extended precision.
    double fwd_lo[MAXN][MAXN] = {}, fwd_up[MAXN][MAXN] = {}, rev_lo[MAXN][MAXN] = {}, rev_up[MAXN][MAXN] = {};
    static long double lagrange(const long double* xs, int m, int k, long double x) {
        long double p = 1.0L;
        for (int q = 0; q < m; ++q)
            if (q != k) p *= (x - xs[q]) / (xs[k] - xs[q]);
        return p;
    }
    void init_mortar(const long double* x) {
        // Gauss-Legendre nodes and weights: Newton on L_n from Chebyshev-like guesses
        long double g[MAXN], wg[MAXN];
        const long double pi = 3.141592653589793238462643383279502884L;
        for (int j = 0; j < n; ++j) {
            long double xx = -cosl((j + 0.75L) * pi / (n + 0.5L)), dL = 1;
            for (int it = 0; it < 100; ++it) {
                long double L0 = 1, L1 = xx;
                dL = 1;
                for (int k = 2; k <= n; ++k) {
                    const long double L2 = ((2.0L * k - 1) * xx * L1 - (k - 1.0L) * L0) / k;
                    L0 = L1;
                    L1 = L2;
                }
                if (n == 1) { L1 = xx; L0 = 1; }
                dL = n * (xx * L1 - L0) / (xx * xx - 1.0L);
                const long double d = -L1 / dL;
                xx += d;
                if (fabsl(d) <= 1e-19L) break;
            }
            g[j] = xx;
            wg[j] = 2.0L / ((1.0L - xx * xx) * dL * dL);
        }
        long double l2g[MAXN][MAXN], g2l[MAXN][MAXN], op_lo[MAXN][MAXN], op_up[MAXN][MAXN];
        for (int a = 0; a < n; ++a)
            for (int k = 0; k < n; ++k) {
                l2g[a][k] = lagrange(x, n, k, g[a]);  // GLL data -> Gauss node a
                g2l[a][k] = lagrange(g, n, k, x[a]);  // Gauss data -> GLL node a
                // projection on Gauss nodes: (Pf)(g_a) = Σ_k ½ ℓ_a(½(g_k ± 1)) w_k / w_a f(g_k)
                op_lo[a][k] = 0.5L * lagrange(g, n, a, 0.5L * (g[k] - 1.0L)) * wg[k] / wg[a];
                op_up[a][k] = 0.5L * lagrange(g, n, a, 0.5L * (g[k] + 1.0L)) * wg[k] / wg[a];
            }
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) {
                fwd_lo[i][k] = (double)lagrange(x, n, k, 0.5L * (x[i] - 1.0L));
                fwd_up[i][k] = (double)lagrange(x, n, k, 0.5L * (x[i] + 1.0L));
                long double rl = 0, ru = 0;
                for (int a = 0; a < n; ++a)
                    for (int b = 0; b < n; ++b) {
                        rl += g2l[i][a] * op_lo[a][b] * l2g[b][k];
                        ru += g2l[i][a] * op_up[a][b] * l2g[b][k];
                    }
                rev_lo[i][k] = (double)rl;
                rev_up[i][k] = (double)ru;
            }
    }
};

// ---------------------------------------------------------------------------
// Equations.
//
// Trixi CompressibleEulerEquations2D: flux(u, orientation), cons2prim,
// max_abs_speed_naive, max_abs_speeds.  TrixiLW calls them as
// `fluxes(u, eq) = (flux(u,1), flux(u,2))` (src/solvers/dgsem_tree/dg_2d.jl:368)
// and through FluxPlusDissipation (src/equations/numerical_fluxes.jl:14-38).
// Pinned by the closed form, the x<->y mirror symmetry, free-stream
// preservation and the isentropic-vortex convergence order.
template <class T> struct Euler2D {
    static constexpr int NV = 4;
    double gamma = 1.4;
    void fluxes(const T* u, T* f, T* g) const {
        const T rho = u[0], m1 = u[1], m2 = u[2], E = u[3];
        const T v1 = m1 / rho, v2 = m2 / rho;
        const T p = (gamma - 1.0) * (E - 0.5 * (m1 * v1 + m2 * v2));
        f[0] = m1; f[1] = m1 * v1 + p; f[2] = m1 * v2; f[3] = (E + p) * v1;
        g[0] = m2; g[1] = m2 * v1; g[2] = m2 * v2 + p; g[3] = (E + p) * v2;
    }
    void flux(const T* u, int orientation, T* f) const {
        T ff[4], gg[4];
        fluxes(u, ff, gg);
        for (int v = 0; v < 4; ++v) f[v] = orientation == 1 ? ff[v] : gg[v];
    }
    // Trixi flux(u, normal_direction, ::CompressibleEulerEquations2D): with
    // v_n = v·n, (ρ v_n, ρ v_n v1 + p n1, ρ v_n v2 + p n2, (ρe + p) v_n)
    void flux_normal(const T* u, const double* nrm, T* f) const {
        T rho, v1, v2, p;
        cons2prim(u, rho, v1, v2, p);
        const T vn = v1 * nrm[0] + v2 * nrm[1];
        const T rvn = rho * vn;
        f[0] = rvn;
        f[1] = rvn * v1 + p * nrm[0];
        f[2] = rvn * v2 + p * nrm[1];
        f[3] = (u[3] + p) * vn;
    }
    void cons2prim(const T* u, T& rho, T& v1, T& v2, T& p) const {
        rho = u[0];
        v1 = u[1] / rho;
        v2 = u[2] / rho;
        p = (gamma - 1.0) * (u[3] - 0.5 * (u[1] * v1 + u[2] * v2));
    }
    // |v_n| and c of one state, the two ingredients of max_abs_speed_naive
    void speed_parts(const T* u, int orientation, T& vn_abs, T& c) const {
        T rho, v1, v2, p;
        cons2prim(u, rho, v1, v2, p);
        vn_abs = abs_(orientation == 1 ? v1 : v2);
        c = sqrt_(gamma * p / rho);
    }
    T max_abs_speed_naive(const T* ul, const T* ur, int orientation) const {
        T al, cl, ar, cr;
        speed_parts(ul, orientation, al, cl);
        speed_parts(ur, orientation, ar, cr);
        return max_(al, ar) + max_(cl, cr);
    }
    void max_abs_speeds(const T* u, T& l1, T& l2) const {
        T rho, v1, v2, p;
        cons2prim(u, rho, v1, v2, p);
        const T c = sqrt_(gamma * p / rho);
        l1 = abs_(v1) + c;
        l2 = abs_(v2) + c;
    }
    // max_abs_speed_naive with a normal direction (curvilinear meshes):
    // max(|v_ll·n|, |v_rr·n|) + max(c_ll, c_rr) |n|
    T max_abs_speed_naive_normal(const T* ul, const T* ur, const double* nrm) const {
        T rl, v1l, v2l, pl, rr, v1r, v2r, pr;
        cons2prim(ul, rl, v1l, v2l, pl);
        cons2prim(ur, rr, v1r, v2r, pr);
        const T vl = v1l * nrm[0] + v2l * nrm[1];
        const T vr = v1r * nrm[0] + v2r * nrm[1];
        const T cl = sqrt_(gamma * pl / rl), cr = sqrt_(gamma * pr / rr);
        return max_(abs_(vl), abs_(vr)) + max_(cl, cr) * std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1]);
    }
    // Trixi density and pressure, the variables of the positivity limiter
    T density(const T* u) const { return u[0]; }
    T pressure(const T* u) const { return (gamma - 1.0) * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]) / u[0]); }
    // Trixi density_pressure: ρ p = (γ-1)(ρ ρe - ½(ρv1² + ρv2²)), the
    // indicator variable of the shock-capturing configuration
    T density_pressure(const T* u) const {
        return (gamma - 1.0) * (u[0] * u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]));
    }
    // src/equations/compressible_euler_2d.jl:4-7
    bool is_admissible(const T* u) const {
        T rho, v1, v2, p;
        cons2prim(u, rho, v1, v2, p);
        return value(rho) > 0.0 && value(p) > 0.0;
    }
};

// Trixi LinearScalarAdvectionEquation2D (flux a_d * u, constant speed).  Used
// by the linear-case pin: with a linear flux the finite-difference time
// derivatives are exact and one step has a closed matrix form.
template <class T> struct LinAdv2D {
    static constexpr int NV = 1;
    double a1 = 1.0, a2 = 1.0;
    void fluxes(const T* u, T* f, T* g) const { f[0] = a1 * u[0]; g[0] = a2 * u[0]; }
    void flux(const T* u, int orientation, T* f) const { f[0] = (orientation == 1 ? a1 : a2) * u[0]; }
    void flux_normal(const T* u, const double* nrm, T* f) const { f[0] = (a1 * nrm[0] + a2 * nrm[1]) * u[0]; }
    void speed_parts(const T*, int orientation, T& vn_abs, T& c) const {
        vn_abs = T(std::fabs(orientation == 1 ? a1 : a2));
        c = T(0.0);
    }
    T max_abs_speed_naive(const T*, const T*, int orientation) const {
        return T(std::fabs(orientation == 1 ? a1 : a2));
    }
    void max_abs_speeds(const T*, T& l1, T& l2) const { l1 = T(std::fabs(a1)); l2 = T(std::fabs(a2)); }
    T max_abs_speed_naive_normal(const T*, const T*, const double* nrm) const {
        return T(std::fabs(a1 * nrm[0] + a2 * nrm[1]));
    }
    T density_pressure(const T* u) const { return u[0]; }
    T density(const T* u) const { return u[0]; }
    T pressure(const T* u) const { return u[0]; }
    bool is_admissible(const T*) const { return true; }
};

// ---------------------------------------------------------------------------
// Boundary value functions for the time-averaged Dirichlet condition
// (src/equations/equations.jl:6-37).  Same set as the HIP library.
// BV_SLIP_WALL is no boundary value but the reference's reflecting wall
// (boundary_condition_slip_wall_vertical on x faces, _horizontal on y faces,
// src/equations/equations.jl:95-139).
// BV_CALLBACK: a caller-defined boundary_value_function (x, y, t) -> conservative
// state, Trixi's BoundaryConditionDirichlet(f) for any f (equations.jl:6-37)
// BV_CONVERGENCE_TEST: Trixi's initial_condition_convergence_test for
// CompressibleEulerEquations2D as a boundary value function (the reference's
// elixir_euler_source_terms_nonperiodic.jl): ρ = ρv1 = ρv2 = φ, ρe = φ² with
// φ = 2 + 0.1 sin(π(x + y - t)).
// BV_OUTFLOW, BV_SLIP_WALL_APPROX: boundary_condition_outflow and
// slip_wall_approximate (equations.jl:185-197, 251-260), conditions on uⁿ with
// a normal direction (curvilinear mesh only; the solver applies them, eval
// gives no value).
enum BoundaryValueKind {
    BV_NONE = 0, BV_VORTEX = 1, BV_CONSTANT = 2, BV_RIEMANN4 = 3, BV_SLIP_WALL = 4, BV_CALLBACK = 5,
    BV_CONVERGENCE_TEST = 6, BV_OUTFLOW = 7, BV_SLIP_WALL_APPROX = 8
};

// The manufactured solution of Trixi's convergence test for the Euler
// equations (initial_condition_convergence_test, c = 2, A = 0.1, L = 2, ω = π).
inline void convergence_test_state(double x, double y, double t, double* u) {
    const double phi = 2.0 + 0.1 * std::sin(3.141592653589793 * (x + y - t));
    u[0] = u[1] = u[2] = phi;
    u[3] = phi * phi;
}

// The manufactured solution of the Navier-Stokes convergence test
// (examples/tree_2d_dgsem/elixir_navierstokes_convergence.jl:31-45):
// ρ = 2 + ½ sin πx cos πy cos πt, v1 = v2 = sin πx g(y) cos πt with
// g(y) = log(y + 2)(1 - e^{-(y-1)/2}), p = ρ².
inline void ns_convergence_state(double x, double y, double t, double gamma, double* u) {
    const double pi = 3.141592653589793;
    const double rho = 2.0 + 0.5 * std::sin(pi * x) * std::cos(pi * y) * std::cos(pi * t);
    const double v = std::sin(pi * x) * std::log(y + 2.0) * (1.0 - std::exp(-0.5 * (y - 1.0))) * std::cos(pi * t);
    const double p = rho * rho;
    u[0] = rho;
    u[1] = rho * v;
    u[2] = rho * v;
    u[3] = p / (gamma - 1.0) + 0.5 * rho * (v * v + v * v);
}

// Its source (the example's source_terms_navier_stokes_convergence_test,
// elixir_navierstokes_convergence.jl:49-205): the residual of the manufactured
// solution in the compressible Navier-Stokes equations with Trixi's viscous
// fluxes (compressible_navier_stokes_2d.jl: τ with Stokes' hypothesis, heat
// flux μ κ ∇T with κ = γ/((γ-1) Pr) and T = p/ρ),
//   s = ∂t u + ∇·f_a(u) - ∇·f_v(u, ∇u),
// assembled here from the closed-form first and second derivatives of ρ, v1,
// v2 and p by the product rule, term by term of each conservation law.  The
// source reads the example's μ and Pr, which are the equations' own.
inline void ns_convergence_source(double x, double y, double t, double gamma, double mu, double prandtl,
                                  double* s) {
    const double pi = 3.141592653589793, A = 0.5;
    const double sx = std::sin(pi * x), cx = std::cos(pi * x);
    const double sy = std::sin(pi * y), cy = std::cos(pi * y);
    const double st = std::sin(pi * t), ct = std::cos(pi * t);
    // density and its derivatives
    const double r = 2.0 + A * sx * cy * ct;
    const double r_t = -pi * A * sx * cy * st;
    const double r_x = pi * A * cx * cy * ct;
    const double r_y = -pi * A * sx * sy * ct;
    const double r_xx = -pi * pi * A * sx * cy * ct;
    const double r_yy = r_xx;
    // velocity φ = sin πx g(y) cos πt for both components
    const double e = std::exp(-A * (y - 1.0)), L = std::log(y + 2.0), q = 1.0 / (y + 2.0);
    const double g = L * (1.0 - e);
    const double g1 = (1.0 - e) * q + A * L * e;
    const double g2 = -(1.0 - e) * q * q + 2.0 * A * e * q - A * A * L * e;
    const double v = sx * g * ct, v_t = -pi * sx * g * st;
    const double v_x = pi * cx * g * ct, v_y = sx * g1 * ct;
    const double v_xx = -pi * pi * sx * g * ct, v_xy = pi * cx * g1 * ct, v_yy = sx * g2 * ct;
    const double v1 = v, v2 = v;
    const double v1_t = v_t, v2_t = v_t, v1_x = v_x, v2_x = v_x, v1_y = v_y, v2_y = v_y;
    const double v1_xx = v_xx, v2_xx = v_xx, v1_xy = v_xy, v2_xy = v_xy, v1_yy = v_yy, v2_yy = v_yy;
    // pressure p = ρ²
    const double p = r * r, p_x = 2.0 * r * r_x, p_y = 2.0 * r * r_y;
    const double p_xx = 2.0 * (r_x * r_x + r * r_xx), p_yy = 2.0 * (r_y * r_y + r * r_yy);
    // total energy E = p/(γ-1) + ρ|v|²/2
    const double ig = 1.0 / (gamma - 1.0), k2 = v1 * v1 + v2 * v2;
    const double E = p * ig + 0.5 * r * k2;
    const double E_t = 2.0 * r * r_t * ig + 0.5 * r_t * k2 + r * (v1 * v1_t + v2 * v2_t);
    const double E_x = p_x * ig + 0.5 * r_x * k2 + r * (v1 * v1_x + v2 * v2_x);
    const double E_y = p_y * ig + 0.5 * r_y * k2 + r * (v1 * v1_y + v2 * v2_y);
    // viscous stresses (without μ) and their divergence pieces
    const double t11 = (4.0 / 3.0) * v1_x - (2.0 / 3.0) * v2_y, t12 = v1_y + v2_x;
    const double t22 = (4.0 / 3.0) * v2_y - (2.0 / 3.0) * v1_x;
    const double t11_x = (4.0 / 3.0) * v1_xx - (2.0 / 3.0) * v2_xy, t12_x = v1_xy + v2_xx;
    const double t12_y = v1_yy + v2_xy, t22_y = (4.0 / 3.0) * v2_yy - (2.0 / 3.0) * v1_xy;
    // ∇²T of T = p/ρ: T_xx = (p_xx ρ² - 2 p_x ρ ρ_x + 2 p ρ_x² - p ρ ρ_xx)/ρ³, same in y
    const double ir3 = 1.0 / (r * r * r);
    const double T_xx = (p_xx * r * r - 2.0 * p_x * r * r_x + 2.0 * p * r_x * r_x - p * r * r_xx) * ir3;
    const double T_yy = (p_yy * r * r - 2.0 * p_y * r * r_y + 2.0 * p * r_y * r_y - p * r * r_yy) * ir3;
    const double kappa = gamma * ig / prandtl;
    // mass: ρ_t + (ρ v1)_x + (ρ v2)_y
    s[0] = r_t + r_x * v1 + r * v1_x + r_y * v2 + r * v2_y;
    // x momentum: (ρ v1)_t + (ρ v1² + p)_x + (ρ v1 v2)_y - μ (τ11_x + τ12_y)
    s[1] = r_t * v1 + r * v1_t + r_x * v1 * v1 + 2.0 * r * v1 * v1_x + p_x + r_y * v1 * v2 + r * v1_y * v2 +
           r * v1 * v2_y - mu * (t11_x + t12_y);
    // y momentum: (ρ v2)_t + (ρ v1 v2)_x + (ρ v2² + p)_y - μ (τ12_x + τ22_y)
    s[2] = r_t * v2 + r * v2_t + r_x * v1 * v2 + r * v1_x * v2 + r * v1 * v2_x + r_y * v2 * v2 +
           2.0 * r * v2 * v2_y + p_y - mu * (t12_x + t22_y);
    // energy: E_t + ((E + p) v1)_x + ((E + p) v2)_y
    //         - μ ((v1 τ11 + v2 τ12)_x + (v1 τ12 + v2 τ22)_y + κ (T_xx + T_yy))
    const double work_x = v1_x * t11 + v1 * t11_x + v2_x * t12 + v2 * t12_x;
    const double work_y = v1_y * t12 + v1 * t12_y + v2_y * t22 + v2 * t22_y;
    s[3] = E_t + (E_x + p_x) * v1 + (E + p) * v1_x + (E_y + p_y) * v2 + (E + p) * v2_y -
           mu * (work_x + work_y + kappa * (T_xx + T_yy));
}

// Source terms s(u, x, t) (the `source_terms` of SemidiscretizationHyperbolic),
// which the LW kernels fold into the Taylor procedure (dg_2d.jl:89-176).
// SRC_CONVERGENCE_TEST: Trixi's source_terms_convergence_test for
// CompressibleEulerEquations2D, the residual ∂t u + ∇·f(u) of the state above
// (with φ' = ∂x φ = π 0.1 cos(π(x + y - t)) and p = (γ-1)(φ² - φ)):
// (φ', φ'(1 + τ), φ'(1 + τ), 2φ'(φ + τ)), τ = (2φ - 1)(γ - 1); it does not
// depend on u.  SRC_CALLBACK: a caller's function of (u, x, y, t).
// SRC_NS_CONVERGENCE: the Navier-Stokes convergence source above (with the
// hyperbolic-parabolic solver; it does not depend on u either).
enum SourceKind { SRC_NONE = 0, SRC_CONVERGENCE_TEST = 1, SRC_CALLBACK = 2, SRC_NS_CONVERGENCE = 3 };

struct SourceTerm {
    int kind = SRC_NONE;
    double gamma = 1.4;
    double mu = 0.0, prandtl = 1.0;  // SRC_NS_CONVERGENCE
    void (*fn)(const double*, double, double, double, double*) = nullptr;
    // NV = 4 (Euler); other equations take no source terms
    template <int NV, class T> void eval(const T* u, double x, double y, double t, T* s) const {
        if constexpr (NV != 4) {
            for (int v = 0; v < NV; ++v) s[v] = T(0);
            return;
        } else if (kind == SRC_CONVERGENCE_TEST) {
            const double w = 3.141592653589793;
            const double arg = w * (x + y - t);
            const double rho = 2.0 + 0.1 * std::sin(arg);
            const double rho_x = w * 0.1 * std::cos(arg);
            const double tmp = (2.0 * rho - 1.0) * (gamma - 1.0);
            s[0] = T(rho_x);
            s[1] = T(rho_x * (1.0 + tmp));
            s[2] = s[1];
            s[3] = T(2.0 * rho_x * (rho + tmp));
        } else if (kind == SRC_NS_CONVERGENCE) {
            double sv[4];
            ns_convergence_source(x, y, t, gamma, mu, prandtl, sv);
            for (int v = 0; v < 4; ++v) s[v] = T(sv[v]);
        } else if (kind == SRC_CALLBACK) {
            double uv[4], sv[4];
            for (int v = 0; v < 4; ++v) uv[v] = value(u[v]);
            fn(uv, x, y, t, sv);
            for (int v = 0; v < 4; ++v) s[v] = T(sv[v]);
        } else {
            for (int v = 0; v < 4; ++v) s[v] = T(0);
        }
    }
};

struct BoundaryValue {
    int kind = BV_NONE;
    double par[24] = {};
    // isentropic vortex: par = {eps, mach, xc0, yc0, Lx, Ly, gamma}
    // constant: par = {rho, v1, v2, p, gamma}
    // four-quadrant Riemann: par = {x0, y0, gamma, then (rho, v1, v2, p) for
    //   quadrants NE, NW, SW, SE}
    void (*fn)(double, double, double, double*) = nullptr;  // BV_CALLBACK
    void eval(double x, double y, double t, double* u) const {
        double rho, v1, v2, p, gamma;
        if (kind == BV_CALLBACK) {
            fn(x, y, t, u);
            return;
        }
        if (kind == BV_CONVERGENCE_TEST) {
            convergence_test_state(x, y, t, u);
            return;
        }
        if (kind == BV_VORTEX) {
            const double eps = par[0], mach = par[1], Lx = par[4], Ly = par[5];
            gamma = par[6];
            const double vinf = mach * std::sqrt(gamma) / std::sqrt(2.0);
            double dx = x - (par[2] + vinf * t), dy = y - (par[3] + vinf * t);
            dx -= Lx * std::floor(dx / Lx + 0.5);
            dy -= Ly * std::floor(dy / Ly + 0.5);
            const double r2 = dx * dx + dy * dy;
            const double pi = 3.141592653589793;
            const double du = eps / (2 * pi) * std::exp(0.5 * (1 - r2));
            const double T = 1 - (gamma - 1) * eps * eps / (8 * gamma * pi * pi) * std::exp(1 - r2);
            rho = std::pow(T, 1 / (gamma - 1));
            v1 = vinf - du * dy;
            v2 = vinf + du * dx;
            p = rho * T;
        } else if (kind == BV_CONSTANT) {
            rho = par[0]; v1 = par[1]; v2 = par[2]; p = par[3]; gamma = par[4];
        } else if (kind == BV_RIEMANN4) {
            gamma = par[2];
            int q;
            if (x >= par[0]) q = (y >= par[1]) ? 0 : 3;
            else q = (y >= par[1]) ? 1 : 2;
            rho = par[3 + 4 * q]; v1 = par[4 + 4 * q]; v2 = par[5 + 4 * q]; p = par[6 + 4 * q];
        } else {
            u[0] = u[1] = u[2] = u[3] = std::numeric_limits<double>::quiet_NaN();
            return;
        }
        u[0] = rho;
        u[1] = rho * v1;
        u[2] = rho * v2;
        u[3] = p / (gamma - 1) + 0.5 * rho * (v1 * v1 + v2 * v2);
    }
};

// ---------------------------------------------------------------------------
// Uniform 2-D TreeMesh (square elements), elements numbered e = ix + nx*iy.
// Trixi numbers TreeMesh leaves in depth-first (Z) order; every routine here
// uses the neighbour and interface tables, so the element order is a free
// choice and parity is compared by node coordinate.
struct Mesh {
    int nx = 0, ny = 0, E = 0;
    double xmin = 0, ymin = 0, h = 0;  // element edge length
    bool per_x = true, per_y = true;
    std::vector<double> inv_jac;       // 2/h per element (Trixi TreeMesh)
    std::vector<int> nbr;              // [e*4 + face], -1 on a domain boundary
    // Trixi interface container: neighbor_ids[1|2], orientations
    std::vector<int> if_left, if_right, if_orient;
    // Trixi boundary container, sorted by direction 1..4
    std::vector<int> bd_elem, bd_orient, bd_side, bd_dir;

    void init(int nx_, int ny_, double xmin_, double ymin_, double h_, bool px, bool py) {
        nx = nx_; ny = ny_; E = nx * ny; xmin = xmin_; ymin = ymin_; h = h_; per_x = px; per_y = py;
        inv_jac.assign(E, 2.0 / h);
        nbr.assign(4 * E, -1);
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const int e = ix + nx * iy;
                int l = ix - 1, r = ix + 1, b = iy - 1, t = iy + 1;
                if (per_x) { l = (l + nx) % nx; r = r % nx; }
                if (per_y) { b = (b + ny) % ny; t = t % ny; }
                nbr[4 * e + 0] = (l >= 0 && l < nx) ? l + nx * iy : -1;
                nbr[4 * e + 1] = (r >= 0 && r < nx) ? r + nx * iy : -1;
                nbr[4 * e + 2] = (b >= 0 && b < ny) ? ix + nx * b : -1;
                nbr[4 * e + 3] = (t >= 0 && t < ny) ? ix + nx * t : -1;
            }
        if_left.clear(); if_right.clear(); if_orient.clear();
        for (int e = 0; e < E; ++e) {
            if (nbr[4 * e + 1] >= 0) { if_left.push_back(e); if_right.push_back(nbr[4 * e + 1]); if_orient.push_back(1); }
            if (nbr[4 * e + 3] >= 0) { if_left.push_back(e); if_right.push_back(nbr[4 * e + 3]); if_orient.push_back(2); }
        }
        bd_elem.clear(); bd_orient.clear(); bd_side.clear(); bd_dir.clear();
        for (int dir = 1; dir <= 4; ++dir)
            for (int e = 0; e < E; ++e)
                if (nbr[4 * e + (dir - 1)] < 0) {
                    bd_elem.push_back(e);
                    bd_orient.push_back(dir <= 2 ? 1 : 2);
                    // neighbor_sides: 1 = element left of the boundary (+ face)
                    bd_side.push_back((dir % 2 == 0) ? 1 : 2);
                    bd_dir.push_back(dir);
                }
    }
    // A TreeMesh with refined cells (Trixi TreeMesh leaves of different levels):
    // element centres and edge lengths, the face neighbours (-1 domain boundary,
    // -2 a face on a mortar) and Trixi's L2 mortar container (neighbor_ids
    // lower, upper, large; orientations; large_sides), one mortar per large face
    // with two small neighbours.
    struct MortarRec { int lower, upper, large, orient, large_side; };
    std::vector<double> cx, cy, len;
    std::vector<MortarRec> mortars;
    void init_general(int E_, const double* cx_, const double* cy_, const double* len_, const int* nbr_, int M,
                      const int* mort) {
        E = E_; nx = E_; ny = 1; h = 0.0; per_x = per_y = false;
        cx.assign(cx_, cx_ + E);
        cy.assign(cy_, cy_ + E);
        len.assign(len_, len_ + E);
        inv_jac.resize(E);
        for (int e = 0; e < E; ++e) inv_jac[e] = 2.0 / len[e];
        nbr.assign(nbr_, nbr_ + 4 * E);
        mortars.resize(M);
        for (int m = 0; m < M; ++m)
            mortars[m] = {mort[5 * m], mort[5 * m + 1], mort[5 * m + 2], mort[5 * m + 3], mort[5 * m + 4]};
        if_left.clear(); if_right.clear(); if_orient.clear();
        for (int e = 0; e < E; ++e) {
            if (nbr[4 * e + 1] >= 0) { if_left.push_back(e); if_right.push_back(nbr[4 * e + 1]); if_orient.push_back(1); }
            if (nbr[4 * e + 3] >= 0) { if_left.push_back(e); if_right.push_back(nbr[4 * e + 3]); if_orient.push_back(2); }
        }
        bd_elem.clear(); bd_orient.clear(); bd_side.clear(); bd_dir.clear();
        for (int dir = 1; dir <= 4; ++dir)
            for (int e = 0; e < E; ++e)
                if (nbr[4 * e + (dir - 1)] == -1) {
                    bd_elem.push_back(e);
                    bd_orient.push_back(dir <= 2 ? 1 : 2);
                    bd_side.push_back((dir % 2 == 0) ? 1 : 2);
                    bd_dir.push_back(dir);
                }
    }
    double node_x(const Basis& b, int i, int e) const {
        if (!cx.empty()) return cx[e] + 0.5 * len[e] * b.nodes[i];
        return xmin + h * ((e % nx) + 0.5) + 0.5 * h * b.nodes[i];
    }
    double node_y(const Basis& b, int j, int e) const {
        if (!cy.empty()) return cy[e] + 0.5 * len[e] * b.nodes[j];
        return ymin + h * ((e / nx) + 0.5) + 0.5 * h * b.nodes[j];
    }
};

// ---------------------------------------------------------------------------
// Solver state (the subset of `semi.cache` on the LW path,
// src/solvers/dgsem_tree/containers.jl:7-165).
template <class T, class Eq> struct Solver {
    static constexpr int NV = Eq::NV;
    Basis basis;
    Mesh mesh;
    Eq eq;
    BoundaryValue bval;
    int n = 0;
    // 0: TreeMesh order (λ from Ũ, jump on uⁿ; dg_2d.jl:1928)
    // 1: Structured/P4est order (λ from uⁿ, jump on Ũ; structured/dg_2d.jl:2271)
    // 2: the StructuredMesh conventions on this Cartesian mesh: Structured order
    //    at interfaces, TreeMesh order at Dirichlet boundaries
    //    (structured/dg_2d.jl:2374) and a full-step boundary time average in MDRK
    //    stage 1 (structured/multi_derivative_2d.jl:18-20 passes no scaling)
    int flux_order = 0;
    double abstol = 0.0, reltol = 0.0;
    // shock capturing (VolumeIntegralFRShockCapturing with IndicatorHennemannGassner
    // on density_pressure and first-order subcell FV, src/solvers/fr.jl:63-86)
    bool sc = false;
    bool muscl_hancock = false;  // MUSCLHancockReconstruction instead of FirstOrderReconstruction
    double alpha_max = 0.5, alpha_min = 0.001;
    bool alpha_smooth = true;
    std::vector<double> alpha, alpha_tmp, alpha_override;
    double sc_dt = 0.0;  // dt passed to compute_alp
    // PositivityPreservingLimiterZhangShu(thresholds, variables = (density, pressure))
    // passed to solve_lwfr as a stage limiter (lwfr.jl:131-137, 226, 251, 257)
    bool positivity = false;
    double pos_threshold[2] = {5.0e-6, 5.0e-6};
    std::vector<T> fn_low;  // [v, l, face, e]
    std::vector<T> U, F, sfv, temporal_errors;
    std::vector<T> us, U2, F2;  // MDRK stage cache (containers.jl:83-98)
    // source terms and the t of the current rhs call; S2 is the MDRK stage
    // cache's source part (s + s_t/6 of stage 1, read by stage 2)
    SourceTerm src;
    double src_t = 0.0;
    std::vector<T> S2;

    virtual ~Solver() = default;
    // node coordinates (TreeMesh: from the element's centre and edge length)
    virtual void node_xy(int i, int j, int e, double& x, double& y) const {
        x = mesh.node_x(basis, i, e);
        y = mesh.node_y(basis, j, e);
    }
    // gradients and viscous caches of the hyperbolic-parabolic solver (tlwo_get_aux 9-12)
    virtual const std::vector<T>* extra_aux(int) const { return nullptr; }

    void setup() {
        n = basis.n;
        const size_t nd = (size_t)mesh.E * n * n * NV;
        U.assign(nd, T(0));
        F.assign(2 * nd, T(0));
        sfv.assign((size_t)mesh.E * 4 * n * NV, T(0));
        temporal_errors.assign(mesh.E, T(0));
        alpha.assign(mesh.E, 0.0);
        alpha_tmp.assign(mesh.E, 0.0);
        fn_low.assign((size_t)mesh.E * 4 * n * NV, T(0));
    }
    // the MDRK stage cache is allocated on first use
    void ensure_mdrk() {
        const size_t nd = ndofs_total();
        if (us.size() != nd) us.assign(nd, T(0));
        if (U2.size() != nd) U2.assign(nd, T(0));
        if (F2.size() != 2 * nd) F2.assign(2 * nd, T(0));
        if (S2.size() != nd) S2.assign(nd, T(0));
    }

    // The source-term finite differences in time (dg_2d.jl:89-176) at the node
    // (x, y): s(w, Δt) = source_terms(w, x, t + Δt), each stencil state at the
    // time level it approximates.  The argument order of each helper is the
    // reference's (u, up, upp, um, umm), so that its call sites below can pass
    // the states exactly as the reference's kernels do.
    void src_at(const T* w, double x, double y, double dtoff, T* s) const { src.template eval<NV>(w, x, y, src_t + dtoff, s); }
    void src_t_N34(const T* up, const T* upp, const T* um, const T* umm, double x, double y, double dt, double t0,
                   T* st) const {
        T a[NV], b[NV], c[NV], d[NV];
        src.template eval<NV>(upp, x, y, t0 + 2.0 * dt, a);
        src.template eval<NV>(up, x, y, t0 + dt, b);
        src.template eval<NV>(um, x, y, t0 - dt, c);
        src.template eval<NV>(umm, x, y, t0 - 2.0 * dt, d);
        for (int v = 0; v < NV; ++v) st[v] = (1.0 / 12.0) * (-a[v] + 8.0 * b[v] - 8.0 * c[v] + d[v]);
    }
    void src_tt_N23(const T* u, const T* up, const T* um, double x, double y, double dt, T* s) const {
        T a[NV], b[NV], c[NV];
        src.template eval<NV>(up, x, y, src_t + dt, a);
        src.template eval<NV>(u, x, y, src_t, b);
        src.template eval<NV>(um, x, y, src_t - dt, c);
        for (int v = 0; v < NV; ++v) s[v] = a[v] - 2.0 * b[v] + c[v];
    }
    void src_ttt_N34(const T* up, const T* upp, const T* um, const T* umm, double x, double y, double dt,
                     T* s) const {
        T a[NV], b[NV], c[NV], d[NV];
        src.template eval<NV>(upp, x, y, src_t + 2.0 * dt, a);
        src.template eval<NV>(up, x, y, src_t + dt, b);
        src.template eval<NV>(um, x, y, src_t - dt, c);
        src.template eval<NV>(umm, x, y, src_t - 2.0 * dt, d);
        for (int v = 0; v < NV; ++v) s[v] = 0.5 * (a[v] - 2.0 * b[v] + 2.0 * c[v] - d[v]);
    }
    void src_tttt_N4(const T* u, const T* up, const T* upp, const T* um, const T* umm, double x, double y, double dt,
                     T* s) const {
        T a[NV], b[NV], c[NV], d[NV], z[NV];
        src.template eval<NV>(upp, x, y, src_t + 2.0 * dt, a);
        src.template eval<NV>(up, x, y, src_t + dt, b);
        src.template eval<NV>(u, x, y, src_t, z);
        src.template eval<NV>(um, x, y, src_t - dt, c);
        src.template eval<NV>(umm, x, y, src_t - 2.0 * dt, d);
        for (int v = 0; v < NV; ++v) s[v] = a[v] - 4.0 * b[v] + 6.0 * z[v] - 4.0 * c[v] + d[v];
    }
    using NodeArr = T[MAXN][MAXN][NV];
    // S = s(uⁿ, t), ut += dt s (dg_2d.jl:980-987: the source has no Jacobian factor)
    void src_init(const T* u, int e, double dt, NodeArr& ut, NodeArr& Sc) const {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double x, y;
                node_xy(i, j, e, x, y);
                T s[NV];
                src_at(&u[idx(0, i, j, e)], x, y, 0.0, s);
                for (int v = 0; v < NV; ++v) {
                    Sc[j][i][v] = s[v];
                    ut[j][i][v] = ut[j][i][v] + dt * s[v];
                }
            }
    }
    // s_t of the stage-1 states: S += s_t/2, utt += dt s_t (1033-1045)
    void src_stage_t(int e, double dt, const NodeArr& up, const NodeArr& upp, const NodeArr& um, const NodeArr& umm,
                     NodeArr& utt, NodeArr& Sc) const {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double x, y;
                node_xy(i, j, e, x, y);
                T s[NV];
                src_t_N34(up[j][i], upp[j][i], um[j][i], umm[j][i], x, y, dt, src_t, s);
                for (int v = 0; v < NV; ++v) {
                    Sc[j][i][v] = Sc[j][i][v] + 0.5 * s[v];
                    utt[j][i][v] = utt[j][i][v] + dt * s[v];
                }
            }
    }
    // s_tt (three-point) of the stage-2 states: S += s_tt/6, uttt += dt s_tt (1093-1103)
    void src_stage_tt(const T* u, int e, double dt, const NodeArr& up, const NodeArr& um, NodeArr& uttt,
                      NodeArr& Sc) const {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double x, y;
                node_xy(i, j, e, x, y);
                T s[NV];
                src_tt_N23(&u[idx(0, i, j, e)], up[j][i], um[j][i], x, y, dt, s);
                for (int v = 0; v < NV; ++v) {
                    Sc[j][i][v] = Sc[j][i][v] + (1.0 / 6.0) * s[v];
                    uttt[j][i][v] = uttt[j][i][v] + dt * s[v];
                }
            }
    }
    // s_ttt of the stage-3 states: S += s_ttt/24, utttt += dt s_ttt (1148-1160)
    void src_stage_ttt(int e, double dt, const NodeArr& up, const NodeArr& upp, const NodeArr& um,
                       const NodeArr& umm, NodeArr& utttt, NodeArr& Sc) const {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double x, y;
                node_xy(i, j, e, x, y);
                T s[NV];
                src_ttt_N34(up[j][i], upp[j][i], um[j][i], umm[j][i], x, y, dt, s);
                for (int v = 0; v < NV; ++v) {
                    Sc[j][i][v] = Sc[j][i][v] + (1.0 / 24.0) * s[v];
                    utttt[j][i][v] = utttt[j][i][v] + dt * s[v];
                }
            }
    }
    size_t ndofs_total() const { return (size_t)mesh.E * n * n * NV; }
    size_t idx(int v, int i, int j, int e) const { return (((size_t)e * n + j) * n + i) * NV + v; }
    size_t fidx(int v, int d, int i, int j, int e) const { return ((((size_t)e * n + j) * n + i) * 2 + d) * NV + v; }
    size_t sidx(int v, int l, int face, int e) const { return (((size_t)e * 4 + face) * n + l) * NV + v; }

    // ---------------- LW volume kernel, N = 4 ----------------
    // lw_volume_kernel_4!, src/solvers/dgsem_tree/dg_2d.jl:918-1267.
    // With source terms (src.kind != SRC_NONE) each time derivative gets dt times
    // the source's own derivative at the stencil states of that point of the
    // procedure (980-987, 1033-1045, 1093-1103, 1148-1160) and S = s + s_t/2 +
    // s_tt/6 + s_ttt/24 + s_tttt/120 enters du as -S/J⁻¹ (1229-1247), so that the
    // Jacobian turns it into +S.  s_tt takes the three-point difference
    // (calc_source_tt_N23) and s_tttt the reference's argument order
    // (u, up, um, upp, umm), which swaps the states of t + 2dt and t - dt; a
    // source that does not depend on u, as the convergence test's, is not
    // affected.  Without source terms these steps are skipped (they add zero).
    void lw_volume_kernel_4(T* du, const T* u, double dt, int e, double alpha) {
        const auto& D = basis.D;
        const auto& Dh = basis.Dhat;
        const double ij = mesh.inv_jac[e];
        const bool has_src = src.kind != SRC_NONE;
        T f[MAXN][MAXN][NV], g[MAXN][MAXN][NV], Fc[MAXN][MAXN][NV], Gc[MAXN][MAXN][NV];
        T ut[MAXN][MAXN][NV], utt[MAXN][MAXN][NV], uttt[MAXN][MAXN][NV], utttt[MAXN][MAXN][NV];
        T Uc[MAXN][MAXN][NV], up[MAXN][MAXN][NV], um[MAXN][MAXN][NV], upp[MAXN][MAXN][NV], umm[MAXN][MAXN][NV];
        T unp1[MAXN][MAXN][NV], unp1low[MAXN][MAXN][NV], dlow[MAXN][MAXN][NV], Sc[MAXN][MAXN][NV];
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) ut[j][i][v] = utt[j][i][v] = uttt[j][i][v] = utttt[j][i][v] = dlow[j][i][v] = T(0);
        // 953-987: fluxes at uⁿ and ut = -dt*J⁻¹*(D f + D g)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const T* un = &u[idx(0, i, j, e)];
                T fl[NV], gl[NV];
                eq.fluxes(un, fl, gl);
                for (int v = 0; v < NV; ++v) { f[j][i][v] = fl[v]; g[j][i][v] = gl[v]; Fc[j][i][v] = fl[v]; }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) ut[j][ii][v] = ut[j][ii][v] + (-dt * D[ii][i]) * fl[v];
                for (int v = 0; v < NV; ++v) Gc[j][i][v] = gl[v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) ut[jj][i][v] = ut[jj][i][v] + (-dt * D[jj][j]) * gl[v];
                for (int v = 0; v < NV; ++v)
                    unp1[j][i][v] = unp1low[j][i][v] = um[j][i][v] = up[j][i][v] = umm[j][i][v] = upp[j][i][v] =
                        Uc[j][i][v] = un[v];
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) ut[j][i][v] = ut[j][i][v] * ij;
        if (has_src) src_init(u, e, dt, ut, Sc);
        // 995-1022: first Taylor states, f_t, utt
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                for (int v = 0; v < NV; ++v) {
                    const T a = ut[j][i][v];
                    Uc[j][i][v] = Uc[j][i][v] + 0.5 * a;
                    up[j][i][v] = up[j][i][v] + 1.0 * a;
                    um[j][i][v] = um[j][i][v] + (-1.0) * a;
                    umm[j][i][v] = umm[j][i][v] + (-2.0) * a;
                    upp[j][i][v] = upp[j][i][v] + 2.0 * a;
                }
                T fm[NV], gm[NV], fp[NV], gp[NV], fmm[NV], gmm[NV], fpp[NV], gpp[NV];
                eq.fluxes(um[j][i], fm, gm);
                eq.fluxes(up[j][i], fp, gp);
                eq.fluxes(umm[j][i], fmm, gmm);
                eq.fluxes(upp[j][i], fpp, gpp);
                T ft[NV], gt[NV];
                for (int v = 0; v < NV; ++v) {
                    ft[v] = 1.0 / 12.0 * (-fpp[v] + 8.0 * fp[v] - 8.0 * fm[v] + fmm[v]);
                    gt[v] = 1.0 / 12.0 * (-gpp[v] + 8.0 * gp[v] - 8.0 * gm[v] + gmm[v]);
                    Fc[j][i][v] = Fc[j][i][v] + 0.5 * ft[v];
                    Gc[j][i][v] = Gc[j][i][v] + 0.5 * gt[v];
                }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) utt[j][ii][v] = utt[j][ii][v] + (-dt * D[ii][i]) * ft[v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) utt[jj][i][v] = utt[jj][i][v] + (-dt * D[jj][j]) * gt[v];
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) utt[j][i][v] = utt[j][i][v] * ij;
        if (has_src) src_stage_t(e, dt, up, upp, um, umm, utt, Sc);
        // 1047-1080: f_tt, uttt
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                for (int v = 0; v < NV; ++v) {
                    const T a = utt[j][i][v];
                    Uc[j][i][v] = Uc[j][i][v] + (1.0 / 6.0) * a;
                    up[j][i][v] = up[j][i][v] + 0.5 * a;
                    um[j][i][v] = um[j][i][v] + 0.5 * a;
                    upp[j][i][v] = upp[j][i][v] + 2.0 * a;
                    umm[j][i][v] = umm[j][i][v] + 2.0 * a;
                }
                T fm[NV], gm[NV], fp[NV], gp[NV], fmm[NV], gmm[NV], fpp[NV], gpp[NV];
                eq.fluxes(um[j][i], fm, gm);
                eq.fluxes(up[j][i], fp, gp);
                eq.fluxes(umm[j][i], fmm, gmm);
                eq.fluxes(upp[j][i], fpp, gpp);
                T ftt[NV], gtt[NV];
                for (int v = 0; v < NV; ++v) {
                    ftt[v] = (1.0 / 12.0) * (-fpp[v] + 16.0 * fp[v] - 30.0 * f[j][i][v] + 16.0 * fm[v] - fmm[v]);
                    gtt[v] = (1.0 / 12.0) * (-gpp[v] + 16.0 * gp[v] - 30.0 * g[j][i][v] + 16.0 * gm[v] - gmm[v]);
                    Fc[j][i][v] = Fc[j][i][v] + (1.0 / 6.0) * ftt[v];
                    Gc[j][i][v] = Gc[j][i][v] + (1.0 / 6.0) * gtt[v];
                }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) uttt[j][ii][v] = uttt[j][ii][v] + (-dt * D[ii][i]) * ftt[v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) uttt[jj][i][v] = uttt[jj][i][v] + (-dt * D[jj][j]) * gtt[v];
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) uttt[j][i][v] = uttt[j][i][v] * ij;
        if (has_src) src_stage_tt(u, e, dt, up, um, uttt, Sc);
        // 1103-1136: f_ttt, utttt
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                for (int v = 0; v < NV; ++v) {
                    const T a = uttt[j][i][v];
                    Uc[j][i][v] = Uc[j][i][v] + (1.0 / 24.0) * a;
                    um[j][i][v] = um[j][i][v] + (-1.0 / 6.0) * a;
                    up[j][i][v] = up[j][i][v] + (1.0 / 6.0) * a;
                    umm[j][i][v] = umm[j][i][v] + (-4.0 / 3.0) * a;
                    upp[j][i][v] = upp[j][i][v] + (4.0 / 3.0) * a;
                }
                T fm[NV], gm[NV], fp[NV], gp[NV], fmm[NV], gmm[NV], fpp[NV], gpp[NV];
                eq.fluxes(um[j][i], fm, gm);
                eq.fluxes(up[j][i], fp, gp);
                eq.fluxes(umm[j][i], fmm, gmm);
                eq.fluxes(upp[j][i], fpp, gpp);
                T fttt[NV], gttt[NV];
                for (int v = 0; v < NV; ++v) {
                    fttt[v] = 0.5 * (fpp[v] - 2.0 * fp[v] + 2.0 * fm[v] - fmm[v]);
                    Fc[j][i][v] = Fc[j][i][v] + (1.0 / 24.0) * fttt[v];
                    gttt[v] = 0.5 * (gpp[v] - 2.0 * gp[v] + 2.0 * gm[v] - gmm[v]);
                    Gc[j][i][v] = Gc[j][i][v] + (1.0 / 24.0) * gttt[v];
                }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) utttt[j][ii][v] = utttt[j][ii][v] + (-dt * D[ii][i]) * fttt[v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) utttt[jj][i][v] = utttt[jj][i][v] + (-dt * D[jj][j]) * gttt[v];
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) utttt[j][i][v] = utttt[j][i][v] * ij;
        if (has_src) src_stage_ttt(e, dt, up, upp, um, umm, utttt, Sc);
        // 1160-1249: f_tttt (with the reference's 0.5 factor), F̃, volume term
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                for (int v = 0; v < NV; ++v) {
                    const T a = utttt[j][i][v];
                    Uc[j][i][v] = Uc[j][i][v] + (1.0 / 120.0) * a;
                    um[j][i][v] = um[j][i][v] + (1.0 / 24.0) * a;
                    up[j][i][v] = up[j][i][v] + (1.0 / 24.0) * a;
                    umm[j][i][v] = umm[j][i][v] + (2.0 / 3.0) * a;
                    upp[j][i][v] = upp[j][i][v] + (2.0 / 3.0) * a;
                }
                T fm[NV], gm[NV], fp[NV], gp[NV], fmm[NV], gmm[NV], fpp[NV], gpp[NV];
                eq.fluxes(um[j][i], fm, gm);
                eq.fluxes(up[j][i], fp, gp);
                eq.fluxes(umm[j][i], fmm, gmm);
                eq.fluxes(upp[j][i], fpp, gpp);
                // u_np1_low uses F̃ before the f_tttt term (1183-1199)
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v)
                        unp1low[j][ii][v] = unp1low[j][ii][v] + (-dt * ij * D[ii][i]) * Fc[j][i][v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v)
                        unp1low[jj][i][v] = unp1low[jj][i][v] + (-dt * ij * D[jj][j]) * Gc[j][i][v];
                for (int v = 0; v < NV; ++v) {
                    const T ftttt = 0.5 * (fpp[v] - 4.0 * fp[v] + 6.0 * f[j][i][v] - 4.0 * fm[v] + fmm[v]);
                    const T gtttt = 0.5 * (gpp[v] - 4.0 * gp[v] + 6.0 * g[j][i][v] - 4.0 * gm[v] + gmm[v]);
                    Fc[j][i][v] = Fc[j][i][v] + (1.0 / 120.0) * ftttt;
                    Gc[j][i][v] = Gc[j][i][v] + (1.0 / 120.0) * gtttt;
                    // u_np1 - u_np1_low: the f_tttt term's own update
                    for (int ii = 0; ii < n; ++ii)
                        dlow[j][ii][v] = dlow[j][ii][v] + (-dt * ij * D[ii][i]) * ((1.0 / 120.0) * ftttt);
                    for (int jj = 0; jj < n; ++jj)
                        dlow[jj][i][v] = dlow[jj][i][v] + (-dt * ij * D[jj][j]) * ((1.0 / 120.0) * gtttt);
                }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) {
                        du[idx(v, ii, j, e)] = du[idx(v, ii, j, e)] + (alpha * Dh[ii][i]) * Fc[j][i][v];
                        unp1[j][ii][v] = unp1[j][ii][v] + (-dt * ij * D[ii][i]) * Fc[j][i][v];
                    }
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) {
                        du[idx(v, i, jj, e)] = du[idx(v, i, jj, e)] + (alpha * Dh[jj][j]) * Gc[j][i][v];
                        unp1[jj][i][v] = unp1[jj][i][v] + (-dt * ij * D[jj][j]) * Gc[j][i][v];
                    }
                if (has_src) {
                    // 1228-1233: the reference passes (u, up, um, upp, umm) to
                    // calc_source_tttt_N4(u, up, upp, um, umm, ...)
                    double x, y;
                    node_xy(i, j, e, x, y);
                    T s4[NV];
                    src_tttt_N4(&u[idx(0, i, j, e)], up[j][i], um[j][i], upp[j][i], umm[j][i], x, y, dt, s4);
                    for (int v = 0; v < NV; ++v) Sc[j][i][v] = Sc[j][i][v] + (1.0 / 120.0) * s4[v];
                }
                for (int v = 0; v < NV; ++v) {
                    F[fidx(v, 0, i, j, e)] = Fc[j][i][v];
                    F[fidx(v, 1, i, j, e)] = Gc[j][i][v];
                    U[idx(v, i, j, e)] = Uc[j][i][v];
                    if (has_src) du[idx(v, i, j, e)] = du[idx(v, i, j, e)] + (-1.0 / ij) * Sc[j][i][v];
                }
            }
        // 1251-1265: temporal error estimate for the adaptive controller.  The
        // numerator u_np1 - u_np1_low is the update of the f_tttt term itself
        // (the same quantity without the cancellation of the difference).
        T err = T(0);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) {
                    const T d = dlow[j][i][v];
                    const T s = abstol + reltol * max_(abs_(unp1[j][i][v]), abs_(unp1low[j][i][v]));
                    const T r = d / s;
                    err = err + r * r;
                }
        temporal_errors[e] = err;
    }

    // ---------------- LW volume kernel, N = 3 ----------------
    // lw_volume_kernel_3!, src/solvers/dgsem_tree/dg_2d.jl:674-916.  Source
    // terms as in the N = 4 kernel (732-740, 785-797, 841-851), with
    // S = s + s_t/2 + s_tt/6 + s_ttt/24; s_ttt at the end takes the reference's
    // arguments (u, up, upp, upp, umm) (894-896: the state of t + 2dt also at
    // t - dt), which a source independent of u does not see.
    void lw_volume_kernel_3(T* du, const T* u, double dt, int e, double alpha) {
        const auto& D = basis.D;
        const auto& Dh = basis.Dhat;
        const double ij = mesh.inv_jac[e];
        const bool has_src = src.kind != SRC_NONE;
        T f[MAXN][MAXN][NV], g[MAXN][MAXN][NV], Fc[MAXN][MAXN][NV], Gc[MAXN][MAXN][NV];
        T ut[MAXN][MAXN][NV], utt[MAXN][MAXN][NV], uttt[MAXN][MAXN][NV], Sc[MAXN][MAXN][NV];
        T Uc[MAXN][MAXN][NV], up[MAXN][MAXN][NV], um[MAXN][MAXN][NV], upp[MAXN][MAXN][NV], umm[MAXN][MAXN][NV];
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) ut[j][i][v] = utt[j][i][v] = uttt[j][i][v] = T(0);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const T* un = &u[idx(0, i, j, e)];
                T fl[NV], gl[NV];
                eq.fluxes(un, fl, gl);
                for (int v = 0; v < NV; ++v) { f[j][i][v] = fl[v]; g[j][i][v] = gl[v]; Fc[j][i][v] = fl[v]; }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) ut[j][ii][v] = ut[j][ii][v] + (-dt * D[ii][i]) * fl[v];
                for (int v = 0; v < NV; ++v) Gc[j][i][v] = gl[v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) ut[jj][i][v] = ut[jj][i][v] + (-dt * D[jj][j]) * gl[v];
                for (int v = 0; v < NV; ++v)
                    um[j][i][v] = up[j][i][v] = umm[j][i][v] = upp[j][i][v] = Uc[j][i][v] = un[v];
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) ut[j][i][v] = ut[j][i][v] * ij;
        if (has_src) src_init(u, e, dt, ut, Sc);
        // 740-769
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                for (int v = 0; v < NV; ++v) {
                    const T a = ut[j][i][v];
                    Uc[j][i][v] = Uc[j][i][v] + 0.5 * a;
                    up[j][i][v] = up[j][i][v] + 1.0 * a;
                    um[j][i][v] = um[j][i][v] + (-1.0) * a;
                    umm[j][i][v] = umm[j][i][v] + (-2.0) * a;
                    upp[j][i][v] = upp[j][i][v] + 2.0 * a;
                }
                T fm[NV], gm[NV], fp[NV], gp[NV], fmm[NV], gmm[NV], fpp[NV], gpp[NV];
                eq.fluxes(um[j][i], fm, gm);
                eq.fluxes(up[j][i], fp, gp);
                eq.fluxes(umm[j][i], fmm, gmm);
                eq.fluxes(upp[j][i], fpp, gpp);
                T ft[NV], gt[NV];
                for (int v = 0; v < NV; ++v) {
                    ft[v] = 1.0 / 12.0 * (-fpp[v] + 8.0 * fp[v] - 8.0 * fm[v] + fmm[v]);
                    gt[v] = 1.0 / 12.0 * (-gpp[v] + 8.0 * gp[v] - 8.0 * gm[v] + gmm[v]);
                    Fc[j][i][v] = Fc[j][i][v] + 0.5 * ft[v];
                    Gc[j][i][v] = Gc[j][i][v] + 0.5 * gt[v];
                }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) utt[j][ii][v] = utt[j][ii][v] + (-dt * D[ii][i]) * ft[v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) utt[jj][i][v] = utt[jj][i][v] + (-dt * D[jj][j]) * gt[v];
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) utt[j][i][v] = utt[j][i][v] * ij;
        if (has_src) src_stage_t(e, dt, up, upp, um, umm, utt, Sc);
        // 799-827: f_tt = fp - 2f + fm (two-point second difference)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                for (int v = 0; v < NV; ++v) {
                    const T a = utt[j][i][v];
                    Uc[j][i][v] = Uc[j][i][v] + (1.0 / 6.0) * a;
                    up[j][i][v] = up[j][i][v] + 0.5 * a;
                    um[j][i][v] = um[j][i][v] + 0.5 * a;
                    upp[j][i][v] = upp[j][i][v] + 2.0 * a;
                    umm[j][i][v] = umm[j][i][v] + 2.0 * a;
                }
                T fm[NV], gm[NV], fp[NV], gp[NV];
                eq.fluxes(um[j][i], fm, gm);
                eq.fluxes(up[j][i], fp, gp);
                T ftt[NV], gtt[NV];
                for (int v = 0; v < NV; ++v) {
                    ftt[v] = fp[v] - 2.0 * f[j][i][v] + fm[v];
                    gtt[v] = gp[v] - 2.0 * g[j][i][v] + gm[v];
                    Fc[j][i][v] = Fc[j][i][v] + (1.0 / 6.0) * ftt[v];
                    Gc[j][i][v] = Gc[j][i][v] + (1.0 / 6.0) * gtt[v];
                }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) uttt[j][ii][v] = uttt[j][ii][v] + (-dt * D[ii][i]) * ftt[v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) uttt[jj][i][v] = uttt[jj][i][v] + (-dt * D[jj][j]) * gtt[v];
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) uttt[j][i][v] = uttt[j][i][v] * ij;
        if (has_src) src_stage_tt(u, e, dt, up, um, uttt, Sc);
        // 853-913: f_ttt, F̃, volume term
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                for (int v = 0; v < NV; ++v) {
                    const T a = uttt[j][i][v];
                    Uc[j][i][v] = Uc[j][i][v] + (1.0 / 24.0) * a;
                    um[j][i][v] = um[j][i][v] + (-1.0 / 6.0) * a;
                    up[j][i][v] = up[j][i][v] + (1.0 / 6.0) * a;
                    umm[j][i][v] = umm[j][i][v] + (-4.0 / 3.0) * a;
                    upp[j][i][v] = upp[j][i][v] + (4.0 / 3.0) * a;
                }
                T fm[NV], gm[NV], fp[NV], gp[NV], fmm[NV], gmm[NV], fpp[NV], gpp[NV];
                eq.fluxes(um[j][i], fm, gm);
                eq.fluxes(up[j][i], fp, gp);
                eq.fluxes(umm[j][i], fmm, gmm);
                eq.fluxes(upp[j][i], fpp, gpp);
                for (int v = 0; v < NV; ++v) {
                    const T fttt = 0.5 * (fpp[v] - 2.0 * fp[v] + 2.0 * fm[v] - fmm[v]);
                    Fc[j][i][v] = Fc[j][i][v] + (1.0 / 24.0) * fttt;
                    const T gttt = 0.5 * (gpp[v] - 2.0 * gp[v] + 2.0 * gm[v] - gmm[v]);
                    Gc[j][i][v] = Gc[j][i][v] + (1.0 / 24.0) * gttt;
                }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, ii, j, e)] = du[idx(v, ii, j, e)] + (alpha * Dh[ii][i]) * Fc[j][i][v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, i, jj, e)] = du[idx(v, i, jj, e)] + (alpha * Dh[jj][j]) * Gc[j][i][v];
                if (has_src) {
                    double x, y;
                    node_xy(i, j, e, x, y);
                    T s3[NV];
                    src_ttt_N34(up[j][i], upp[j][i], upp[j][i], umm[j][i], x, y, dt, s3);
                    for (int v = 0; v < NV; ++v) Sc[j][i][v] = Sc[j][i][v] + (1.0 / 24.0) * s3[v];
                }
                for (int v = 0; v < NV; ++v) {
                    U[idx(v, i, j, e)] = Uc[j][i][v];
                    F[fidx(v, 0, i, j, e)] = Fc[j][i][v];
                    F[fidx(v, 1, i, j, e)] = Gc[j][i][v];
                    if (has_src) du[idx(v, i, j, e)] = du[idx(v, i, j, e)] + (-1.0 / ij) * Sc[j][i][v];
                }
            }
    }

    // ---------------- MDRK stage kernels (TreeMesh) ----------------
    // mdrk_kernel_1!, src/solvers/dgsem_tree/multi_derivative_2d.jl:263-420.
    // The reference kernel also writes `mdrk_cache.u_low`, a field that
    // MDRKElementCache does not have (containers.jl:83-98), so the reference
    // raises there; that low-order update feeds only the admissibility
    // machinery and is left out.  du, Ũ, F̃, U2, F2 follow the kernel.
    // Source terms (323-331, 351-379): ut += dt s(uⁿ, t); S = s/2 + s_t/8 enters
    // du as -S/J⁻¹ and S2 = s + s_t/6 goes to the stage cache.
    void mdrk_kernel_1(T* du, const T* u, double dt, int e, double alpha) {
        const auto& D = basis.D;
        const auto& Dh = basis.Dhat;
        const double ij = mesh.inv_jac[e];
        const bool has_src = src.kind != SRC_NONE;
        T Fc[MAXN][MAXN][NV], F2c[MAXN][MAXN][NV], Gc[MAXN][MAXN][NV], G2c[MAXN][MAXN][NV];
        T ut[MAXN][MAXN][NV], Uc[MAXN][MAXN][NV], U2c[MAXN][MAXN][NV], Sc[MAXN][MAXN][NV];
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) ut[j][i][v] = T(0);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const T* un = &u[idx(0, i, j, e)];
                T fl[NV], gl[NV];
                eq.fluxes(un, fl, gl);
                for (int v = 0; v < NV; ++v) { Fc[j][i][v] = 0.5 * fl[v]; F2c[j][i][v] = fl[v]; }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) ut[j][ii][v] = ut[j][ii][v] + (-dt * D[ii][i]) * fl[v];
                for (int v = 0; v < NV; ++v) { Gc[j][i][v] = 0.5 * gl[v]; G2c[j][i][v] = gl[v]; }
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) ut[jj][i][v] = ut[jj][i][v] + (-dt * D[jj][j]) * gl[v];
                for (int v = 0; v < NV; ++v) { Uc[j][i][v] = 0.5 * un[v]; U2c[j][i][v] = un[v]; }
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) ut[j][i][v] = ut[j][i][v] * ij;
        if (has_src) src_init(u, e, dt, ut, Sc);  // Sc = s; halved below (S = s/2, S2 = s)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const T* un = &u[idx(0, i, j, e)];
                T um[NV], up[NV], umm[NV], upp[NV];
                for (int v = 0; v < NV; ++v) {
                    const T a = ut[j][i][v];
                    um[v] = un[v] - a;
                    up[v] = un[v] + a;
                    umm[v] = un[v] - 2.0 * a;
                    upp[v] = un[v] + 2.0 * a;
                }
                T fm[NV], gm[NV], fp[NV], gp[NV], fmm[NV], gmm[NV], fpp[NV], gpp[NV];
                eq.fluxes(um, fm, gm);
                eq.fluxes(up, fp, gp);
                eq.fluxes(umm, fmm, gmm);
                eq.fluxes(upp, fpp, gpp);
                for (int v = 0; v < NV; ++v) {
                    const T ft = 1.0 / 12.0 * (-fpp[v] + 8.0 * fp[v] - 8.0 * fm[v] + fmm[v]);
                    const T gt = 1.0 / 12.0 * (-gpp[v] + 8.0 * gp[v] - 8.0 * gm[v] + gmm[v]);
                    Fc[j][i][v] = Fc[j][i][v] + 0.125 * ft;
                    Gc[j][i][v] = Gc[j][i][v] + 0.125 * gt;
                    Uc[j][i][v] = Uc[j][i][v] + 0.125 * ut[j][i][v];
                    F2c[j][i][v] = F2c[j][i][v] + (1.0 / 6.0) * ft;
                    G2c[j][i][v] = G2c[j][i][v] + (1.0 / 6.0) * gt;
                    U2c[j][i][v] = U2c[j][i][v] + (1.0 / 6.0) * ut[j][i][v];
                }
                T S1[NV], S2n[NV];
                if (has_src) {
                    double x, y;
                    node_xy(i, j, e, x, y);
                    T st[NV];
                    src_t_N34(up, upp, um, umm, x, y, dt, src_t, st);
                    for (int v = 0; v < NV; ++v) {
                        S1[v] = 0.5 * Sc[j][i][v] + 0.125 * st[v];
                        S2n[v] = Sc[j][i][v] + (1.0 / 6.0) * st[v];
                    }
                }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, ii, j, e)] = du[idx(v, ii, j, e)] + (alpha * Dh[ii][i]) * Fc[j][i][v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, i, jj, e)] = du[idx(v, i, jj, e)] + (alpha * Dh[jj][j]) * Gc[j][i][v];
                for (int v = 0; v < NV; ++v) {
                    U[idx(v, i, j, e)] = Uc[j][i][v];
                    F[fidx(v, 0, i, j, e)] = Fc[j][i][v];
                    F[fidx(v, 1, i, j, e)] = Gc[j][i][v];
                    U2[idx(v, i, j, e)] = U2c[j][i][v];
                    F2[fidx(v, 0, i, j, e)] = F2c[j][i][v];
                    F2[fidx(v, 1, i, j, e)] = G2c[j][i][v];
                    if (has_src) {
                        du[idx(v, i, j, e)] = du[idx(v, i, j, e)] + (-1.0 / ij) * S1[v];
                        S2[idx(v, i, j, e)] = S2n[v];
                    }
                }
            }
    }

    // mdrk_kernel_2!, src/solvers/dgsem_tree/multi_derivative_2d.jl:423-555
    // Source terms (487-522): S starts from the stage cache's S2, ust += dt
    // s(us, t + dt/2), S += s_t/3 with s_t of the us stencil about t + dt/2.
    void mdrk_kernel_2(T* du, double dt, int e, double alpha) {
        const auto& D = basis.D;
        const auto& Dh = basis.Dhat;
        const double ij = mesh.inv_jac[e];
        const bool has_src = src.kind != SRC_NONE;
        T Fc[MAXN][MAXN][NV], Gc[MAXN][MAXN][NV], ust[MAXN][MAXN][NV], Uc[MAXN][MAXN][NV];
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) {
                    Fc[j][i][v] = F2[fidx(v, 0, i, j, e)];
                    Gc[j][i][v] = F2[fidx(v, 1, i, j, e)];
                    Uc[j][i][v] = U2[idx(v, i, j, e)];
                    ust[j][i][v] = T(0);
                }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const T* usn = &us[idx(0, i, j, e)];
                T fl[NV], gl[NV];
                eq.fluxes(usn, fl, gl);
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v) ust[j][ii][v] = ust[j][ii][v] + (-dt * D[ii][i]) * fl[v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v) ust[jj][i][v] = ust[jj][i][v] + (-dt * D[jj][j]) * gl[v];
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v) ust[j][i][v] = ust[j][i][v] * ij;
        if (has_src)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double x, y;
                    node_xy(i, j, e, x, y);
                    T s[NV];
                    src_at(&us[idx(0, i, j, e)], x, y, 0.5 * dt, s);
                    for (int v = 0; v < NV; ++v) ust[j][i][v] = ust[j][i][v] + dt * s[v];
                }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const T* usn = &us[idx(0, i, j, e)];
                T um[NV], up[NV], umm[NV], upp[NV];
                for (int v = 0; v < NV; ++v) {
                    const T a = ust[j][i][v];
                    um[v] = usn[v] - a;
                    up[v] = usn[v] + a;
                    umm[v] = usn[v] - 2.0 * a;
                    upp[v] = usn[v] + 2.0 * a;
                }
                T fm[NV], gm[NV], fp[NV], gp[NV], fmm[NV], gmm[NV], fpp[NV], gpp[NV];
                eq.fluxes(um, fm, gm);
                eq.fluxes(up, fp, gp);
                eq.fluxes(umm, fmm, gmm);
                eq.fluxes(upp, fpp, gpp);
                for (int v = 0; v < NV; ++v) {
                    const T ft = 1.0 / 12.0 * (-fpp[v] + 8.0 * fp[v] - 8.0 * fm[v] + fmm[v]);
                    const T gt = 1.0 / 12.0 * (-gpp[v] + 8.0 * gp[v] - 8.0 * gm[v] + gmm[v]);
                    Fc[j][i][v] = Fc[j][i][v] + (1.0 / 3.0) * ft;
                    Gc[j][i][v] = Gc[j][i][v] + (1.0 / 3.0) * gt;
                    Uc[j][i][v] = Uc[j][i][v] + (1.0 / 3.0) * ust[j][i][v];
                }
                T Sn[NV];
                if (has_src) {
                    double x, y;
                    node_xy(i, j, e, x, y);
                    T st[NV];
                    src_t_N34(up, upp, um, umm, x, y, dt, src_t + 0.5 * dt, st);
                    for (int v = 0; v < NV; ++v) Sn[v] = S2[idx(v, i, j, e)] + (1.0 / 3.0) * st[v];
                }
                for (int ii = 0; ii < n; ++ii)
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, ii, j, e)] = du[idx(v, ii, j, e)] + (alpha * Dh[ii][i]) * Fc[j][i][v];
                for (int jj = 0; jj < n; ++jj)
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, i, jj, e)] = du[idx(v, i, jj, e)] + (alpha * Dh[jj][j]) * Gc[j][i][v];
                for (int v = 0; v < NV; ++v) {
                    U[idx(v, i, j, e)] = Uc[j][i][v];
                    F[fidx(v, 0, i, j, e)] = Fc[j][i][v];
                    F[fidx(v, 1, i, j, e)] = Gc[j][i][v];
                    if (has_src) du[idx(v, i, j, e)] = du[idx(v, i, j, e)] + (-1.0 / ij) * Sn[v];
                }
            }
    }


    // ---------------- shock capturing ----------------
    // Trixi IndicatorHennemannGassner (Hennemann et al., JCP 426, 2021):
    // modal energy of density_pressure, threshold 0.5·10^(-1.8 n^0.25),
    // sharpness s = ln((1-1e-4)/1e-4), clipping to 0 below alpha_min and to 1
    // above 1-alpha_min, alpha_max cap, then smoothing over interfaces
    // (alpha_e >= ½ alpha_neighbour), called at dg_2d.jl:1665.
    void indicator_hg(const T* u) {
        const double threshold = 0.5 * std::pow(10.0, -1.8 * std::pow((double)n, 0.25));
        const double par_s = std::log((1 - 0.0001) / 0.0001);
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) {
            double ind[MAXN][MAXN], tmp[MAXN][MAXN], modal[MAXN][MAXN];
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) ind[j][i] = value(eq.density_pressure(&u[idx(0, i, j, e)]));
            // modal[j][i] = Σ Vinv[i][ii] Vinv[j][jj] ind[jj][ii]
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double acc = 0;
                    for (int ii = 0; ii < n; ++ii) acc += basis.Vinv[i][ii] * ind[j][ii];
                    tmp[j][i] = acc;
                }
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double acc = 0;
                    for (int jj = 0; jj < n; ++jj) acc += basis.Vinv[j][jj] * tmp[jj][i];
                    modal[j][i] = acc;
                }
            double total = 0, clip1 = 0, clip2 = 0;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) total += modal[j][i] * modal[j][i];
            for (int j = 0; j < n - 1; ++j)
                for (int i = 0; i < n - 1; ++i) clip1 += modal[j][i] * modal[j][i];
            for (int j = 0; j < n - 2; ++j)
                for (int i = 0; i < n - 2; ++i) clip2 += modal[j][i] * modal[j][i];
            const double f1 = total != 0 ? (total - clip1) / total : 0.0;
            const double f2 = clip1 != 0 ? (clip1 - clip2) / clip1 : 0.0;
            const double energy = std::max(f1, f2);
            double a = 1.0 / (1.0 + std::exp(-par_s / threshold * (energy - threshold)));
            if (a < alpha_min) a = 0.0;
            if (a > 1 - alpha_min) a = 1.0;
            alpha[e] = std::min(alpha_max, a);
        }
        if (alpha_smooth) {
            alpha_tmp = alpha;
            const int nI = (int)mesh.if_left.size();
            for (int I = 0; I < nI; ++I) {
                const int L = mesh.if_left[I], R = mesh.if_right[I];
                alpha[L] = std::max(std::max(alpha_tmp[L], 0.5 * alpha_tmp[R]), alpha[L]);
                alpha[R] = std::max(std::max(alpha_tmp[R], 0.5 * alpha_tmp[L]), alpha[R]);
            }
        }
    }

    // Trixi flux_lax_friedrichs on two states (FluxPlusDissipation of
    // flux_central and DissipationLocalLaxFriedrichs with max_abs_speed_naive)
    void rusanov(const T* ul, const T* ur, int orientation, T* out) const {
        T fl[NV], fr[NV];
        eq.flux(ul, orientation, fl);
        eq.flux(ur, orientation, fr);
        const T lam = eq.max_abs_speed_naive(ul, ur, orientation);
        for (int v = 0; v < NV; ++v) out[v] = 0.5 * (fl[v] + fr[v]) + (-0.5 * lam) * (ur[v] - ul[v]);
    }

    // First-order subcell FV of a blended element (fv_kernel!, FirstOrderReconstruction,
    // dg_2d.jl:251-282, with Trixi calcflux_fv!) and the fluxes next to the element
    // faces stored in fn_low (load_fn_low!, dg_2d.jl:220-234).  With fv_scale = 0
    // only fn_low is produced (calc_fn_low!, dg_2d.jl:1948-1991).
    void fv_kernel_fo(T* du, const T* u, int e, double fv_scale, bool add) {
        // fx[j][i]: flux between nodes i-1 and i (i = 1..n-1); fx[j][0] = fx[j][n] = 0
        T fx[MAXN][MAXN + 1][NV], gy[MAXN + 1][MAXN][NV];
        for (int j = 0; j < n; ++j)
            for (int v = 0; v < NV; ++v) { fx[j][0][v] = T(0); fx[j][n][v] = T(0); }
        for (int i = 0; i < n; ++i)
            for (int v = 0; v < NV; ++v) { gy[0][i][v] = T(0); gy[n][i][v] = T(0); }
        for (int j = 0; j < n; ++j)
            for (int i = 1; i < n; ++i) rusanov(&u[idx(0, i - 1, j, e)], &u[idx(0, i, j, e)], 1, fx[j][i]);
        for (int j = 1; j < n; ++j)
            for (int i = 0; i < n; ++i) rusanov(&u[idx(0, i, j - 1, e)], &u[idx(0, i, j, e)], 2, gy[j][i]);
        if (add) {
            const auto& iw = basis.inv_w;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v)
                        du[idx(v, i, j, e)] = du[idx(v, i, j, e)] +
                                              fv_scale * (iw[i] * (fx[j][i + 1][v] - fx[j][i][v]) +
                                                          iw[j] * (gy[j + 1][i][v] - gy[j][i][v]));
        }
        for (int k = 0; k < n; ++k)
            for (int v = 0; v < NV; ++v) {
                fn_low[sidx(v, k, 0, e)] = fx[k][1][v];
                fn_low[sidx(v, k, 1, e)] = fx[k][n - 1][v];
                fn_low[sidx(v, k, 2, e)] = gy[1][k][v];
                fn_low[sidx(v, k, 3, e)] = gy[n - 1][k][v];
            }
    }

    // ---------------- MUSCL-Hancock subcell reconstruction ----------------
    // fv_kernel! with MUSCLHancockReconstruction (dg_2d.jl:284-324) calls
    // calcflux_mh!, which the reference defines for curvilinear meshes only
    // (structured/dg_2d.jl:313-538).  Restated here for the TreeMesh with its
    // Cartesian metric: contravariant vectors (1, 0) and (0, 1) (every subcell
    // normal Ja + Σ w D Ja is then the same unit vector) and the element's
    // inverse Jacobian 2/h at every node.  Kept as the reference computes it:
    //   - subcell geometry of compute_subcells (dg_2d.jl:1806-1824): faces
    //     xs[0] = -1, xs[i] = -1 + w_1 + ... + w_i, extended nodes ξe[0] = -1,
    //     ξe[1..n] = GLL nodes, ξe[n+1] = 1;
    //   - uext copies the element's edge values outward, so at the end subcells
    //     (ξe[0] = ξe[1] = -1) the one-sided differences are 0/0 = NaN and minmod
    //     returns 0 (Julia: sign(NaN) is NaN and NaN != NaN);
    //   - minmod(back, centred, forward, β = 2 - α, 0) (structured/dg_2d.jl:75-95);
    //   - limit_slope on density then pressure (compressible_euler_2d.jl:9-21,
    //     equations.jl:262-287), which scales u* by θ a second time;
    //   - the half-step (structured/dg_2d.jl:396-475): subcell k (1-based) gets
    //     -Δt/2 J⁻¹ f(its right face value) / w_{k+1} for k < n and
    //     +Δt/2 J⁻¹ f(its left face value) / w_{k+1} for 1 < k < n, in both
    //     directions and added to all four face values; the end subcells get no
    //     other term;
    //   - Rusanov fluxes between the evolved face values at the interior subcell
    //     faces, zero at the element faces;
    //   - elements with a face on the domain boundary use first-order FV
    //     (structured/dg_2d.jl:319-327).
    // Euler only (limit_slope has no other method in the reference).
    static T sign_julia(const T& x) {
        const double v = value(x);
        if (v != v) return x;  // NaN
        return v > 0 ? T(1.0) : (v < 0 ? T(-1.0) : T(0.0));
    }
    static bool isnan_(const T& x) { const double v = value(x); return v != v; }
    static T min3_julia(const T& a, const T& b, const T& c) {
        if (isnan_(a)) return a;
        if (isnan_(b)) return b;
        if (isnan_(c)) return c;
        T m = value(b) < value(a) ? b : a;
        return value(c) < value(m) ? c : m;
    }
    static T minmod_mh(const T& a, const T& b, const T& c, double beta) {
        const T slope = min3_julia(beta * abs_(a), abs_(b), beta * abs_(c));
        const T s1 = sign_julia(a), s2 = sign_julia(b), s3 = sign_julia(c);
        const bool ne12 = isnan_(s1) || isnan_(s2) || value(s1) != value(s2);
        const bool ne23 = isnan_(s2) || isnan_(s3) || value(s2) != value(s3);
        if (ne12 || ne23) return T(0.0);
        return s1 * slope;
    }
    // limit_variable_slope (equations.jl:262-287) for var = density (0) or pressure (1)
    void limit_variable_slope(int var, T* slope, T* us_ll, T* us_rr, const T* ue, double xl, double xr) const {
        auto vf = [&](const T* q) { return var == 0 ? eq.density(q) : eq.pressure(q); };
        const T v_ll = vf(us_ll), v_rr = vf(us_rr), v_low = vf(ue);
        const double eps = 1e-10;
        const T threshold = 0.1 * v_low;
        if (value(v_ll) < eps || value(v_rr) < eps) {
            const T r_ll = abs_(threshold - v_low) / (abs_(v_ll - v_low) + 1e-13);
            const T r_rr = abs_(threshold - v_low) / (abs_(v_rr - v_low) + 1e-13);
            T theta = value(r_rr) < value(r_ll) ? r_rr : r_ll;
            if (!(value(theta) < 1.0)) theta = T(1.0);
            for (int v = 0; v < NV; ++v) {
                slope[v] = slope[v] * theta;
                us_ll[v] = ue[v] + (2.0 * theta) * (xl * slope[v]);
                us_rr[v] = ue[v] + (2.0 * theta) * (xr * slope[v]);
            }
        }
    }
    // limit_slope (compressible_euler_2d.jl:9-21): face values ue + slope·x of the limited slope
    void limit_slope(T* slope, T* ufl, T* us_ll, T* ufr, T* us_rr, const T* ue, double xl, double xr) const {
        limit_variable_slope(0, slope, us_ll, us_rr, ue, xl, xr);
        limit_variable_slope(1, slope, us_ll, us_rr, ue, xl, xr);
        for (int v = 0; v < NV; ++v) {
            ufl[v] = ue[v] + slope[v] * xl;
            ufr[v] = ue[v] + slope[v] * xr;
        }
    }
    bool boundary_adjacent(int e) const {
        for (int f = 0; f < 4; ++f)
            if (mesh.nbr[4 * e + f] < 0) return true;
        return false;
    }
    // fv_kernel! with MUSCL-Hancock: the subcell fluxes, α-scaled FV term and fn_low
    void fv_kernel_mh(T* du, const T* u, int e, double fv_scale, double dtmh) {
        if (boundary_adjacent(e)) {
            fv_kernel_fo(du, u, e, fv_scale, true);
            return;
        }
        const auto& w = basis.weights;
        double xs[MAXN + 1], xe[MAXN + 2];
        xs[0] = -1.0;
        for (int i = 1; i <= n; ++i) {
            double acc = w[0];
            for (int k = 1; k < i; ++k) acc += w[k];  // sum(weights[1:i]), left to right
            xs[i] = -1.0 + acc;
        }
        xe[0] = -1.0;
        for (int i = 1; i <= n; ++i) xe[i] = basis.nodes[i - 1];
        xe[n + 1] = 1.0;
        const double ij = mesh.inv_jac[e];
        // uext[j][i] with 1-based subcells and the copies at 0 and n+1
        T uext[MAXN + 2][MAXN + 2][NV];
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= n; ++i)
                for (int v = 0; v < NV; ++v) uext[j][i][v] = u[idx(v, i - 1, j - 1, e)];
        for (int k = 1; k <= n; ++k)
            for (int v = 0; v < NV; ++v) {
                uext[k][0][v] = uext[k][1][v];
                uext[k][n + 1][v] = uext[k][n][v];
                uext[0][k][v] = uext[1][k][v];
                uext[n + 1][k][v] = uext[n][k][v];
            }
        T uf[MAXN + 1][MAXN + 1][4][NV], unph[MAXN + 1][MAXN + 1][4][NV];
        const double beta = 2.0 - fv_scale;
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= n; ++i) {
                const T* u_ = uext[j][i];
                const T *ul = uext[j][i - 1], *ur = uext[j][i + 1], *ud = uext[j - 1][i], *uu = uext[j + 1][i];
                const double dx1 = xe[i] - xe[i - 1], dx2 = xe[i + 1] - xe[i];
                const double dy1 = xe[j] - xe[j - 1], dy2 = xe[j + 1] - xe[j];
                T sx[NV], sy[NV];
                for (int v = 0; v < NV; ++v) {
                    // finite_differences (structured/dg_2d.jl:75-81)
                    {
                        const double a = -(dx2 / (dx1 * (dx1 + dx2))), b = (dx2 - dx1) / (dx1 * dx2),
                                     c = (dx1 / (dx2 * (dx1 + dx2)));
                        const T back = (u_[v] - ul[v]) / dx1, fwd = (ur[v] - u_[v]) / dx2;
                        const T cent = a * ul[v] + b * u_[v] + c * ur[v];
                        sx[v] = minmod_mh(back, cent, fwd, beta);
                    }
                    {
                        const double a = -(dy2 / (dy1 * (dy1 + dy2))), b = (dy2 - dy1) / (dy1 * dy2),
                                     c = (dy1 / (dy2 * (dy1 + dy2)));
                        const T back = (u_[v] - ud[v]) / dy1, fwd = (uu[v] - u_[v]) / dy2;
                        const T cent = a * ud[v] + b * u_[v] + c * uu[v];
                        sy[v] = minmod_mh(back, cent, fwd, beta);
                    }
                }
                const double xl = xs[i - 1] - xe[i], xr = xs[i] - xe[i];
                const double yd = xs[j - 1] - xe[j], yu = xs[j] - xe[j];
                T us_l[NV], us_r[NV], us_d[NV], us_u[NV];
                for (int v = 0; v < NV; ++v) {
                    uf[j][i][0][v] = u_[v] + sx[v] * xl;
                    uf[j][i][1][v] = u_[v] + sx[v] * xr;
                    uf[j][i][2][v] = u_[v] + sy[v] * yd;
                    uf[j][i][3][v] = u_[v] + sy[v] * yu;
                    us_l[v] = u_[v] + 2.0 * sx[v] * xl;
                    us_r[v] = u_[v] + 2.0 * sx[v] * xr;
                    us_d[v] = u_[v] + 2.0 * sy[v] * yd;
                    us_u[v] = u_[v] + 2.0 * sy[v] * yu;
                }
                limit_slope(sx, uf[j][i][0], us_l, uf[j][i][1], us_r, u_, xl, xr);
                limit_slope(sy, uf[j][i][2], us_d, uf[j][i][3], us_u, u_, yd, yu);
                for (int q = 0; q < 4; ++q)
                    for (int v = 0; v < NV; ++v) unph[j][i][q][v] = uf[j][i][q][v];
            }
        auto add_all = [&](int i, int j, const T* r) {
            for (int q = 0; q < 4; ++q)
                for (int v = 0; v < NV; ++v) unph[j][i][q][v] = unph[j][i][q][v] + r[v];
        };
        T f[NV], r[NV];
        for (int j = 1; j <= n; ++j)
            for (int i = 2; i <= n; ++i) {
                const double dxi = xs[i] - xs[i - 1];
                eq.flux(uf[j][i - 1][1], 1, f);
                for (int v = 0; v < NV; ++v) r[v] = -0.5 * dtmh * ij / dxi * f[v];
                add_all(i - 1, j, r);
                if (i < n) {
                    eq.flux(uf[j][i][0], 1, f);
                    const double dxip1 = xs[i + 1] - xs[i];
                    for (int v = 0; v < NV; ++v) r[v] = 0.5 * dtmh * ij / dxip1 * f[v];
                    add_all(i, j, r);
                }
            }
        for (int i = 1; i <= n; ++i)
            for (int j = 2; j <= n; ++j) {
                const double dyi = xs[j] - xs[j - 1];
                eq.flux(uf[j - 1][i][3], 2, f);
                for (int v = 0; v < NV; ++v) r[v] = -0.5 * dtmh * ij / dyi * f[v];
                add_all(i, j - 1, r);
                if (j < n) {
                    eq.flux(uf[j][i][2], 2, f);
                    const double dyip1 = xs[j + 1] - xs[j];
                    for (int v = 0; v < NV; ++v) r[v] = 0.5 * dtmh * ij / dyip1 * f[v];
                    add_all(i, j, r);
                }
            }
        // subcell fluxes between evolved face values (0-based node indices as in fv_kernel_fo)
        T fx[MAXN][MAXN + 1][NV], gy[MAXN + 1][MAXN][NV];
        for (int j = 0; j < n; ++j)
            for (int v = 0; v < NV; ++v) { fx[j][0][v] = T(0); fx[j][n][v] = T(0); }
        for (int i = 0; i < n; ++i)
            for (int v = 0; v < NV; ++v) { gy[0][i][v] = T(0); gy[n][i][v] = T(0); }
        for (int j = 1; j <= n; ++j)
            for (int i = 2; i <= n; ++i) rusanov(unph[j][i - 1][1], unph[j][i][0], 1, fx[j - 1][i - 1]);
        for (int i = 1; i <= n; ++i)
            for (int j = 2; j <= n; ++j) rusanov(unph[j - 1][i][3], unph[j][i][2], 2, gy[j - 1][i - 1]);
        const auto& iw = basis.inv_w;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int v = 0; v < NV; ++v)
                    du[idx(v, i, j, e)] = du[idx(v, i, j, e)] + fv_scale * (iw[i] * (fx[j][i + 1][v] - fx[j][i][v]) +
                                                                             iw[j] * (gy[j + 1][i][v] - gy[j][i][v]));
        for (int k = 0; k < n; ++k)
            for (int v = 0; v < NV; ++v) {
                fn_low[sidx(v, k, 0, e)] = fx[k][1][v];
                fn_low[sidx(v, k, 1, e)] = fx[k][n - 1][v];
                fn_low[sidx(v, k, 2, e)] = gy[1][k][v];
                fn_low[sidx(v, k, 3, e)] = gy[n - 1][k][v];
            }
    }

    // compute_alp for VolumeIntegralFRShockCapturing on TreeMesh (dg_2d.jl:1879-1901):
    // average of the two blending factors, reset to 1 when the low-order update
    // of either adjacent node with the blended flux is not admissible.  J is the
    // interface cache's inverse Jacobian, the first element's value on a TreeMesh
    // (containers_2d.jl load_inverse_jacobian!).
    double compute_alp(const T* ul, const T* ur, int L, int R, double J, double dt, const T* fn, const T* Fn_,
                       const T* fin_ll, const T* fin_rr) const {
        const double alp = 0.5 * (alpha[L] + alpha[R]);
        T Fn[NV], low[NV];
        for (int v = 0; v < NV; ++v) Fn[v] = (1.0 - alp) * Fn_[v] + alp * fn[v];
        const double lx = 0.5;
        const double cl = dt * J / (basis.weights[n - 1] * lx);
        for (int v = 0; v < NV; ++v) low[v] = ul[v] - cl * (Fn[v] - fin_ll[v]);
        if (!eq.is_admissible(low)) return 1.0;
        const double cr = dt * J / (basis.weights[0] * lx);
        for (int v = 0; v < NV; ++v) low[v] = ur[v] - cr * (fin_rr[v] - Fn[v]);
        if (!eq.is_admissible(low)) return 1.0;
        return alp;
    }

    // ---------------- interface flux ----------------
    // The 8-argument FluxPlusDissipation with flux_central and the 6-argument
    // DissipationLocalLaxFriedrichs (src/equations/numerical_fluxes.jl:7-38):
    //   F* = 0.5*(f_ll + f_rr) + (-0.5*λ)*(q_rr - q_ll).
    // TreeMesh passes (F̃, F̃, Ũ, Ũ, uⁿ, uⁿ), so λ = max_abs_speed_naive(Ũ)
    // and the jump is on uⁿ (dg_2d.jl:1928); the Structured/P4est order is
    // (F̃n, F̃n, uⁿ, uⁿ, Ũ, Ũ) (structured/dg_2d.jl:2271-2274).
    void lw_numflux(const T* fl, const T* fr, const T* Ul, const T* Ur, const T* ul, const T* ur, int orientation,
                    T* out) const {
        lw_numflux_ord(fl, fr, Ul, Ur, ul, ur, orientation, flux_order == 0 ? 0 : 1, out);
    }
    void lw_numflux_ord(const T* fl, const T* fr, const T* Ul, const T* Ur, const T* ul, const T* ur, int orientation,
                        int order, T* out) const {
        const T* sl = order == 0 ? Ul : ul;  // state for λ
        const T* sr = order == 0 ? Ur : ur;
        const T* jl = order == 0 ? ul : Ul;  // state for the jump
        const T* jr = order == 0 ? ur : Ur;
        const T lam = eq.max_abs_speed_naive(sl, sr, orientation);
        for (int v = 0; v < NV; ++v) out[v] = 0.5 * (fl[v] + fr[v]) + (-0.5 * lam) * (jr[v] - jl[v]);
    }

    // prolong2interfaces! + calc_interface_flux! for TreeMesh and plain FR
    // (dg_2d.jl:1269-1326, 1903-1946).  compute_alp returns 0 for
    // VolumeIntegralFR (dg_2d.jl:1873-1877), so Fn = F*; the low-order
    // Rusanov flux `fn` it also evaluates is multiplied by zero.
    void calc_interface_flux(const T* u) {
        const int nI = (int)mesh.if_left.size();
#pragma omp parallel for schedule(static)
        for (int I = 0; I < nI; ++I) {
            const int L = mesh.if_left[I], R = mesh.if_right[I], o = mesh.if_orient[I];
            const int ldir = 2 * o - 1, rdir = 2 * o - 2;  // 0-based: +face of L, -face of R
            for (int l = 0; l < n; ++l) {
                int li, lj, ri, rj;
                if (o == 1) { li = n - 1; lj = l; ri = 0; rj = l; }
                else { li = l; lj = n - 1; ri = l; rj = 0; }
                const T* ul = &u[idx(0, li, lj, L)];
                const T* ur = &u[idx(0, ri, rj, R)];
                const T* Ul = &U[idx(0, li, lj, L)];
                const T* Ur = &U[idx(0, ri, rj, R)];
                const T* fl = &F[fidx(0, o - 1, li, lj, L)];
                const T* fr = &F[fidx(0, o - 1, ri, rj, R)];
                T Fn[NV];
                lw_numflux(fl, fr, Ul, Ur, ul, ur, o, Fn);
                if (sc) {
                    T fn[NV];
                    rusanov(ul, ur, o, fn);
                    const T* fin_ll = &fn_low[sidx(0, l, ldir, L)];
                    const T* fin_rr = &fn_low[sidx(0, l, rdir, R)];
                    const double J = mesh.inv_jac[0];
                    const double alp = compute_alp(ul, ur, L, R, J, sc_dt, fn, Fn, fin_ll, fin_rr);
                    for (int v = 0; v < NV; ++v) Fn[v] = alp * fn[v] + (1.0 - alp) * Fn[v];
                }
                for (int v = 0; v < NV; ++v) {
                    sfv[sidx(v, l, ldir, L)] = Fn[v];
                    sfv[sidx(v, l, rdir, R)] = Fn[v];
                }
            }
        }
    }

    // prolong2boundaries! + calc_boundary_flux! with the LW time-averaged
    // Dirichlet condition (dg_2d.jl:1328-1454; equations.jl:6-37).
    // `scaling_factor` is 0.5 in MDRK stage 1 (tree/multi_derivative_2d.jl:34).
    void calc_boundary_flux(const T* u, double t, double dt, double scaling_factor) {
        const int nB = (int)mesh.bd_elem.size();
        const auto& nodes = basis.nodes;
        const auto& w = basis.weights;
#pragma omp parallel for schedule(static)
        for (int B = 0; B < nB; ++B) {
            const int e = mesh.bd_elem[B], o = mesh.bd_orient[B], side = mesh.bd_side[B], dir = mesh.bd_dir[B];
            for (int l = 0; l < n; ++l) {
                int ii, jj;
                if (o == 1) { ii = (side == 1) ? n - 1 : 0; jj = l; }
                else { ii = l; jj = (side == 1) ? n - 1 : 0; }
                const T* u_in = &u[idx(0, ii, jj, e)];
                const T* U_in = &U[idx(0, ii, jj, e)];
                const T* F_in = &F[fidx(0, o - 1, ii, jj, e)];
                const double x = mesh.node_x(basis, ii, e), y = mesh.node_y(basis, jj, e);
                T Uo[NV], Fo[NV], uo[NV];
                for (int v = 0; v < NV; ++v) { Uo[v] = T(0); Fo[v] = T(0); }
                const double dts = scaling_factor * dt;
                if (bval.kind == BV_SLIP_WALL) {
                    // mirror states: the normal momentum changes sign; the normal flux
                    // keeps only its normal-momentum component (F_outer = (-F1, F2, -F3,
                    // -F4) on x faces, (-F1, -F2, F3, -F4) on y faces)
                    for (int v = 0; v < NV; ++v) {
                        const bool m = (v == o);
                        Uo[v] = m ? -U_in[v] : U_in[v];
                        uo[v] = m ? -u_in[v] : u_in[v];
                        Fo[v] = m ? F_in[v] : -F_in[v];
                    }
                    T Fn[NV];
                    const int bo = flux_order == 2 ? 0 : flux_order;
                    if (dir % 2 == 0) lw_numflux_ord(F_in, Fo, U_in, Uo, u_in, uo, o, bo, Fn);
                    else lw_numflux_ord(Fo, F_in, Uo, U_in, uo, u_in, o, bo, Fn);
                    for (int v = 0; v < NV; ++v) sfv[sidx(v, l, dir - 1, e)] = Fn[v];
                    continue;
                }
                for (int q = 0; q < n; ++q) {
                    const double ts = t + 0.5 * dts * (nodes[q] + 1.0);
                    double ubd[4];
                    bval.eval(x, y, ts, ubd);
                    T ub[NV], fb[NV];
                    for (int v = 0; v < NV; ++v) ub[v] = T(ubd[v]);
                    eq.flux(ub, o, fb);
                    for (int v = 0; v < NV; ++v) {
                        Uo[v] = Uo[v] + 0.5 * scaling_factor * ub[v] * w[q];
                        Fo[v] = Fo[v] + 0.5 * scaling_factor * fb[v] * w[q];
                    }
                }
                double uod[4];
                bval.eval(x, y, t, uod);
                for (int v = 0; v < NV; ++v) uo[v] = T(uod[v]);
                T Fn[NV];
                // flux_order 2: Structured order at interfaces, TreeMesh order at the
                // boundary (a StructuredMesh's Dirichlet call, structured/dg_2d.jl:2374-2378
                // into equations.jl:6-37)
                const int bo = flux_order == 2 ? 0 : flux_order;
                if (dir % 2 == 0) lw_numflux_ord(F_in, Fo, U_in, Uo, u_in, uo, o, bo, Fn);
                else lw_numflux_ord(Fo, F_in, Uo, U_in, uo, u_in, o, bo, Fn);
                for (int v = 0; v < NV; ++v) sfv[sidx(v, l, dir - 1, e)] = Fn[v];
            }
        }
    }

    // prolong2mortars! + calc_mortar_flux! for the TreeMesh (dg_2d.jl:1456-1646):
    // on each half of the mortar, uⁿ, Ũ and the normal component of F̃ of the
    // small element's face node next to the large element's values interpolated
    // by forward_lower/upper, the LW flux in the interface order (the large
    // element on the left when large_side is 1); then Trixi's
    // mortar_fluxes_to_elements!: each small element takes its half's fluxes,
    // the large element reverse_upper·F*_upper + reverse_lower·F*_lower.
    void calc_mortar_flux(const T* u) {
        const int nM = (int)mesh.mortars.size();
        auto face_node = [&](int dir, int l, int& i, int& j) {
            if (dir < 2) { i = (dir == 0) ? 0 : n - 1; j = l; }
            else { i = l; j = (dir == 2) ? 0 : n - 1; }
        };
#pragma omp parallel for schedule(static)
        for (int m = 0; m < nM; ++m) {
            const auto& M = mesh.mortars[m];
            const int o = M.orient;
            // 0-based faces: the large element's + face and the small elements' - faces
            // when the large element is on the left (large_side 1), and the reverse
            const int ldir = (M.large_side == 1) ? 2 * o - 1 : 2 * o - 2;
            const int sdir = (M.large_side == 1) ? 2 * o - 2 : 2 * o - 1;
            T fstar[2][MAXN][NV];
            for (int half = 0; half < 2; ++half) {  // 0 lower, 1 upper
                const int es = half ? M.upper : M.lower;
                const double(*P)[MAXN] = half ? basis.fwd_up : basis.fwd_lo;
                for (int l = 0; l < n; ++l) {
                    int i, j;
                    face_node(sdir, l, i, j);
                    const T* us_ = &u[idx(0, i, j, es)];
                    const T* Us_ = &U[idx(0, i, j, es)];
                    const T* Fs_ = &F[fidx(0, o - 1, i, j, es)];
                    T uL[NV], UL[NV], FL[NV];
                    for (int v = 0; v < NV; ++v) { uL[v] = T(0); UL[v] = T(0); FL[v] = T(0); }
                    for (int k = 0; k < n; ++k) {
                        int ii, jj;
                        face_node(ldir, k, ii, jj);
                        for (int v = 0; v < NV; ++v) {
                            uL[v] = uL[v] + P[l][k] * u[idx(v, ii, jj, M.large)];
                            UL[v] = UL[v] + P[l][k] * U[idx(v, ii, jj, M.large)];
                            FL[v] = FL[v] + P[l][k] * F[fidx(v, o - 1, ii, jj, M.large)];
                        }
                    }
                    T Fn[NV];
                    if (M.large_side == 1) lw_numflux(FL, Fs_, UL, Us_, uL, us_, o, Fn);
                    else lw_numflux(Fs_, FL, Us_, UL, us_, uL, o, Fn);
                    for (int v = 0; v < NV; ++v) {
                        fstar[half][l][v] = Fn[v];
                        sfv[sidx(v, l, sdir, es)] = Fn[v];
                    }
                }
            }
            for (int l = 0; l < n; ++l)
                for (int v = 0; v < NV; ++v) {
                    T acc = T(0);
                    for (int k = 0; k < n; ++k) acc = acc + basis.rev_up[l][k] * fstar[1][k][v];
                    for (int k = 0; k < n; ++k) acc = acc + basis.rev_lo[l][k] * fstar[0][k][v];
                    sfv[sidx(v, l, ldir, M.large)] = acc;
                }
        }
    }

    // Trixi calc_surface_integral! (weak form, TreeMesh 2-D) followed by
    // apply_jacobian! (du *= -J⁻¹), called at dg_2d.jl:75-80.
    void surface_integral_and_jacobian(T* du) {
        const double f1 = basis.bnd_lo, f2 = basis.bnd_hi;
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) {
            for (int l = 0; l < n; ++l)
                for (int v = 0; v < NV; ++v) {
                    du[idx(v, 0, l, e)] = du[idx(v, 0, l, e)] - sfv[sidx(v, l, 0, e)] * f1;
                    du[idx(v, n - 1, l, e)] = du[idx(v, n - 1, l, e)] + sfv[sidx(v, l, 1, e)] * f2;
                    du[idx(v, l, 0, e)] = du[idx(v, l, 0, e)] - sfv[sidx(v, l, 2, e)] * f1;
                    du[idx(v, l, n - 1, e)] = du[idx(v, l, n - 1, e)] + sfv[sidx(v, l, 3, e)] * f2;
                }
            const double factor = -mesh.inv_jac[e];
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) du[idx(v, i, j, e)] = du[idx(v, i, j, e)] * factor;
        }
    }

    // rhs!(du, u, t, dt, mesh::TreeMesh{2}, ...), dg_2d.jl:28-83, plain FR.
    void reset_du(T* du) const {
        const long long nd = (long long)ndofs_total();
#pragma omp parallel for schedule(static)
        for (long long k = 0; k < nd; ++k) du[k] = T(0);
    }

    virtual void rhs_lw(T* du, const T* u, double t, double dt) {
        src_t = t;
        reset_du(du);  // reset_du!
        // calc_volume_integral! for VolumeIntegralFRShockCapturing (dg_2d.jl:1648-1804)
        if (sc) sc_prepare(u, dt);
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) {
            const double a_e = sc ? alpha[e] : 0.0;
            if (basis.N == 4) lw_volume_kernel_4(du, u, dt, e, 1.0 - a_e);
            else lw_volume_kernel_3(du, u, dt, e, 1.0 - a_e);
            if (sc) sc_element(du, u, e, dt);
        }
        calc_interface_flux(u);
        if (!mesh.bd_elem.empty()) calc_boundary_flux(u, t, dt, 1.0);
        if (!mesh.mortars.empty()) calc_mortar_flux(u);
        surface_integral_and_jacobian(du);
    }

    // rhs_mdrk1! / rhs_mdrk2!, tree/multi_derivative_2d.jl:1-116.  Both stages
    // prolong uⁿ (the integrator's u) to the interfaces; stage 2 reads `us`.
    // shock-capturing prologue of an rhs call: blending factors from uⁿ
    void sc_prepare(const T* u, double dt) {
        if (alpha_override.size() == (size_t)mesh.E) alpha = alpha_override;
        else indicator_hg(u);
        sc_dt = dt;
    }
    // pure_and_blended_element_ids!: alpha ≈ 0 (atol 1e-12) is pure DG (calc_fn_low!
    // only); a blended element adds fv_kernel! of the reconstruction, whose
    // MUSCL-Hancock half-step uses dtmh (0.5 dt in MDRK stage 1,
    // tree/multi_derivative_2d.jl:199)
    void sc_element(T* du, const T* u, int e, double dtmh) {
        const double a_e = alpha[e];
        const bool dg_only = std::fabs(a_e) <= 1e-12;
        if (!dg_only && muscl_hancock) fv_kernel_mh(du, u, e, a_e, dtmh);
        else fv_kernel_fo(du, u, e, a_e, !dg_only);
    }

    // calc_volume_integral_mdrk1!/2! with shock capturing (tree/multi_derivative_2d.jl:151-261):
    // the MDRK kernel scaled by 1 - α plus first-order FV of uⁿ scaled by α.
    virtual void rhs_mdrk1(T* du, const T* u, double t, double dt) {
        src_t = t;
        ensure_mdrk();
        reset_du(du);
        if (sc) sc_prepare(u, dt);
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) {
            mdrk_kernel_1(du, u, dt, e, sc ? 1.0 - alpha[e] : 1.0);
            if (sc) sc_element(du, u, e, 0.5 * dt);
        }
        calc_interface_flux(u);
        if (!mesh.bd_elem.empty()) calc_boundary_flux(u, t, dt, flux_order == 2 ? 1.0 : 0.5);
        if (!mesh.mortars.empty()) calc_mortar_flux(u);
        surface_integral_and_jacobian(du);
    }
    virtual void rhs_mdrk2(T* du, const T* u, double t, double dt) {
        src_t = t;
        ensure_mdrk();
        reset_du(du);
        if (sc) sc_prepare(u, dt);
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) {
            mdrk_kernel_2(du, dt, e, sc ? 1.0 - alpha[e] : 1.0);
            if (sc) sc_element(du, u, e, dt);
        }
        calc_interface_flux(u);
        if (!mesh.bd_elem.empty()) calc_boundary_flux(u, t, dt, 1.0);
        if (!mesh.mortars.empty()) calc_mortar_flux(u);
        surface_integral_and_jacobian(du);
    }

    // my_limiter_zhang_shu! on a conforming TreeMesh (no mortars)
    // (src/callbacks/callbacks_stage/positivity_zhang_shu_dg2d.jl:8-45): where
    // the minimum of the variable over an element's nodes is below the
    // threshold, every node is pulled toward the element mean,
    // u <- θ u + (1-θ) ū with θ = (var(ū) - threshold) / (var(ū) - min var).
    // The mean is Σ u w_i w_j / 4 in the reference's loop order.
    template <class Var> void zhang_shu_pass(T* u, double threshold, Var var) {
        const auto& w = basis.weights;
#pragma omp parallel for schedule(static)
        for (int e = 0; e < mesh.E; ++e) {
            double vmin = std::numeric_limits<double>::infinity();
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) vmin = std::min(vmin, value(var(&u[idx(0, i, j, e)])));
            if (!(vmin < threshold)) continue;
            T mean[NV];
            for (int v = 0; v < NV; ++v) mean[v] = T(0);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v) mean[v] = mean[v] + u[idx(v, i, j, e)] * w[i] * w[j];
            for (int v = 0; v < NV; ++v) mean[v] = mean[v] / 4.0;
            const double vmean = value(var(mean));
            const double theta = (vmean - threshold) / (vmean - vmin);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int v = 0; v < NV; ++v)
                        u[idx(v, i, j, e)] = theta * u[idx(v, i, j, e)] + (1.0 - theta) * mean[v];
        }
    }
    // thresholds and variables are processed in order (positivity_zhang_shu.jl:26-41)
    void positivity_limiter(T* u) {
        zhang_shu_pass(u, pos_threshold[0], [this](const T* q) { return eq.density(q); });
        zhang_shu_pass(u, pos_threshold[1], [this](const T* q) { return eq.pressure(q); });
    }

    // Trixi max_dt(u, t, mesh::TreeMesh{2}, constant_speed::False, ...):
    // 2 / (nnodes * max_e J⁻¹ (max λ1 + max λ2)); called by compute_dt
    // (src/solvers/lwfr.jl:176-190).
    virtual double max_dt(const T* u) const {
        double mss = std::nextafter(0.0, 1.0);
#pragma omp parallel for schedule(static) reduction(max : mss)
        for (int e = 0; e < mesh.E; ++e) {
            double m1 = 0, m2 = 0;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    T l1, l2;
                    eq.max_abs_speeds(&u[idx(0, i, j, e)], l1, l2);
                    m1 = max_(m1, value(l1));
                    m2 = max_(m2, value(l2));
                }
            mss = max_(mss, mesh.inv_jac[e] * (m1 + m2));
        }
        return 2.0 / (n * mss);
    }
};

// Julia isapprox(x, y) with default tolerances for Float64 (atol = 0,
// rtol = sqrt(eps)).  Used by isfinished (src/solvers/lwfr.jl:146-148).
inline bool julia_isapprox(double x, double y) {
    if (x == y) return true;
    if (!std::isfinite(x) || !std::isfinite(y)) return false;
    return std::fabs(x - y) <= 1.4901161193847656e-08 * std::max(std::fabs(x), std::fabs(y));
}

// solve_lwfr loop with CFLBased time stepping (src/solvers/lwfr.jl:76-103,
// 176-190, 212-263; src/solvers/mdrk.jl:1-29).  Callbacks and limiters are
// empty.  Returns the number of steps taken; dts[k] receives the k-th dt.
template <class T, class Eq>
int solve(Solver<T, Eq>& s, T* u, double t0, double tf, double cfl, int max_steps, int mdrk, double* t_out,
          double* dts) {
    const long long nd = (long long)s.ndofs_total();
    std::vector<T> du(nd), uprev(nd);
    double t = t0;
    int it = 0;
    if (s.positivity) s.positivity_limiter(u);  // apply_limiters! before the loop (lwfr.jl:251)
    while (!julia_isapprox(t, tf) && it < max_steps) {
        double dt = cfl * s.max_dt(u);
        if (t + dt > tf) dt = tf - t;  // set_dt!
#pragma omp parallel for schedule(static)
        for (long long k = 0; k < nd; ++k) uprev[k] = u[k];
        if (!mdrk) {
            s.rhs_lw(du.data(), u, t, dt);
#pragma omp parallel for schedule(static)
            for (long long k = 0; k < nd; ++k) u[k] = uprev[k] + dt * du[k];
        } else {
            s.rhs_mdrk1(du.data(), u, t, dt);
#pragma omp parallel for schedule(static)
            for (long long k = 0; k < nd; ++k) s.us[k] = uprev[k] + dt * du[k];
            s.rhs_mdrk2(du.data(), u, t, dt);
#pragma omp parallel for schedule(static)
            for (long long k = 0; k < nd; ++k) u[k] = uprev[k] + dt * du[k];
        }
        // perform_step! applies the limiters after the update (lwfr.jl:226) and the loop
        // again after the callbacks (lwfr.jl:257); the second pass leaves a limited
        // state unchanged.  The MDRK step applies them to uⁿ between the stages
        // (mdrk.jl:19), which was already limited.
        if (s.positivity) s.positivity_limiter(u);
        if (dts) dts[it] = dt;
        t += dt;  // set_t_and_iter!
        ++it;
    }
    if (t_out) *t_out = t;
    return it;
}

// The LW (or MDRK) loop of solve_lwfr with a prescribed dt per step, as the
// gradient checks use it: the CFL dts of a reference run are frozen, so the map
// u0 -> u_K is differentiable (compute_dt's max is not).  MDRK: us = uprev +
// dt rhs_mdrk1!(u), then u = uprev + dt rhs_mdrk2!(u) (mdrk.jl:13-25).
template <class T, class Eq>
void solve_fixed_dt(Solver<T, Eq>& s, T* u, double t0, const double* dts, int nsteps, bool mdrk = false) {
    const long long nd = (long long)s.ndofs_total();
    std::vector<T> du(nd), uprev(nd);
    double t = t0;
    if (s.positivity) s.positivity_limiter(u);
    for (int k = 0; k < nsteps; ++k) {
        const double dt = dts[k];
#pragma omp parallel for schedule(static)
        for (long long q = 0; q < nd; ++q) uprev[q] = u[q];
        if (mdrk) {
            s.rhs_mdrk1(du.data(), u, t, dt);
#pragma omp parallel for schedule(static)
            for (long long q = 0; q < nd; ++q) s.us[q] = uprev[q] + dt * du[q];
            s.rhs_mdrk2(du.data(), u, t, dt);
        } else {
            s.rhs_lw(du.data(), u, t, dt);
        }
#pragma omp parallel for schedule(static)
        for (long long q = 0; q < nd; ++q) u[q] = uprev[q] + dt * du[q];
        if (s.positivity) s.positivity_limiter(u);
        t += dt;
    }
}

// solve_lwfr with time_step_computation = Adaptive() for LW (perform_step!,
// src/solvers/adaptive.jl:95-173; test_updated_solution 30-46;
// load_temporal_errors! 48-64; dt_factor 88-93; controller β = (0.6, -0.2, 0),
// epsilon = (1, 1, 1) at the start, lwfr.jl:162-168).  The temporal errors come
// from lw_volume_kernel_4 with the solver's abstol/reltol.  A rejected attempt
// (non-admissible state or factor <= 0.81) restarts from uⁿ with
// min(factor, 0.95)·dt; a non-finite factor there gives 0.95·dt.  Returns the
// number of accepted steps; dts[k] receives the k-th accepted dt, *nreject the
// rejected attempts.
template <class T, class Eq>
int solve_adaptive(Solver<T, Eq>& s, T* u, double t0, double tf, double dt_initial, int max_attempts, double* t_out,
                   double* dts, long long* nreject) {
    const long long nd = (long long)s.ndofs_total();
    const double ndofs = (double)s.mesh.E * s.n * s.n;  // Trixi.ndofs
    const double k = s.n;                                // nnodes
    std::vector<T> du(nd), uprev(nd);
    double eps[3] = {1.0, 1.0, 1.0};  // epsilon[-1], epsilon[0], epsilon[1]
    double t = t0;
    double dt = dt_initial;
    if (t + dt > tf) dt = tf - t;  // set_dt!
    if (s.positivity) s.positivity_limiter(u);
    int it = 0, attempts = 0, nfail = 0;
    long long nrej = 0;
    while (!julia_isapprox(t, tf) && attempts < max_attempts) {
        ++attempts;
#pragma omp parallel for schedule(static)
        for (long long q = 0; q < nd; ++q) uprev[q] = u[q];
        s.rhs_lw(du.data(), u, t, dt);
#pragma omp parallel for schedule(static)
        for (long long q = 0; q < nd; ++q) u[q] = uprev[q] + dt * du[q];
        if (s.positivity) s.positivity_limiter(u);
        bool valid = true;
        for (int e = 0; e < s.mesh.E && valid; ++e)
            for (int q = 0; q < s.n * s.n && valid; ++q) valid = s.eq.is_admissible(&u[(size_t)(e * s.n * s.n + q) * Eq::NV]);
        double total = 0.0;
        for (int e = 0; e < s.mesh.E; ++e) total += value(s.temporal_errors[e]);
        const double err = std::sqrt(total / ndofs);
        if (nfail == 0) {
            eps[0] = eps[1];
            eps[1] = eps[2];
        }
        eps[2] = 1.0 / (err + 1e-12);
        const double fac = std::pow(eps[2], 0.6 / k) * std::pow(eps[1], -0.2 / k) * std::pow(eps[0], 0.0 / k);
        const double factor = 1.0 + std::atan(fac - 1.0);
        if (!valid || !(factor > 0.81)) {
            ++nfail;
            dt = (factor < 0.95 ? factor : 0.95) * dt;
            if (t + dt > tf) dt = tf - t;
#pragma omp parallel for schedule(static)
            for (long long q = 0; q < nd; ++q) u[q] = uprev[q];
            continue;
        }
        nrej += nfail;
        nfail = 0;
        double dtn = (factor < 1.5 ? factor : 1.5) * dt;
        double dts_ = dt;
        if (t + dts_ > tf) dts_ = tf - t;  // set_t_and_iter!
        if (dts) dts[it] = dts_;
        t += dts_;
        ++it;
        if (t + dtn > tf) dtn = tf - t;  // set_dt!
        dt = dtn;
    }
    if (t_out) *t_out = t;
    if (nreject) *nreject = nrej;
    return it;
}

}  // namespace tlwo
