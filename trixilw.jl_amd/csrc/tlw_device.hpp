// tlw_device.hpp -- gfx950 kernels of the TreeMesh Lax-Wendroff / MDRK step.
//
// Work decomposition: one thread per GLL node, EB elements per 128-thread
// workgroup (N=4: 5 elements = 125 lanes; N=3: 8 elements = 128 lanes).  Each
// element's fluxes and flux time derivatives are exchanged through LDS for the
// derivative-matrix contractions; everything else stays in registers.
//
// One step is two kernels:
//   volume   uⁿ -> approximate LW procedure (Taylor states, finite-difference
//            flux time derivatives, Ũ, F̃), volume term D̂F̃, partial update of
//            every node, and a compact face record per element face node
//   surface  Rusanov flux on the time-averaged F̃/Ũ from the two face records,
//            FR correction of the face nodes, finalisation of the update,
//            CFL reduction of the new state and (last workgroup) dt, t update
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

namespace tlw {

constexpr int kBlock = 128;
constexpr int kTraceQ = 10;  // face record: jump state (4), normal flux (4), |v_n|, c

template <int N> struct Geo {
    static constexpr int n = N + 1;
    static constexpr int NN = n * n;
    static constexpr int EB = kBlock / NN;  // elements per workgroup
    static constexpr int S = EB * NN;       // LDS stride of one variable
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
};

// Device-resident integrator state (LWIntegrator fields t, dt, iter and the
// isfinished test), updated by the last workgroup of each surface kernel.
struct DevState {
    double t, dt, t_final, cfl;
    double query;                 // Trixi max_dt result of a query reduction
    int finished;
    int cur;                      // state buffer holding uⁿ
    int bad_state;                // set when a non-admissible state was met
    long long iter;
};

enum BufSel { SEL_EXPLICIT = 0, SEL_CUR = 1, SEL_OTHER = 2 };
enum Reduce { RED_NONE = 0, RED_STEP = 1, RED_INIT = 2, RED_QUERY = 3 };

struct KArgs {
    int E;
    const double* u;      // uⁿ (explicit)
    const double* us;     // MDRK stage state
    double* out;          // volume: partial result; surface: finalised in place
    double* bufs[2];      // ping-pong state buffers for the device-resident loop
    int in_sel, out_sel;  // BufSel
    double* trace;        // [E][4][n][kTraceQ]
    const int* nbr;       // [E][4]
    const double* inv_jac;
    const double* center;  // [E][2]
    const double* length;  // [E]
    const DevConst* c;
    DevState* st;
    double* auxU;  // optional full Ũ
    double* auxF;  // optional full F̃
    double* partial;  // per-workgroup max of J⁻¹(max λ1 + max λ2) for the CFL reduction
    double t_arg, dt_arg;
    int use_state_dt;  // 1: t, dt from DevState (device loop); 0: from t_arg/dt_arg
    int step_mode;     // 1: out = u + dt * rhs; 0: out = rhs
    int reduce;        // Reduce
    double bc_scale;   // scaling_factor of the Dirichlet time average
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

__device__ __forceinline__ void euler_flux_dir(const double* u, double gm1, int orientation, double* f) {
    double ff[4], gg[4];
    euler_fluxes(u, gm1, ff, gg);
    for (int v = 0; v < 4; ++v) f[v] = orientation == 1 ? ff[v] : gg[v];
}

// |v_n| and the sound speed: the per-side ingredients of Trixi's
// max_abs_speed_naive (λ = max(|v_n,L|,|v_n,R|) + max(c_L, c_R)).
__device__ __forceinline__ void speed_parts(const double* s, double gamma, int orientation, double& a, double& c) {
    const double rho = s[0];
    const double v1 = s[1] / rho, v2 = s[2] / rho;
    const double p = (gamma - 1.0) * (s[3] - 0.5 * (s[1] * v1 + s[2] * v2));
    a = fabs(orientation == 1 ? v1 : v2);
    c = sqrt(gamma * p / rho);
}

__device__ __forceinline__ double dmax(double x, double y) { return (y > x) ? y : x; }

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

// Derivative contraction of one element: out[v] = Σ_k Di[k] s_f[v][j,k] + Σ_k Dj[k] s_g[v][k,i]
// where s holds f (variables 0..3) and g (4..7) of every node of the workgroup.
template <int N>
__device__ __forceinline__ void contract(const double* s, int base, int i, int j, const double* Di, const double* Dj,
                                         double* out) {
    constexpr int n = Geo<N>::n, S = Geo<N>::S;
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

template <int N> __device__ __forceinline__ void put8(double* s, int idx, const double* f, const double* g) {
    constexpr int S = Geo<N>::S;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        s[v * S + idx] = f[v];
        s[(4 + v) * S + idx] = g[v];
    }
}

// Face record of one element face node.
__device__ __forceinline__ void write_record(double* rec, const double* jump, const double* Fn, const double* lam_state,
                                             double gamma, int orientation) {
    double a, c;
    speed_parts(lam_state, gamma, orientation, a, c);
    double2* r2 = reinterpret_cast<double2*>(rec);
    r2[0] = make_double2(jump[0], jump[1]);
    r2[1] = make_double2(jump[2], jump[3]);
    r2[2] = make_double2(Fn[0], Fn[1]);
    r2[3] = make_double2(Fn[2], Fn[3]);
    r2[4] = make_double2(a, c);
}

// Writes the face records of node (i, j) (jump state, normal component of
// F̃, λ ingredients): the data prolong2interfaces!/prolong2boundaries! copy.
template <int N>
__device__ __forceinline__ void write_traces(const KArgs& a, int e, int i, int j, const double* uN, const double* Uc,
                                             const double* F, const double* G) {
    constexpr int n = N + 1;
    const bool tree = a.c->flux_order == 0;
    const double* jump = tree ? uN : Uc;
    const double* lam = tree ? Uc : uN;
    const double gamma = a.c->gamma;
    double* tr = a.trace + (size_t)e * 4 * n * kTraceQ;
    if (i == 0) write_record(tr + (0 * n + j) * kTraceQ, jump, F, lam, gamma, 1);
    if (i == n - 1) write_record(tr + (1 * n + j) * kTraceQ, jump, F, lam, gamma, 1);
    if (j == 0) write_record(tr + (2 * n + i) * kTraceQ, jump, G, lam, gamma, 2);
    if (j == n - 1) write_record(tr + (3 * n + i) * kTraceQ, jump, G, lam, gamma, 2);
}

__device__ __forceinline__ void load4(const double* p, double* u) {
    const double2* p2 = reinterpret_cast<const double2*>(p);
    const double2 a = p2[0], b = p2[1];
    u[0] = a.x; u[1] = a.y; u[2] = b.x; u[3] = b.y;
}
__device__ __forceinline__ void store4(double* p, const double* u) {
    double2* p2 = reinterpret_cast<double2*>(p);
    p2[0] = make_double2(u[0], u[1]);
    p2[1] = make_double2(u[2], u[3]);
}

__device__ __forceinline__ const double* sel_in(const KArgs& a, int cur) {
    return a.in_sel == SEL_CUR ? a.bufs[cur] : (a.in_sel == SEL_OTHER ? a.bufs[1 - cur] : a.u);
}
__device__ __forceinline__ double* sel_out(const KArgs& a, int cur) {
    return a.out_sel == SEL_CUR ? a.bufs[cur] : (a.out_sel == SEL_OTHER ? a.bufs[1 - cur] : a.out);
}

// Writes the volume result: rhs mode out = -J⁻¹ vol; step mode out = u + dt (-J⁻¹ vol).
__device__ __forceinline__ void emit_volume(const KArgs& a, double* out, size_t off, const double* uN, const double* vol,
                                            double invJ, double dt) {
    double r[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double d = (-invJ) * vol[v];
        r[v] = a.step_mode ? fma(dt, d, uN[v]) : d;
    }
    store4(out + off, r);
}

template <int N>
__device__ __forceinline__ void load_drows(const DevConst* c, int i, int j, double scale, double* Di, double* Dj) {
    constexpr int n = N + 1;
#pragma unroll
    for (int k = 0; k < n; ++k) {
        Di[k] = scale * c->D[i][k];
        Dj[k] = scale * c->D[j][k];
    }
}

// Shared prologue of the element kernels.
#define TLW_ELEMENT_PROLOGUE(N)                                  \
    constexpr int n = Geo<N>::n, NN = Geo<N>::NN, EB = Geo<N>::EB; \
    const int tid = threadIdx.x;                                   \
    const int le = tid / NN, node = tid - (tid / NN) * NN;         \
    const int i = node % n, j = node / n;                          \
    const int e = blockIdx.x * EB + le;                            \
    const bool act = (le < EB) && (e < a.E);                       \
    const int ec = e < a.E ? e : a.E - 1;                          \
    const int lec = le < EB ? le : EB - 1;                         \
    const int sidx = lec * NN + node;                              \
    const int sbase = lec * NN;                                    \
    (void)sidx; (void)sbase; (void)act;

// ---------------------------------------------------------------------------
// LW volume kernel, N = 4: lw_volume_kernel_4!, dg_2d.jl:918-1267 (the
// temporal-error estimate at 1183-1199, 1251-1265 is an adaptive-mode output
// and is not computed).
template <int N>
__global__ __launch_bounds__(kBlock) void k_volume_lw4(KArgs a) {
    TLW_ELEMENT_PROLOGUE(N)
    static_assert(N == 4, "N = 4 kernel");
    __shared__ double s0[8 * Geo<N>::S], s1[8 * Geo<N>::S], s2[8 * Geo<N>::S];
    DevState* st = a.st;
    if (a.use_state_dt && st->finished) return;
    const int cur = a.use_state_dt ? st->cur : 0;
    const double* uin = sel_in(a, cur);
    double* out = sel_out(a, cur);
    const double dt = a.use_state_dt ? st->dt : a.dt_arg;
    const DevConst* c = a.c;
    const double gm1 = c->gamma - 1.0;
    const double invJ = a.inv_jac[ec];
    const size_t off = ((size_t)ec * NN + node) * 4;

    double Di[n], Dj[n];
    load_drows<N>(c, i, j, -dt, Di, Dj);
    double u[4];
    load4(uin + off, u);
    double F[4], G[4];
    euler_fluxes(u, gm1, F, G);
    if (act) put8<N>(s0, sidx, F, G);
    __syncthreads();
    double ut[4];
    contract<N>(s0, sbase, i, j, Di, Dj, ut);
    double Uc[4], up[4], um[4], upp[4], umm[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        ut[v] *= invJ;
        Uc[v] = fma(0.5, ut[v], u[v]);
        up[v] = u[v] + ut[v];
        um[v] = u[v] - ut[v];
        umm[v] = fma(-2.0, ut[v], u[v]);
        upp[v] = fma(2.0, ut[v], u[v]);
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
    if (act) put8<N>(s1, sidx, dx, dy);
    __syncthreads();
    double uk[4];
    contract<N>(s1, sbase, i, j, Di, Dj, uk);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double w = uk[v] * invJ;  // utt
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
        dx[v] = fma(-30.0, s0[v * Geo<N>::S + sidx], dx[v]);
        dy[v] = fma(-30.0, s0[(4 + v) * Geo<N>::S + sidx], dy[v]);
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
    if (act) put8<N>(s2, sidx, dx, dy);
    __syncthreads();
    contract<N>(s2, sbase, i, j, Di, Dj, uk);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double w = uk[v] * invJ;  // uttt
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
    if (act) put8<N>(s1, sidx, dx, dy);
    __syncthreads();
    contract<N>(s1, sbase, i, j, Di, Dj, uk);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double w = uk[v] * invJ;  // utttt
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
        dx[v] = fma(6.0, s0[v * Geo<N>::S + sidx], dx[v]);
        dy[v] = fma(6.0, s0[(4 + v) * Geo<N>::S + sidx], dy[v]);
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
    if (act) put8<N>(s2, sidx, F, G);
    __syncthreads();
    double Hi[n], Hj[n];
#pragma unroll
    for (int k = 0; k < n; ++k) { Hi[k] = c->Dhat[i][k]; Hj[k] = c->Dhat[j][k]; }
    double vol[4];
    contract<N>(s2, sbase, i, j, Hi, Hj, vol);
    if (act) {
        emit_volume(a, out, off, u, vol, invJ, dt);
        write_traces<N>(a, e, i, j, u, Uc, F, G);
        if (a.auxU) {
            store4(a.auxU + off, Uc);
            store4(a.auxF + ((size_t)e * NN + node) * 8, F);
            store4(a.auxF + ((size_t)e * NN + node) * 8 + 4, G);
        }
    }
}

// LW volume kernel, N = 3: lw_volume_kernel_3!, dg_2d.jl:674-916.
template <int N>
__global__ __launch_bounds__(kBlock) void k_volume_lw3(KArgs a) {
    TLW_ELEMENT_PROLOGUE(N)
    static_assert(N == 3, "N = 3 kernel");
    __shared__ double s0[8 * Geo<N>::S], s1[8 * Geo<N>::S], s2[8 * Geo<N>::S];
    DevState* st = a.st;
    if (a.use_state_dt && st->finished) return;
    const int cur = a.use_state_dt ? st->cur : 0;
    const double* uin = sel_in(a, cur);
    double* out = sel_out(a, cur);
    const double dt = a.use_state_dt ? st->dt : a.dt_arg;
    const DevConst* c = a.c;
    const double gm1 = c->gamma - 1.0;
    const double invJ = a.inv_jac[ec];
    const size_t off = ((size_t)ec * NN + node) * 4;

    double Di[n], Dj[n];
    load_drows<N>(c, i, j, -dt, Di, Dj);
    double u[4];
    load4(uin + off, u);
    double F[4], G[4];
    euler_fluxes(u, gm1, F, G);
    if (act) put8<N>(s0, sidx, F, G);
    __syncthreads();
    double ut[4];
    contract<N>(s0, sbase, i, j, Di, Dj, ut);
    double Uc[4], up[4], um[4], upp[4], umm[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        ut[v] *= invJ;
        Uc[v] = fma(0.5, ut[v], u[v]);
        up[v] = u[v] + ut[v];
        um[v] = u[v] - ut[v];
        umm[v] = fma(-2.0, ut[v], u[v]);
        upp[v] = fma(2.0, ut[v], u[v]);
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
    if (act) put8<N>(s1, sidx, dx, dy);
    __syncthreads();
    double uk[4];
    contract<N>(s1, sbase, i, j, Di, Dj, uk);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double w = uk[v] * invJ;  // utt
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
        dx[v] = fma(-2.0, s0[v * Geo<N>::S + sidx], fa[v]);
        dy[v] = fma(-2.0, s0[(4 + v) * Geo<N>::S + sidx], ga[v]);
    }
    euler_fluxes(um, gm1, fa, ga);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        dx[v] = dx[v] + fa[v];
        dy[v] = dy[v] + ga[v];
        F[v] = fma(1.0 / 6.0, dx[v], F[v]);
        G[v] = fma(1.0 / 6.0, dy[v], G[v]);
    }
    if (act) put8<N>(s2, sidx, dx, dy);
    __syncthreads();
    contract<N>(s2, sbase, i, j, Di, Dj, uk);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double w = uk[v] * invJ;  // uttt
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
    if (act) put8<N>(s1, sidx, F, G);
    __syncthreads();
    double Hi[n], Hj[n];
#pragma unroll
    for (int k = 0; k < n; ++k) { Hi[k] = c->Dhat[i][k]; Hj[k] = c->Dhat[j][k]; }
    double vol[4];
    contract<N>(s1, sbase, i, j, Hi, Hj, vol);
    if (act) {
        emit_volume(a, out, off, u, vol, invJ, dt);
        write_traces<N>(a, e, i, j, u, Uc, F, G);
        if (a.auxU) {
            store4(a.auxU + off, Uc);
            store4(a.auxF + ((size_t)e * NN + node) * 8, F);
            store4(a.auxF + ((size_t)e * NN + node) * 8 + 4, G);
        }
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

// MDRK stage 1 volume kernel: mdrk_kernel_1!, tree/multi_derivative_2d.jl:263-420.
// F̃ = f/2 + f_t/8, Ũ = u/2 + ut/8; F2 = f + f_t/6 and U2 = u + ut/6 are
// recomputed by the stage-2 kernel instead of being stored.
template <int N>
__global__ __launch_bounds__(kBlock) void k_volume_mdrk1(KArgs a) {
    TLW_ELEMENT_PROLOGUE(N)
    __shared__ double s0[8 * Geo<N>::S], s1[8 * Geo<N>::S];
    DevState* st = a.st;
    if (a.use_state_dt && st->finished) return;
    const int cur = a.use_state_dt ? st->cur : 0;
    const double* uin = sel_in(a, cur);
    double* out = sel_out(a, cur);
    const double dt = a.use_state_dt ? st->dt : a.dt_arg;
    const DevConst* c = a.c;
    const double gm1 = c->gamma - 1.0;
    const double invJ = a.inv_jac[ec];
    const size_t off = ((size_t)ec * NN + node) * 4;

    double Di[n], Dj[n];
    load_drows<N>(c, i, j, -dt, Di, Dj);
    double u[4], f[4], g[4];
    load4(uin + off, u);
    euler_fluxes(u, gm1, f, g);
    if (act) put8<N>(s0, sidx, f, g);
    __syncthreads();
    double ut[4];
    contract<N>(s0, sbase, i, j, Di, Dj, ut);
#pragma unroll
    for (int v = 0; v < 4; ++v) ut[v] *= invJ;
    double dx[4], dy[4], F[4], G[4], Uc[4];
    mdrk_ft(u, ut, gm1, dx, dy);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F[v] = fma(0.125, dx[v], 0.5 * f[v]);
        G[v] = fma(0.125, dy[v], 0.5 * g[v]);
        Uc[v] = fma(0.125, ut[v], 0.5 * u[v]);
    }
    if (act) put8<N>(s1, sidx, F, G);
    __syncthreads();
    double Hi[n], Hj[n];
#pragma unroll
    for (int k = 0; k < n; ++k) { Hi[k] = c->Dhat[i][k]; Hj[k] = c->Dhat[j][k]; }
    double vol[4];
    contract<N>(s1, sbase, i, j, Hi, Hj, vol);
    if (act) {
        emit_volume(a, out, off, u, vol, invJ, dt);
        write_traces<N>(a, e, i, j, u, Uc, F, G);
        if (a.auxU) {
            store4(a.auxU + off, Uc);
            store4(a.auxF + ((size_t)e * NN + node) * 8, F);
            store4(a.auxF + ((size_t)e * NN + node) * 8 + 4, G);
        }
    }
}

// MDRK stage 2 volume kernel: mdrk_kernel_2!, tree/multi_derivative_2d.jl:423-555.
// F̃ = F2 + f_t(us)/3, Ũ = U2 + ust/3, with F2, U2 recomputed from uⁿ.
template <int N>
__global__ __launch_bounds__(kBlock) void k_volume_mdrk2(KArgs a) {
    TLW_ELEMENT_PROLOGUE(N)
    __shared__ double s0[8 * Geo<N>::S], s1[8 * Geo<N>::S], s2[8 * Geo<N>::S];
    DevState* st = a.st;
    if (a.use_state_dt && st->finished) return;
    const int cur = a.use_state_dt ? st->cur : 0;
    const double* uin = sel_in(a, cur);
    double* out = sel_out(a, cur);
    const double dt = a.use_state_dt ? st->dt : a.dt_arg;
    const DevConst* c = a.c;
    const double gm1 = c->gamma - 1.0;
    const double invJ = a.inv_jac[ec];
    const size_t off = ((size_t)ec * NN + node) * 4;

    double Di[n], Dj[n];
    load_drows<N>(c, i, j, -dt, Di, Dj);
    double u[4], f[4], g[4], usn[4], fs[4], gs[4];
    load4(uin + off, u);
    load4(a.us + off, usn);
    euler_fluxes(u, gm1, f, g);
    euler_fluxes(usn, gm1, fs, gs);
    if (act) {
        put8<N>(s0, sidx, f, g);
        put8<N>(s1, sidx, fs, gs);
    }
    __syncthreads();
    double ut[4], ust[4];
    contract<N>(s0, sbase, i, j, Di, Dj, ut);
    contract<N>(s1, sbase, i, j, Di, Dj, ust);
#pragma unroll
    for (int v = 0; v < 4; ++v) { ut[v] *= invJ; ust[v] *= invJ; }
    double dx[4], dy[4], F[4], G[4], Uc[4];
    mdrk_ft(u, ut, gm1, dx, dy);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F[v] = fma(1.0 / 6.0, dx[v], f[v]);   // F2
        G[v] = fma(1.0 / 6.0, dy[v], g[v]);
        Uc[v] = fma(1.0 / 6.0, ut[v], u[v]);  // U2
    }
    mdrk_ft(usn, ust, gm1, dx, dy);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F[v] = fma(1.0 / 3.0, dx[v], F[v]);
        G[v] = fma(1.0 / 3.0, dy[v], G[v]);
        Uc[v] = fma(1.0 / 3.0, ust[v], Uc[v]);
    }
    if (act) put8<N>(s2, sidx, F, G);
    __syncthreads();
    double Hi[n], Hj[n];
#pragma unroll
    for (int k = 0; k < n; ++k) { Hi[k] = c->Dhat[i][k]; Hj[k] = c->Dhat[j][k]; }
    double vol[4];
    contract<N>(s2, sbase, i, j, Hi, Hj, vol);
    if (act) {
        emit_volume(a, out, off, u, vol, invJ, dt);
        write_traces<N>(a, e, i, j, u, Uc, F, G);
        if (a.auxU) {
            store4(a.auxU + off, Uc);
            store4(a.auxF + ((size_t)e * NN + node) * 8, F);
            store4(a.auxF + ((size_t)e * NN + node) * 8 + 4, G);
        }
    }
}

// ---------------------------------------------------------------------------
// Rusanov flux of one face node from the two face records (left = -side).
__device__ __forceinline__ void face_numflux(const double* L, const double* R, double* Fs) {
    const double lam = dmax(L[8], R[8]) + dmax(L[9], R[9]);
#pragma unroll
    for (int v = 0; v < 4; ++v) Fs[v] = 0.5 * (L[4 + v] + R[4 + v]) + (-0.5 * lam) * (R[v] - L[v]);
}

__device__ __forceinline__ void load_record(const double* p, double* r) {
    const double2* p2 = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double2 x = p2[k];
        r[2 * k] = x.x;
        r[2 * k + 1] = x.y;
    }
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
        boundary_value(c, x, y, ts, ub);
        euler_flux_dir(ub, gm1, orientation, fb);
        for (int v = 0; v < 4; ++v) {
            Uo[v] += 0.5 * sf * ub[v] * c->weights[q];
            Fo[v] += 0.5 * sf * fb[v] * c->weights[q];
        }
    }
    boundary_value(c, x, y, t, uo);
    const bool tree = c->flux_order == 0;
    const double* jump = tree ? uo : Uo;
    const double* lam = tree ? Uo : uo;
    double a_, c_;
    speed_parts(lam, c->gamma, orientation, a_, c_);
    for (int v = 0; v < 4; ++v) { r[v] = jump[v]; r[4 + v] = Fo[v]; }
    r[8] = a_;
    r[9] = c_;
}

template <int N, bool BC>
__device__ __forceinline__ void face_flux(const KArgs& a, int e, int face, int l, double t, double dt, double* Fs) {
    constexpr int n = N + 1;
    double own[kTraceQ], oth[kTraceQ];
    load_record(a.trace + ((size_t)e * 4 * n + face * n + l) * kTraceQ, own);
    const int ne = a.nbr[4 * e + face];
    if (!BC || ne >= 0) load_record(a.trace + ((size_t)ne * 4 * n + (face ^ 1) * n + l) * kTraceQ, oth);
    else boundary_record<N>(a, e, face, l, t, dt, oth);
    if (face & 1) face_numflux(own, oth, Fs);  // +x, +y face: own side is left
    else face_numflux(oth, own, Fs);
}

// Final reduction of the per-workgroup partial maxima of J⁻¹(max λ1 + max λ2)
// by one workgroup, then compute_dt / set_dt! / set_t_and_iter! / isfinished
// on the device state (lwfr.jl:76-103, 146-148, 176-190; Trixi max_dt).
constexpr int kFinalizeBlock = 1024;
template <int N>
__global__ __launch_bounds__(kFinalizeBlock) void k_finalize(DevState* st, const double* partial, int nparts,
                                                           int reduce) {
    constexpr int n = N + 1;
    __shared__ double s[kFinalizeBlock];
    if (reduce == RED_STEP && st->finished) return;
    // eight independent accumulators keep several loads in flight per lane
    double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int k = threadIdx.x;
    for (; k + 7 * kFinalizeBlock < nparts; k += 8 * kFinalizeBlock) {
#pragma unroll
        for (int q = 0; q < 8; ++q) m[q] = dmax(m[q], partial[k + q * kFinalizeBlock]);
    }
    for (; k < nparts; k += kFinalizeBlock) m[0] = dmax(m[0], partial[k]);
#pragma unroll
    for (int q = 1; q < 8; ++q) m[0] = dmax(m[0], m[q]);
    s[threadIdx.x] = m[0];
    __syncthreads();
    for (int w = kFinalizeBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = dmax(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double tiny = 4.9406564584124654e-324;  // nextfloat(0.0), Trixi's start value
    const double M = dmax(tiny, s[0]);
    const double max_dt = 2.0 / (n * M);
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
    if (t + dt > tf) dt = tf - t;  // set_dt!
    st->dt = dt;
    st->finished = julia_isapprox(t, tf);
}

// Per-element max of λ1 = |v1| + c and λ2 = |v2| + c (Trixi max_abs_speeds),
// scaled max over the workgroup.
template <int N>
__device__ __forceinline__ double block_speed_max(const KArgs& a, const double* un, bool act, int lec, int node,
                                                  int ec, double* s_l1, double* s_l2, double* s_m) {
    constexpr int NN = Geo<N>::NN, EB = Geo<N>::EB;
    const double gamma = a.c->gamma;
    const double rho = un[0], v1 = un[1] / rho, v2 = un[2] / rho;
    const double p = (gamma - 1.0) * (un[3] - 0.5 * (un[1] * v1 + un[2] * v2));
    const bool bad = !(rho > 0.0) || !(p > 0.0);
    const double cs = sqrt(gamma * p / rho);
    if (act && bad) a.st->bad_state = 1;
    if (act) {
        s_l1[lec * NN + node] = fabs(v1) + cs;
        s_l2[lec * NN + node] = fabs(v2) + cs;
    }
    __syncthreads();
    if (act && node == 0) {
        double m1 = 0.0, m2 = 0.0;
        for (int k = 0; k < NN; ++k) {
            m1 = dmax(m1, s_l1[lec * NN + k]);
            m2 = dmax(m2, s_l2[lec * NN + k]);
        }
        s_m[lec] = a.inv_jac[ec] * (m1 + m2);
    }
    __syncthreads();
    double bm = 0.0;
    if (threadIdx.x == 0) {
        const int nel = min(EB, a.E - (int)blockIdx.x * EB);
        for (int k = 0; k < nel; ++k) bm = dmax(bm, s_m[k]);
    }
    return bm;
}

// Surface kernel: Rusanov flux at every face node of the element (each
// interface is evaluated by both neighbours from identical inputs, so both
// see the same bits), weak-form surface integral with ℓ̂ = 1/w at the end
// nodes, Jacobian, finalisation of the update, CFL reduction.
template <int N, bool BC>
__global__ __launch_bounds__(kBlock) void k_surface(KArgs a) {
    TLW_ELEMENT_PROLOGUE(N)
    __shared__ double s_l1[Geo<N>::S], s_l2[Geo<N>::S], s_m[Geo<N>::EB];
    DevState* st = a.st;
    if (a.use_state_dt && st->finished) return;
    const int cur = a.use_state_dt ? st->cur : 0;
    double* out = sel_out(a, cur);
    const double dt = a.use_state_dt ? st->dt : a.dt_arg;
    const double t = a.use_state_dt ? st->t : a.t_arg;
    const DevConst* c = a.c;
    const double invJ = a.inv_jac[ec];
    const size_t off = ((size_t)ec * NN + node) * 4;
    double val[4];
    load4(out + off, val);
    if (act && (i == 0 || i == n - 1 || j == 0 || j == n - 1)) {
        double surf[4] = {0.0, 0.0, 0.0, 0.0}, Fs[4];
        if (i == 0) {
            face_flux<N, BC>(a, e, 0, j, t, dt, Fs);
#pragma unroll
            for (int v = 0; v < 4; ++v) surf[v] -= Fs[v] * c->bnd_lo;
        }
        if (i == n - 1) {
            face_flux<N, BC>(a, e, 1, j, t, dt, Fs);
#pragma unroll
            for (int v = 0; v < 4; ++v) surf[v] += Fs[v] * c->bnd_hi;
        }
        if (j == 0) {
            face_flux<N, BC>(a, e, 2, i, t, dt, Fs);
#pragma unroll
            for (int v = 0; v < 4; ++v) surf[v] -= Fs[v] * c->bnd_lo;
        }
        if (j == n - 1) {
            face_flux<N, BC>(a, e, 3, i, t, dt, Fs);
#pragma unroll
            for (int v = 0; v < 4; ++v) surf[v] += Fs[v] * c->bnd_hi;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double d = (-invJ) * surf[v];
            val[v] = a.step_mode ? fma(dt, d, val[v]) : val[v] + d;
        }
        store4(out + off, val);
    }
    if (a.reduce != RED_NONE) {
        const double bm = block_speed_max<N>(a, val, act, lec, node, ec, s_l1, s_l2, s_m);
        if (threadIdx.x == 0) a.partial[blockIdx.x] = bm;
    }
}

// CFL reduction over a state (Trixi max_dt) without an update: first dt of a
// solve (RED_INIT) or a max_dt query (RED_QUERY).
template <int N>
__global__ __launch_bounds__(kBlock) void k_max_speed(KArgs a) {
    TLW_ELEMENT_PROLOGUE(N)
    __shared__ double s_l1[Geo<N>::S], s_l2[Geo<N>::S], s_m[Geo<N>::EB];
    const int cur = a.st->cur;
    const double* uin = sel_in(a, cur);
    const size_t off = ((size_t)ec * NN + node) * 4;
    double val[4];
    load4(uin + off, val);
    const double bm = block_speed_max<N>(a, val, act, lec, node, ec, s_l1, s_l2, s_m);
    if (threadIdx.x == 0) a.partial[blockIdx.x] = bm;
}

// Initialisation of the device integrator state (one thread).
__global__ void k_init_state(DevState* st, double t0, double tf, double cfl, int cur) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->t = t0;
        st->dt = 0.0;
        st->t_final = tf;
        st->cfl = cfl;
        st->query = 0.0;
        st->finished = julia_isapprox(t0, tf);
        st->cur = cur;
        st->bad_state = 0;
        st->iter = 0;
    }
}

__global__ void k_clear_bad_state(DevState* st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) st->bad_state = 0;
}

}  // namespace tlw
