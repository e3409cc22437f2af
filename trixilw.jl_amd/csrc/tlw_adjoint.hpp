// tlw_adjoint.hpp -- discrete adjoint of one LW step on a periodic TreeMesh
// (plain FR, TreeMesh flux order):  u' = u + dt R(u)  =>  λ = λ' + dt (∂R/∂u)ᵀ λ'.
//
// R is the reference's LW residual (rhs!, src/solvers/dgsem_tree/dg_2d.jl:28-83):
//   R = -J⁻¹ [ D̂·(F̃, G̃) + lifted Rusanov fluxes F* ],
//   F* = ½(F̃n_L + F̃n_R) − ½ λ(Ũ_L, Ũ_R) (u_R − u_L),  λ = max|v_n| + max c,
// with Ũ, F̃ from the approximate LW procedure (lw_volume_kernel_4!/3!,
// dg_2d.jl:674-1267).  The adjoint reverses that chain by hand: the transposed
// D̂ and D contractions through LDS, the transpose of the Euler flux Jacobian at
// each Taylor state (17 states at N=4, 11 at N=3), and the interface flux with
// the derivative of the max going to the side the forward max picked.
// Two kernels per step: k_adj_records recomputes the forward element
// computation and writes every face node's (u, |v_n|(Ũ), c(Ũ), s̄) with
// s̄ = -J⁻¹ dt λ'; k_adjoint recomputes it again, keeping the time
// derivatives, and runs the reverse sweep.
#pragma once

#include "tlw_device.hpp"

namespace tlw {

constexpr int kAdjQ = 10;  // adjoint face record: u 4, |v_n| 1, c 1, s̄ 4

// Taylor states of the FD flux derivatives: w = u + Σ_d a[kind][d] u^(d+1) for
// kind = u+2, u+1, u-1, u-2 (the incremental updates of lw4_element/lw3_element)
struct LWTaylor {
    static constexpr double a[4][4] = {{2.0, 2.0, 4.0 / 3.0, 2.0 / 3.0},
                                       {1.0, 0.5, 1.0 / 6.0, 1.0 / 24.0},
                                       {-1.0, 0.5, -1.0 / 6.0, 1.0 / 24.0},
                                       {-2.0, 2.0, -4.0 / 3.0, 2.0 / 3.0}};
    // Ũ = u + Σ_d w[d] u^(d+1);  F̃ = F0 + Σ_k w[k] dx_k
    static constexpr double w[4] = {0.5, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0};
};
// FD stencil of stage k: dx_k = Σ_kind c[k][kind] f(w_kind) + c[k][4] F0
template <int N> struct LWStencil;
template <> struct LWStencil<4> {
    static constexpr double c[4][5] = {{-1.0 / 12.0, 8.0 / 12.0, -8.0 / 12.0, 1.0 / 12.0, 0.0},
                                       {-1.0 / 12.0, 16.0 / 12.0, 16.0 / 12.0, -1.0 / 12.0, -30.0 / 12.0},
                                       {0.5, -1.0, 1.0, -0.5, 0.0},
                                       {0.5, -2.0, -2.0, 0.5, 3.0}};
};
template <> struct LWStencil<3> {
    static constexpr double c[3][5] = {{-1.0 / 12.0, 8.0 / 12.0, -8.0 / 12.0, 1.0 / 12.0, 0.0},
                                       {0.0, 1.0, 1.0, 0.0, -2.0},
                                       {0.5, -1.0, 1.0, -0.5, 0.0}};
};

// wb = (∂f/∂w)ᵀ fb + (∂g/∂w)ᵀ gb for the Euler fluxes f, g of euler_fluxes.
__device__ __forceinline__ void euler_fluxes_vjp(const double* w, double gm1, const double* fb, const double* gb,
                                                 double* wb) {
    const double m1 = w[1], m2 = w[2], E = w[3];
    const double ir = rcp_nr(w[0]);
    const double v1 = m1 * ir, v2 = m2 * ir;
    const double p = gm1 * (E - 0.5 * (m1 * v1 + m2 * v2));
    const double T = E + p;
    // f = (m1, m1 v1 + p, m1 v2, T v1), g = (m2, m2 v1, m2 v2 + p, T v2)
    double m1b = fb[0] + fb[1] * v1 + fb[2] * v2;
    double m2b = gb[0] + gb[1] * v1 + gb[2] * v2;
    const double Tb = fb[3] * v1 + gb[3] * v2;
    double v1b = fb[1] * m1 + gb[1] * m2 + fb[3] * T;
    double v2b = fb[2] * m1 + gb[2] * m2 + gb[3] * T;
    const double pb = fb[1] + gb[2] + Tb;
    // T = E + p, p = gm1 (E - ½ (m1 v1 + m2 v2))
    const double Eb = Tb + gm1 * pb;
    const double q = -0.5 * gm1 * pb;
    m1b = fma(q, v1, m1b);
    v1b = fma(q, m1, v1b);
    m2b = fma(q, v2, m2b);
    v2b = fma(q, m2, v2b);
    // v = m / ρ
    m1b = fma(v1b, ir, m1b);
    m2b = fma(v2b, ir, m2b);
    const double irb = v1b * m1 + v2b * m2;
    wb[0] = -irb * ir * ir;
    wb[1] = m1b;
    wb[2] = m2b;
    wb[3] = Eb;
}

// Gradients of |v_n| and of c = sqrt(γ p / ρ) with respect to the state.
__device__ __forceinline__ void speed_grads(const double* s, double gamma, int orientation, double* da, double* dc) {
    const double ir = rcp_nr(s[0]);
    const double v1 = s[1] * ir, v2 = s[2] * ir;
    const double gm1 = gamma - 1.0;
    const double p = gm1 * (s[3] - 0.5 * (s[1] * v1 + s[2] * v2));
    const double vn = orientation == 1 ? v1 : v2;
    const double sg = vn < 0.0 ? -1.0 : 1.0;
    da[0] = -sg * vn * ir;
    da[1] = orientation == 1 ? sg * ir : 0.0;
    da[2] = orientation == 1 ? 0.0 : sg * ir;
    da[3] = 0.0;
    // c² = γ p/ρ,  ∂(p/ρ) = gm1/ρ (-E/ρ + |v|², -v1, -v2, 1)
    const double cs = sqrt_gs(gamma * p * ir);
    const double k = 0.5 * gamma * gm1 * ir / cs;
    dc[0] = k * (v1 * v1 + v2 * v2 - s[3] * ir);
    dc[1] = -k * v1;
    dc[2] = -k * v2;
    dc[3] = k;
}

// Transposed derivative contraction: with a[v] of every node of the
// workgroup in s (stride S), Fo[v] = Σ_k M[k,i] a[v](k, j) and
// Go[v] = Σ_k M[k,j] a[v](i, k) -- the adjoint of contract<n, S>.
template <int n, int S>
__device__ __forceinline__ void contract_t(const double* s, int base, int i, int j, const double* M, double* Fo,
                                           double* Go) {
    double Mi[n], Mj[n];
#pragma unroll
    for (int k = 0; k < n; ++k) {
        Mi[k] = M[k * n + i];
        Mj[k] = M[k * n + j];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double af = 0.0, ag = 0.0;
#pragma unroll
        for (int k = 0; k < n; ++k) {
            af = fma(Mi[k], s[v * S + base + j * n + k], af);
            ag = fma(Mj[k], s[v * S + base + k * n + i], ag);
        }
        Fo[v] = af;
        Go[v] = ag;
    }
}

template <int N> constexpr int kEBA = (N == 4) ? 5 : 8;  // elements per adjoint workgroup

template <int N, int EB, bool KEEP>
__device__ __forceinline__ void adj_forward(const double* uN, double dt, double invJ, double gm1, const DevConst* c,
                                            int i, int j, bool act, int sidx, int sbase, double* lds, double* sDm,
                                            double* sDh, double* Uc, double* tder) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, NN = G_::NN, S = G_::S;
    if (threadIdx.x < NN) {
        sDm[threadIdx.x] = -dt * c->D[threadIdx.x / n][threadIdx.x % n];
        sDh[threadIdx.x] = c->Dhat[threadIdx.x / n][threadIdx.x % n];
    }
    double F[4], G[4], vol[4];
    if constexpr (N == 4)
        lw4_element<N, EB, KEEP>(uN, dt, invJ, gm1, c, i, j, act, sidx, sbase, lds, lds + 8 * S, lds + 16 * S, sDm,
                                 sDh, Uc, F, G, vol, tder);
    else
        lw3_element<N, EB, KEEP>(uN, dt, invJ, gm1, c, i, j, act, sidx, sbase, lds, lds + 8 * S, lds + 16 * S, sDm,
                                 sDh, Uc, F, G, vol, tder);
}

// Face records of the adjoint: every face node's u, the λ ingredients of Ũ and
// s̄ = -J⁻¹ dt λ' (a.u = state, a.us = λ').
template <int N, int EB>
__global__ __launch_bounds__((VGeo<N, EB>::NT), 2) void k_adj_records(KArgs a) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, NN = G_::NN;
    __shared__ double lds[G_::LDS];
    __shared__ double sDm[NN], sDh[NN];
    const int tid = threadIdx.x;
    const int le = tid / NN, node = tid - (tid / NN) * NN;
    const int i = node % n, j = node / n;
    const int e0 = blockIdx.x * EB;
    const int nel = min(EB, a.E - e0);
    const bool act = le < nel;
    const int lec = le < nel ? le : nel - 1;
    const int e = e0 + lec;
    const int sidx = lec * NN + node, sbase = lec * NN;
    const DevConst* c = a.c;
    const double dt = a.dt_arg, invJ = a.inv_jac[e];
    const size_t off = ((size_t)e * NN + node) * 4;
    double uN[4], Uc[4];
    load4(a.u + off, uN);
    adj_forward<N, EB, false>(uN, dt, invJ, c->gamma - 1.0, c, i, j, act, sidx, sbase, lds, sDm, sDh, Uc, nullptr);
    const int fx = (i == 0) ? 0 : ((i == n - 1) ? 1 : -1);
    const int fy = (j == 0) ? 2 : ((j == n - 1) ? 3 : -1);
    if (!act || (fx < 0 && fy < 0)) return;
    double lam[4], r[kAdjQ], av1, av2, cs;
    load4(a.us + off, lam);
    speed_parts2(Uc, c->gamma, av1, av2, cs);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        r[v] = uN[v];
        r[6 + v] = (-invJ * dt) * lam[v];
    }
    r[5] = cs;
    if (fx >= 0) {
        r[4] = av1;
        store_record<kAdjQ>(a.trace + (((size_t)e * 4 + fx) * n + j) * kAdjQ, r);
    }
    if (fy >= 0) {
        r[4] = av2;
        store_record<kAdjQ>(a.trace + (((size_t)e * 4 + fy) * n + i) * kAdjQ, r);
    }
}

// One adjoint step: a.out = λ = λ' + dt (∂R/∂u)ᵀ λ'  (a.u = u, a.us = λ').
template <int N, int EB>
__global__ __launch_bounds__((VGeo<N, EB>::NT), 2) void k_adjoint(KArgs a) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, NN = G_::NN, S = G_::S;
    __shared__ double lds[G_::LDS];
    __shared__ double sDm[NN], sDh[NN];
    const int tid = threadIdx.x;
    const int le = tid / NN, node = tid - (tid / NN) * NN;
    const int i = node % n, j = node / n;
    const int e0 = blockIdx.x * EB;
    const int nel = min(EB, a.E - e0);
    const bool act = le < nel;
    const int lec = le < nel ? le : nel - 1;
    const int e = e0 + lec;
    const int sidx = lec * NN + node, sbase = lec * NN;
    const DevConst* c = a.c;
    const double gamma = c->gamma, gm1 = gamma - 1.0;
    const double dt = a.dt_arg, invJ = a.inv_jac[e];
    const size_t off = ((size_t)e * NN + node) * 4;
    double uN[4], lam[4], Uc[4], ud[N][4];
    load4(a.u + off, uN);
    load4(a.us + off, lam);
    adj_forward<N, EB, true>(uN, dt, invJ, gm1, c, i, j, act, sidx, sbase, lds, sDm, sDh, Uc, &ud[0][0]);

    // --- interface fluxes of this node's faces
    double sb[4], Fb[4] = {0, 0, 0, 0}, Gb[4] = {0, 0, 0, 0}, Ucb[4] = {0, 0, 0, 0}, ub[4] = {0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 4; ++v) sb[v] = (-invJ * dt) * lam[v];
    const int fx = (i == 0) ? 0 : ((i == n - 1) ? 1 : -1);
    const int fy = (j == 0) ? 2 : ((j == n - 1) ? 3 : -1);
    auto face = [&](int f, int l, double* Fnb) {
        const int nb = a.nbr[4 * e + f];
        double r[kAdjQ];
        load_record<kAdjQ>(a.trace + (((size_t)nb * 4 + (f ^ 1)) * n + l) * kAdjQ, r);
        const int orientation = f < 2 ? 1 : 2;
        double av1, av2, cs;
        speed_parts2(Uc, gamma, av1, av2, cs);
        const double aown = orientation == 1 ? av1 : av2;
        const bool own_left = (f & 1) != 0;
        const double aL = own_left ? aown : r[4], aR = own_left ? r[4] : aown;
        const double cL = own_left ? cs : r[5], cR = own_left ? r[5] : cs;
        const double lamf = dmax(aL, aR) + dmax(cL, cR);
        double Fsb[4], lamb = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double sL = own_left ? sb[v] : r[6 + v];
            const double sR = own_left ? r[6 + v] : sb[v];
            Fsb[v] = c->bnd_hi * sL - c->bnd_lo * sR;
            const double jump = own_left ? r[v] - uN[v] : uN[v] - r[v];  // u_R - u_L
            lamb = fma(-0.5 * Fsb[v], jump, lamb);
            Fnb[v] = fma(0.5, Fsb[v], Fnb[v]);
            ub[v] = fma(own_left ? 0.5 * lamf : -0.5 * lamf, Fsb[v], ub[v]);
        }
        // dmax(x, y) = y > x ? y : x picks the right side only when it is larger
        const bool a_own = own_left ? !(aR > aL) : (aR > aL);
        const bool c_own = own_left ? !(cR > cL) : (cR > cL);
        if (a_own || c_own) {
            double da[4], dc[4];
            speed_grads(Uc, gamma, orientation, da, dc);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (a_own) Ucb[v] = fma(lamb, da[v], Ucb[v]);
                if (c_own) Ucb[v] = fma(lamb, dc[v], Ucb[v]);
            }
        }
    };
    if (fx >= 0) face(fx, j, Fb);
    if (fy >= 0) face(fy, i, Gb);

    // --- volume term D̂·(F̃, G̃): (F̄, Ḡ) += D̂ᵀ s̄
    double* ex0 = lds;          // exchange buffers, 4 doubles per node each
    double* ex1 = lds + 4 * S;  // (the forward's last reads were of s2)
    if (act) put4<S>(ex0, sidx, sb);
    __syncthreads();
    {
        double cf[4], cg[4];
        contract_t<n, S>(ex0, sbase, i, j, sDh, cf, cg);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            Fb[v] += cf[v];
            Gb[v] += cg[v];
        }
    }
    // --- reverse of the LW procedure
    double F0b[4], G0b[4], udb[N][4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F0b[v] = Fb[v];
        G0b[v] = Gb[v];
        ub[v] += Ucb[v];
#pragma unroll
        for (int d = 0; d < N; ++d) udb[d][v] = LWTaylor::w[d] * Ucb[v];
    }
    int buf = 1;
#pragma unroll
    for (int k = N; k >= 1; --k) {
        double dxb[4], dyb[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dxb[v] = LWTaylor::w[k - 1] * Fb[v];
            dyb[v] = LWTaylor::w[k - 1] * Gb[v];
        }
        if (k < N) {  // u^(k+1) = J⁻¹ C(dx_k, dy_k) with C the (-dt D) contraction
            double* ex = buf ? ex1 : ex0;
            buf ^= 1;
            double t4[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) t4[v] = invJ * udb[k][v];
            if (act) put4<S>(ex, sidx, t4);
            __syncthreads();
            double cf[4], cg[4];
            contract_t<n, S>(ex, sbase, i, j, sDm, cf, cg);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                dxb[v] += cf[v];
                dyb[v] += cg[v];
            }
        }
        const double f0 = LWStencil<N>::c[k - 1][4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            F0b[v] = fma(f0, dxb[v], F0b[v]);
            G0b[v] = fma(f0, dyb[v], G0b[v]);
        }
#pragma unroll
        for (int kind = 0; kind < 4; ++kind) {
            const double cc = LWStencil<N>::c[k - 1][kind];
            if (cc == 0.0) continue;
            double w[4], fb[4], gb[4], wb[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double x = uN[v];
#pragma unroll
                for (int d = 0; d < k; ++d) x = fma(LWTaylor::a[kind][d], ud[d][v], x);
                w[v] = x;
                fb[v] = cc * dxb[v];
                gb[v] = cc * dyb[v];
            }
            euler_fluxes_vjp(w, gm1, fb, gb, wb);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                ub[v] += wb[v];
#pragma unroll
                for (int d = 0; d < k; ++d) udb[d][v] = fma(LWTaylor::a[kind][d], wb[v], udb[d][v]);
            }
        }
    }
    {  // ut = J⁻¹ C(F0, G0)
        double* ex = buf ? ex1 : ex0;
        double t4[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) t4[v] = invJ * udb[0][v];
        if (act) put4<S>(ex, sidx, t4);
        __syncthreads();
        double cf[4], cg[4];
        contract_t<n, S>(ex, sbase, i, j, sDm, cf, cg);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            F0b[v] += cf[v];
            G0b[v] += cg[v];
        }
    }
    double wb[4];
    euler_fluxes_vjp(uN, gm1, F0b, G0b, wb);
    double out[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) out[v] = lam[v] + (ub[v] + wb[v]);
    if (act) store4(a.out + off, out);
}

}  // namespace tlw

namespace tlw {

// ---------------------------------------------------------------------------
// Discrete adjoint of one MDRK step on a periodic TreeMesh (plain FR, TreeMesh
// flux order):  us = u + dt R1(u),  u' = u + dt R2(u, us)  (mdrk.jl:13-25), so
//   λ_us = dt (∂R2/∂us)ᵀ λ',  λ = λ' + dt (∂R2/∂u)ᵀ λ' + λ_us + dt (∂R1/∂u)ᵀ λ_us.
// R1 and R2 are the stage residuals of mdrk1_element / mdrk2_element
// (tree/multi_derivative_2d.jl:263-555 via the structured kernels'
// formulas, structured/multi_derivative_2d.jl:82-445):
//   stage 1: ut = J⁻¹ C(f(u)),  F̃ = f(u)/2 + f_t(u, ut)/8,  Ũ = u/2 + ut/8;
//   stage 2: ut, ust likewise from u and us,  F̃ = f(u) + f_t(u, ut)/6 +
//            f_t(us, ust)/3,  Ũ = u + ut/6 + ust/3,
// with f_t(u, ut) = (-f(u+2ut) + 8 f(u+ut) - 8 f(u-ut) + f(u-2ut))/12, C the
// (-dt D) contraction, and the interface flux of the LW step (jump on uⁿ, λ
// from Ũ).  Each stage is a record kernel (face node u, the λ ingredients of
// Ũ, s̄ = -J⁻¹ dt λ_stage) and a reverse kernel, as for the LW step.

struct AdjArgs {
    int E;
    const double* u;         // uⁿ
    const double* us;        // stage 2: the stage state us
    const double* lam;       // the adjoint of the stage's output (λ' or λ_us)
    const double* lam_base;  // added to the result (stage 1: stage 2's λ), or null
    double* lam_out;         // the adjoint of uⁿ
    double* lam_us;          // stage 2: the adjoint of us
    double* trace;           // [E][4][n][kAdjQ] face records
    const int* nbr;
    const double* inv_jac;
    const DevConst* c;
    double dt;
};

// FD coefficients and Taylor offsets of f_t(u, ut): kinds u+2ut, u+ut, u-ut, u-2ut
struct MdrkFt {
    static constexpr double a[4] = {2.0, 1.0, -1.0, -2.0};
    static constexpr double c[4] = {-1.0 / 12.0, 8.0 / 12.0, -8.0 / 12.0, 1.0 / 12.0};
};

// ub += Σ_k c_k (∂f/∂w)ᵀ(dxb, dyb) at w_k = u + a_k ut, utb += a_k times the same
__device__ __forceinline__ void mdrk_ft_vjp(const double* u, const double* ut, double gm1, const double* dxb,
                                            const double* dyb, double* ub, double* utb) {
#pragma unroll
    for (int kind = 0; kind < 4; ++kind) {
        double w[4], fb[4], gb[4], wb[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            w[v] = fma(MdrkFt::a[kind], ut[v], u[v]);
            fb[v] = MdrkFt::c[kind] * dxb[v];
            gb[v] = MdrkFt::c[kind] * dyb[v];
        }
        euler_fluxes_vjp(w, gm1, fb, gb, wb);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            ub[v] += wb[v];
            utb[v] = fma(MdrkFt::a[kind], wb[v], utb[v]);
        }
    }
}

// Forward quantities of a stage the adjoint needs: ut (and ust) and Ũ.
template <int N, int EB, int STAGE>
__device__ __forceinline__ void mdrk_adj_forward(const double* uN, const double* usn, double invJ, double gm1, int i,
                                                 int j, bool act, int sidx, int sbase, double* lds, const double* sDm,
                                                 double* ut, double* ust, double* Uc) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, S = G_::S;
    double f[4], g[4];
    euler_fluxes(uN, gm1, f, g);
    if (act) put8<S>(lds, sidx, f, g);
    if constexpr (STAGE == 2) {
        euler_fluxes(usn, gm1, f, g);
        if (act) put8<S>(lds + 8 * S, sidx, f, g);
    }
    __syncthreads();
    contract<n, S>(lds, sbase, i, j, sDm, ut);
#pragma unroll
    for (int v = 0; v < 4; ++v) ut[v] *= invJ;
    if constexpr (STAGE == 2) {
        contract<n, S>(lds + 8 * S, sbase, i, j, sDm, ust);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            ust[v] *= invJ;
            Uc[v] = fma(1.0 / 3.0, ust[v], fma(1.0 / 6.0, ut[v], uN[v]));
        }
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) Uc[v] = fma(0.125, ut[v], 0.5 * uN[v]);
    }
    __syncthreads();  // the exchange buffers are reused by the caller
}

template <int N, int EB, int STAGE>
__global__ __launch_bounds__((VGeo<N, EB>::NT), 2) void k_adj_records_md(AdjArgs a) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, NN = G_::NN;
    __shared__ double lds[G_::LDS];
    __shared__ double sDm[NN];
    const int tid = threadIdx.x;
    const int le = tid / NN, node = tid - (tid / NN) * NN;
    const int i = node % n, j = node / n;
    const int e0 = blockIdx.x * EB;
    const int nel = min(EB, a.E - e0);
    const bool act = le < nel;
    const int lec = le < nel ? le : nel - 1;
    const int e = e0 + lec;
    const int sidx = lec * NN + node, sbase = lec * NN;
    const DevConst* c = a.c;
    const double dt = a.dt, invJ = a.inv_jac[e];
    const size_t off = ((size_t)e * NN + node) * 4;
    if (tid < NN) sDm[tid] = -dt * c->D[tid / n][tid % n];
    double uN[4], usn[4] = {0, 0, 0, 0}, ut[4], ust[4], Uc[4];
    load4(a.u + off, uN);
    if constexpr (STAGE == 2) load4(a.us + off, usn);
    mdrk_adj_forward<N, EB, STAGE>(uN, usn, invJ, c->gamma - 1.0, i, j, act, sidx, sbase, lds, sDm, ut, ust, Uc);
    const int fx = (i == 0) ? 0 : ((i == n - 1) ? 1 : -1);
    const int fy = (j == 0) ? 2 : ((j == n - 1) ? 3 : -1);
    if (!act || (fx < 0 && fy < 0)) return;
    double lam[4], r[kAdjQ], av1, av2, cs;
    load4(a.lam + off, lam);
    speed_parts2(Uc, c->gamma, av1, av2, cs);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        r[v] = uN[v];
        r[6 + v] = (-invJ * dt) * lam[v];
    }
    r[5] = cs;
    if (fx >= 0) {
        r[4] = av1;
        store_record<kAdjQ>(a.trace + (((size_t)e * 4 + fx) * n + j) * kAdjQ, r);
    }
    if (fy >= 0) {
        r[4] = av2;
        store_record<kAdjQ>(a.trace + (((size_t)e * 4 + fy) * n + i) * kAdjQ, r);
    }
}

template <int N, int EB, int STAGE>
__global__ __launch_bounds__((VGeo<N, EB>::NT), 2) void k_adjoint_md(AdjArgs a) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, NN = G_::NN, S = G_::S;
    __shared__ double lds[G_::LDS];
    __shared__ double sDm[NN], sDh[NN];
    const int tid = threadIdx.x;
    const int le = tid / NN, node = tid - (tid / NN) * NN;
    const int i = node % n, j = node / n;
    const int e0 = blockIdx.x * EB;
    const int nel = min(EB, a.E - e0);
    const bool act = le < nel;
    const int lec = le < nel ? le : nel - 1;
    const int e = e0 + lec;
    const int sidx = lec * NN + node, sbase = lec * NN;
    const DevConst* c = a.c;
    const double gamma = c->gamma, gm1 = gamma - 1.0;
    const double dt = a.dt, invJ = a.inv_jac[e];
    const size_t off = ((size_t)e * NN + node) * 4;
    if (tid < NN) {
        sDm[tid] = -dt * c->D[tid / n][tid % n];
        sDh[tid] = c->Dhat[tid / n][tid % n];
    }
    double uN[4], usn[4] = {0, 0, 0, 0}, lam[4], ut[4], ust[4] = {0, 0, 0, 0}, Uc[4];
    load4(a.u + off, uN);
    if constexpr (STAGE == 2) load4(a.us + off, usn);
    load4(a.lam + off, lam);
    mdrk_adj_forward<N, EB, STAGE>(uN, usn, invJ, gm1, i, j, act, sidx, sbase, lds, sDm, ut, ust, Uc);

    // --- interface fluxes of this node's faces (the LW step's: jump on uⁿ, λ from Ũ)
    double sb[4], Fb[4] = {0, 0, 0, 0}, Gb[4] = {0, 0, 0, 0}, Ucb[4] = {0, 0, 0, 0}, ub[4] = {0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 4; ++v) sb[v] = (-invJ * dt) * lam[v];
    const int fx = (i == 0) ? 0 : ((i == n - 1) ? 1 : -1);
    const int fy = (j == 0) ? 2 : ((j == n - 1) ? 3 : -1);
    auto face = [&](int f, int l, double* Fnb) {
        const int nb = a.nbr[4 * e + f];
        double r[kAdjQ];
        load_record<kAdjQ>(a.trace + (((size_t)nb * 4 + (f ^ 1)) * n + l) * kAdjQ, r);
        const int orientation = f < 2 ? 1 : 2;
        double av1, av2, cs;
        speed_parts2(Uc, gamma, av1, av2, cs);
        const double aown = orientation == 1 ? av1 : av2;
        const bool own_left = (f & 1) != 0;
        const double aL = own_left ? aown : r[4], aR = own_left ? r[4] : aown;
        const double cL = own_left ? cs : r[5], cR = own_left ? r[5] : cs;
        const double lamf = dmax(aL, aR) + dmax(cL, cR);
        double lamb = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double sL = own_left ? sb[v] : r[6 + v];
            const double sR = own_left ? r[6 + v] : sb[v];
            const double Fsb = c->bnd_hi * sL - c->bnd_lo * sR;
            const double jump = own_left ? r[v] - uN[v] : uN[v] - r[v];  // u_R - u_L
            lamb = fma(-0.5 * Fsb, jump, lamb);
            Fnb[v] = fma(0.5, Fsb, Fnb[v]);
            ub[v] = fma(own_left ? 0.5 * lamf : -0.5 * lamf, Fsb, ub[v]);
        }
        const bool a_own = own_left ? !(aR > aL) : (aR > aL);
        const bool c_own = own_left ? !(cR > cL) : (cR > cL);
        if (a_own || c_own) {
            double da[4], dc[4];
            speed_grads(Uc, gamma, orientation, da, dc);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (a_own) Ucb[v] = fma(lamb, da[v], Ucb[v]);
                if (c_own) Ucb[v] = fma(lamb, dc[v], Ucb[v]);
            }
        }
    };
    if (fx >= 0) face(fx, j, Fb);
    if (fy >= 0) face(fy, i, Gb);

    // --- volume term D̂·(F̃, G̃): (F̄, Ḡ) += D̂ᵀ s̄
    double* ex0 = lds;
    double* ex1 = lds + 4 * S;
    if (act) put4<S>(ex0, sidx, sb);
    __syncthreads();
    {
        double cf[4], cg[4];
        contract_t<n, S>(ex0, sbase, i, j, sDh, cf, cg);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            Fb[v] += cf[v];
            Gb[v] += cg[v];
        }
    }
    // --- reverse of the stage's F̃ and Ũ
    double f0b[4], g0b[4], utb[4], ustb[4] = {0, 0, 0, 0}, usb[4] = {0, 0, 0, 0};
    {
        constexpr double wf = STAGE == 2 ? 1.0 : 0.5;       // weight of f(u) in F̃
        constexpr double wt = STAGE == 2 ? 1.0 / 6.0 : 0.125;  // weight of f_t(u, ut) and of ut in Ũ
        constexpr double wu = STAGE == 2 ? 1.0 : 0.5;       // weight of u in Ũ
        double dxb[4], dyb[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            ub[v] = fma(wu, Ucb[v], ub[v]);
            utb[v] = wt * Ucb[v];
            f0b[v] = wf * Fb[v];
            g0b[v] = wf * Gb[v];
            dxb[v] = wt * Fb[v];
            dyb[v] = wt * Gb[v];
        }
        mdrk_ft_vjp(uN, ut, gm1, dxb, dyb, ub, utb);
        if constexpr (STAGE == 2) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                ustb[v] = (1.0 / 3.0) * Ucb[v];
                dxb[v] = (1.0 / 3.0) * Fb[v];
                dyb[v] = (1.0 / 3.0) * Gb[v];
            }
            mdrk_ft_vjp(usn, ust, gm1, dxb, dyb, usb, ustb);
        }
    }
    // ut = J⁻¹ C(f(u)) (and ust = J⁻¹ C(f(us))): (f̄, ḡ) += Cᵀ J⁻¹ ūt
    __syncthreads();  // every lane has read ex0
    {
        double t4[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) t4[v] = invJ * utb[v];
        if (act) put4<S>(ex1, sidx, t4);
        if constexpr (STAGE == 2) {
#pragma unroll
            for (int v = 0; v < 4; ++v) t4[v] = invJ * ustb[v];
            if (act) put4<S>(ex0, sidx, t4);
        }
    }
    __syncthreads();
    double cf[4], cg[4], wb[4];
    contract_t<n, S>(ex1, sbase, i, j, sDm, cf, cg);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        f0b[v] += cf[v];
        g0b[v] += cg[v];
    }
    euler_fluxes_vjp(uN, gm1, f0b, g0b, wb);
#pragma unroll
    for (int v = 0; v < 4; ++v) ub[v] += wb[v];
    if constexpr (STAGE == 2) {
        contract_t<n, S>(ex0, sbase, i, j, sDm, cf, cg);
        euler_fluxes_vjp(usn, gm1, cf, cg, wb);
#pragma unroll
        for (int v = 0; v < 4; ++v) usb[v] += wb[v];
        if (act) store4(a.lam_us + off, usb);
    }
    double out[4];
    if (a.lam_base) {
        double base[4];
        load4(a.lam_base + off, base);
#pragma unroll
        for (int v = 0; v < 4; ++v) out[v] = base[v] + (lam[v] + ub[v]);
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) out[v] = lam[v] + ub[v];
    }
    if (act) store4(a.lam_out + off, out);
}

}  // namespace tlw

namespace tlw {

// ---------------------------------------------------------------------------
// Discrete adjoint of one LW step with VolumeIntegralFRShockCapturing and
// first-order subcell FV (TreeMesh, periodic), the blending factors α and the
// interface's admissibility decision frozen at the step's state:
//   vol = (1 - α_e) D̂·(F̃, G̃) + α_e fv(u),  fv = w⁻¹ differences of the subcell
//   Rusanov fluxes (fv_kernel!, dg_2d.jl:251-282);
//   F* = alp fn(u_L, u_R) + (1 - alp) F*_LW,  alp = ½(α_L + α_R) or 1 where the
//   low-order update is not admissible (compute_alp, dg_2d.jl:1873-1901).
// α comes from k_indicator on the step's state, smoothed as the volume kernel
// does; fn_low (the subcell fluxes next to the faces) enters only the frozen
// decision.  Records: the forward shock-capturing face record (16 doubles)
// and s̄ = -J⁻¹ dt λ'.

constexpr int kAdjQSC = 20;

// (ulb, urb) += the adjoint of rusanov(ul, ur, orientation) for the output adjoint ob
__device__ __forceinline__ void rusanov_vjp(const double* ul, const double* ur, double gamma, int orientation,
                                            const double* ob, double* ulb, double* urb) {
    double al, cl, ar, cr;
    speed_parts(ul, gamma, orientation, al, cl);
    speed_parts(ur, gamma, orientation, ar, cr);
    const double lam = dmax(al, ar) + dmax(cl, cr);
    double hb[4], zero[4] = {0, 0, 0, 0}, wb[4], lamb = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        hb[v] = 0.5 * ob[v];
        lamb = fma(-0.5 * ob[v], ur[v] - ul[v], lamb);
        ulb[v] = fma(0.5 * lam, ob[v], ulb[v]);
        urb[v] = fma(-0.5 * lam, ob[v], urb[v]);
    }
    const double gm1 = gamma - 1.0;
    euler_fluxes_vjp(ul, gm1, orientation == 1 ? hb : zero, orientation == 1 ? zero : hb, wb);
#pragma unroll
    for (int v = 0; v < 4; ++v) ulb[v] += wb[v];
    euler_fluxes_vjp(ur, gm1, orientation == 1 ? hb : zero, orientation == 1 ? zero : hb, wb);
#pragma unroll
    for (int v = 0; v < 4; ++v) urb[v] += wb[v];
    double da[4], dc[4];
    const bool a_right = ar > al, c_right = cr > cl;  // dmax picks the right side only when larger
    speed_grads(a_right ? ur : ul, gamma, orientation, da, dc);
    double* pa = a_right ? urb : ulb;
#pragma unroll
    for (int v = 0; v < 4; ++v) pa[v] = fma(lamb, da[v], pa[v]);
    speed_grads(c_right ? ur : ul, gamma, orientation, da, dc);
    double* pc = c_right ? urb : ulb;
#pragma unroll
    for (int v = 0; v < 4; ++v) pc[v] = fma(lamb, dc[v], pc[v]);
}

// The interface blending factor of face_numflux_sc for the records L, R.
__device__ __forceinline__ double sc_face_alp(const double* L, const double* R, int orientation, double gamma,
                                              double J, double dt, double w_n, double w_1) {
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
    return ok ? alp : 1.0;
}

// Forward quantities of the blended step at one node: Ũ, F̃, the time
// derivatives (KEEP), α_e and the subcell fluxes next to the node's faces.
template <int N, int EB, bool KEEP>
__device__ __forceinline__ void adj_forward_sc(const AdjArgs& a, const double* alpha_tmp, const double* uN, int e,
                                               int i, int j, bool act, int sidx, int sbase, double* lds, double* sDm,
                                               double* sDh, double* Uc, double* F, double* G, double* tder,
                                               double& alpha_e, double* fnlx, double* fnly) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, NN = G_::NN, S = G_::S;
    const DevConst* c = a.c;
    const double dt = a.dt, invJ = a.inv_jac[e], gamma = c->gamma, gm1 = gamma - 1.0;
    if (threadIdx.x < NN) {
        sDm[threadIdx.x] = -dt * c->D[threadIdx.x / n][threadIdx.x % n];
        sDh[threadIdx.x] = c->Dhat[threadIdx.x / n][threadIdx.x % n];
    }
    double vol[4];
    if constexpr (N == 4)
        lw4_element<N, EB, KEEP>(uN, dt, invJ, gm1, c, i, j, act, sidx, sbase, lds, lds + 8 * S, lds + 16 * S, sDm,
                                 sDh, Uc, F, G, vol, tder);
    else
        lw3_element<N, EB, KEEP>(uN, dt, invJ, gm1, c, i, j, act, sidx, sbase, lds, lds + 8 * S, lds + 16 * S, sDm,
                                 sDh, Uc, F, G, vol, tder);
    alpha_e = alpha_tmp[e];
    if (c->alpha_smooth) {
#pragma unroll
        for (int f = 0; f < 4; ++f) alpha_e = dmax(alpha_e, 0.5 * alpha_tmp[a.nbr[4 * e + f]]);
    }
    __syncthreads();  // the element procedure's last reads of s2 are done
    double* su = lds;
    if (act) put4<S>(su, sidx, uN);
    __syncthreads();
    double w[4];
    // subcell flux between nodes (0, j) and (1, j), or (n-2, j) and (n-1, j); likewise in y
    if (i == 0 || i == n - 1) {
#pragma unroll
        for (int v = 0; v < 4; ++v) w[v] = su[v * S + sidx + (i == 0 ? 1 : -1)];
        if (i == 0) rusanov(uN, w, gamma, 1, fnlx);
        else rusanov(w, uN, gamma, 1, fnlx);
    }
    if (j == 0 || j == n - 1) {
#pragma unroll
        for (int v = 0; v < 4; ++v) w[v] = su[v * S + sidx + (j == 0 ? n : -n)];
        if (j == 0) rusanov(uN, w, gamma, 2, fnly);
        else rusanov(w, uN, gamma, 2, fnly);
    }
    __syncthreads();
}

__device__ __forceinline__ void sc_own_record(const double* uN, const double* Fn, const double* Uc, double gamma,
                                              int orientation, const double* fnl, double alpha_e, double* r) {
    double av1, av2, cs;
    speed_parts2(Uc, gamma, av1, av2, cs);
    make_record(r, uN, Fn, orientation == 1 ? av1 : av2, cs);
#pragma unroll
    for (int v = 0; v < 4; ++v) r[10 + v] = fnl[v];
    r[14] = alpha_e;
    r[15] = 0.0;
}

template <int N, int EB>
__global__ __launch_bounds__((VGeo<N, EB>::NT), 2) void k_adj_records_sc(AdjArgs a, const double* alpha_tmp) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, NN = G_::NN;
    __shared__ double lds[G_::LDS];
    __shared__ double sDm[NN], sDh[NN];
    const int tid = threadIdx.x;
    const int le = tid / NN, node = tid - (tid / NN) * NN;
    const int i = node % n, j = node / n;
    const int e0 = blockIdx.x * EB;
    const int nel = min(EB, a.E - e0);
    const bool act = le < nel;
    const int lec = le < nel ? le : nel - 1;
    const int e = e0 + lec;
    const int sidx = lec * NN + node, sbase = lec * NN;
    const size_t off = ((size_t)e * NN + node) * 4;
    double uN[4], Uc[4], F[4], G[4], fnlx[4] = {0, 0, 0, 0}, fnly[4] = {0, 0, 0, 0}, alpha_e;
    load4(a.u + off, uN);
    adj_forward_sc<N, EB, false>(a, alpha_tmp, uN, e, i, j, act, sidx, sbase, lds, sDm, sDh, Uc, F, G, nullptr,
                                 alpha_e, fnlx, fnly);
    const int fx = (i == 0) ? 0 : ((i == n - 1) ? 1 : -1);
    const int fy = (j == 0) ? 2 : ((j == n - 1) ? 3 : -1);
    if (!act || (fx < 0 && fy < 0)) return;
    const double invJ = a.inv_jac[e], gamma = a.c->gamma;
    double lam[4], r[kAdjQSC];
    load4(a.lam + off, lam);
#pragma unroll
    for (int v = 0; v < 4; ++v) r[16 + v] = (-invJ * a.dt) * lam[v];
    if (fx >= 0) {
        sc_own_record(uN, F, Uc, gamma, 1, fnlx, alpha_e, r);
        store_record<kAdjQSC>(a.trace + (((size_t)e * 4 + fx) * n + j) * kAdjQSC, r);
    }
    if (fy >= 0) {
        sc_own_record(uN, G, Uc, gamma, 2, fnly, alpha_e, r);
        store_record<kAdjQSC>(a.trace + (((size_t)e * 4 + fy) * n + i) * kAdjQSC, r);
    }
}

template <int N, int EB>
__global__ __launch_bounds__((VGeo<N, EB>::NT), 1) void k_adjoint_sc(AdjArgs a, const double* alpha_tmp, double J0) {
    using G_ = VGeo<N, EB>;
    constexpr int n = G_::n, NN = G_::NN, S = G_::S;
    __shared__ double lds[G_::LDS];
    __shared__ double sDm[NN], sDh[NN];
    const int tid = threadIdx.x;
    const int le = tid / NN, node = tid - (tid / NN) * NN;
    const int i = node % n, j = node / n;
    const int e0 = blockIdx.x * EB;
    const int nel = min(EB, a.E - e0);
    const bool act = le < nel;
    const int lec = le < nel ? le : nel - 1;
    const int e = e0 + lec;
    const int sidx = lec * NN + node, sbase = lec * NN;
    const DevConst* c = a.c;
    const double gamma = c->gamma, gm1 = gamma - 1.0;
    const double dt = a.dt, invJ = a.inv_jac[e];
    const size_t off = ((size_t)e * NN + node) * 4;
    double uN[4], lam[4], Uc[4], F[4], G[4], ud[N][4], fnlx[4] = {0, 0, 0, 0}, fnly[4] = {0, 0, 0, 0}, alpha_e;
    load4(a.u + off, uN);
    load4(a.lam + off, lam);
    adj_forward_sc<N, EB, true>(a, alpha_tmp, uN, e, i, j, act, sidx, sbase, lds, sDm, sDh, Uc, F, G, &ud[0][0],
                                alpha_e, fnlx, fnly);
    const bool blended = !(fabs(alpha_e) <= 1e-12);

    // --- blended interface fluxes of this node's faces, alp frozen
    double sb[4], Fb[4] = {0, 0, 0, 0}, Gb[4] = {0, 0, 0, 0}, Ucb[4] = {0, 0, 0, 0}, ub[4] = {0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 4; ++v) sb[v] = (-invJ * dt) * lam[v];
    const int fx = (i == 0) ? 0 : ((i == n - 1) ? 1 : -1);
    const int fy = (j == 0) ? 2 : ((j == n - 1) ? 3 : -1);
    auto face = [&](int f, int l, const double* Fn, const double* fnl, double* Fnb) {
        const int nb = a.nbr[4 * e + f];
        const int orientation = f < 2 ? 1 : 2;
        double r[kAdjQSC], own[kAdjQSC];
        load_record<kAdjQSC>(a.trace + (((size_t)nb * 4 + (f ^ 1)) * n + l) * kAdjQSC, r);
        sc_own_record(uN, Fn, Uc, gamma, orientation, fnl, alpha_e, own);
        const bool own_left = (f & 1) != 0;
        const double* L = own_left ? own : r;
        const double* R = own_left ? r : own;
        const double alp = sc_face_alp(L, R, orientation, gamma, J0, dt, c->weights[n - 1], c->weights[0]);
        const double lamf = dmax(L[8], R[8]) + dmax(L[9], R[9]);
        double Fsb[4], lamb = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double sL = own_left ? sb[v] : r[16 + v];
            const double sR = own_left ? r[16 + v] : sb[v];
            Fsb[v] = c->bnd_hi * sL - c->bnd_lo * sR;
            // the LW part, weighted 1 - alp
            const double hb = (1.0 - alp) * Fsb[v];
            lamb = fma(-0.5 * hb, R[v] - L[v], lamb);
            Fnb[v] = fma(0.5, hb, Fnb[v]);
            ub[v] = fma(own_left ? 0.5 * lamf : -0.5 * lamf, hb, ub[v]);
            Fsb[v] *= alp;  // the low-order part's adjoint
        }
        const double aown = own[8], aoth = r[8], cown = own[9], coth = r[9];
        const bool a_own = own_left ? !(aoth > aown) : (aown > aoth);
        const bool c_own = own_left ? !(coth > cown) : (cown > coth);
        if (a_own || c_own) {
            double da[4], dc[4];
            speed_grads(Uc, gamma, orientation, da, dc);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (a_own) Ucb[v] = fma(lamb, da[v], Ucb[v]);
                if (c_own) Ucb[v] = fma(lamb, dc[v], Ucb[v]);
            }
        }
        if (alp != 0.0) {
            double ulb[4] = {0, 0, 0, 0}, urb[4] = {0, 0, 0, 0};
            rusanov_vjp(L, R, gamma, orientation, Fsb, ulb, urb);
#pragma unroll
            for (int v = 0; v < 4; ++v) ub[v] += own_left ? ulb[v] : urb[v];
        }
    };
    if (fx >= 0) face(fx, j, F, fnlx, Fb);
    if (fy >= 0) face(fy, i, G, fnly, Gb);

    // --- subcell FV of a blended element: vol += α fv(u)
    double* su = lds;               // u, then the adjoint sent to the left / lower neighbour
    double* ss = lds + 4 * S;       // α s̄
    double* sl = lds + 8 * S;       // [x: 4][y: 4] adjoint for the left / lower node
    {
        double sv[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) sv[v] = blended ? alpha_e * sb[v] : 0.0;
        if (act) {
            put4<S>(su, sidx, uN);
            put4<S>(ss, sidx, sv);
        }
        __syncthreads();
        const double iwi = 1.0 / c->weights[i], iwj = 1.0 / c->weights[j];
        double bx[4] = {0, 0, 0, 0}, by[4] = {0, 0, 0, 0};
        if (blended && i > 0) {  // subcell face between (i-1, j) and (i, j)
            double w[4], ob[4];
            const double iwl = 1.0 / c->weights[i - 1];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                w[v] = su[v * S + sidx - 1];
                ob[v] = iwl * ss[v * S + sidx - 1] - iwi * sv[v];
            }
            rusanov_vjp(w, uN, gamma, 1, ob, bx, ub);
        }
        if (blended && j > 0) {
            double w[4], ob[4];
            const double iwl = 1.0 / c->weights[j - 1];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                w[v] = su[v * S + sidx - n];
                ob[v] = iwl * ss[v * S + sidx - n] - iwj * sv[v];
            }
            rusanov_vjp(w, uN, gamma, 2, ob, by, ub);
        }
        if (act) {
            put4<S>(sl, sidx, bx);
            put4<S>(sl + 4 * S, sidx, by);
        }
        __syncthreads();
        if (blended && i < n - 1) {
#pragma unroll
            for (int v = 0; v < 4; ++v) ub[v] += sl[v * S + sidx + 1];
        }
        if (blended && j < n - 1) {
#pragma unroll
            for (int v = 0; v < 4; ++v) ub[v] += sl[4 * S + v * S + sidx + n];
        }
        __syncthreads();
    }

    // --- volume term (1 - α) D̂·(F̃, G̃)
    double* ex0 = lds;
    double* ex1 = lds + 4 * S;
    {
        double t4[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) t4[v] = (1.0 - alpha_e) * sb[v];
        if (act) put4<S>(ex0, sidx, t4);
    }
    __syncthreads();
    {
        double cf[4], cg[4];
        contract_t<n, S>(ex0, sbase, i, j, sDh, cf, cg);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            Fb[v] += cf[v];
            Gb[v] += cg[v];
        }
    }
    // --- reverse of the LW procedure (as k_adjoint)
    double F0b[4], G0b[4], udb[N][4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        F0b[v] = Fb[v];
        G0b[v] = Gb[v];
        ub[v] += Ucb[v];
#pragma unroll
        for (int d = 0; d < N; ++d) udb[d][v] = LWTaylor::w[d] * Ucb[v];
    }
    int buf = 1;
#pragma unroll
    for (int k = N; k >= 1; --k) {
        double dxb[4], dyb[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dxb[v] = LWTaylor::w[k - 1] * Fb[v];
            dyb[v] = LWTaylor::w[k - 1] * Gb[v];
        }
        if (k < N) {
            double* ex = buf ? ex1 : ex0;
            buf ^= 1;
            double t4[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) t4[v] = invJ * udb[k][v];
            if (act) put4<S>(ex, sidx, t4);
            __syncthreads();
            double cf[4], cg[4];
            contract_t<n, S>(ex, sbase, i, j, sDm, cf, cg);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                dxb[v] += cf[v];
                dyb[v] += cg[v];
            }
        }
        const double f0 = LWStencil<N>::c[k - 1][4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            F0b[v] = fma(f0, dxb[v], F0b[v]);
            G0b[v] = fma(f0, dyb[v], G0b[v]);
        }
#pragma unroll
        for (int kind = 0; kind < 4; ++kind) {
            const double cc = LWStencil<N>::c[k - 1][kind];
            if (cc == 0.0) continue;
            double w[4], fb[4], gb[4], wb[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double x = uN[v];
#pragma unroll
                for (int d = 0; d < k; ++d) x = fma(LWTaylor::a[kind][d], ud[d][v], x);
                w[v] = x;
                fb[v] = cc * dxb[v];
                gb[v] = cc * dyb[v];
            }
            euler_fluxes_vjp(w, gm1, fb, gb, wb);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                ub[v] += wb[v];
#pragma unroll
                for (int d = 0; d < k; ++d) udb[d][v] = fma(LWTaylor::a[kind][d], wb[v], udb[d][v]);
            }
        }
    }
    {
        double* ex = buf ? ex1 : ex0;
        double t4[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) t4[v] = invJ * udb[0][v];
        if (act) put4<S>(ex, sidx, t4);
        __syncthreads();
        double cf[4], cg[4];
        contract_t<n, S>(ex, sbase, i, j, sDm, cf, cg);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            F0b[v] += cf[v];
            G0b[v] += cg[v];
        }
    }
    double wb[4];
    euler_fluxes_vjp(uN, gm1, F0b, G0b, wb);
    double out[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) out[v] = lam[v] + (ub[v] + wb[v]);
    if (act) store4(a.lam_out + off, out);
}

}  // namespace tlw
