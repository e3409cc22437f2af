"""Closed-form LW-FR / MDRK step for linear advection (test infrastructure).

For f = a u the finite-difference time derivatives of the approximate LW
procedure are exact on the Taylor states, so with the element-local operator
A = -Δt J⁻¹ (a1 Dx + a2 Dy) the procedure collapses to matrix polynomials:
  N=4 (src/solvers/dgsem_tree/dg_2d.jl:918-1267):
      Ũ = (I + A/2 + A²/6 + A³/24 + A⁴/120) u
      F̃ = a (I + A/2 + A²/6 + A³/24 + A⁴/240) u     (f_tttt carries a factor 0.5,
                                                      dg_2d.jl:1203-1206)
  N=3 (dg_2d.jl:674-916): Ũ = F̃/a = (I + A/2 + A²/6 + A³/24) u
  MDRK stage 1 (tree/multi_derivative_2d.jl:263-420): F̃/a = Ũ = u/2 + Au/8
  MDRK stage 2 (423-555): F̃/a = Ũ = u + Au/6 + A us/3
and the FR update is u + Δt · (-J⁻¹)(D̂ F̃ + lifted Rusanov face fluxes).
This module evaluates those formulas with dense numpy operators, in gather
form, independently of the oracle's loops.  It is a pin, not a restatement:
a wrong Taylor or finite-difference coefficient, a transposed D, a wrong
lifting sign or the wrong jump/λ arguments of the interface flux changes the
result beyond round-off.
"""
from __future__ import annotations

import numpy as np


class LinearLW:
    def __init__(self, nodes, weights, D, Dhat, nx, ny, h, a=(1.0, 0.7), flux_order="tree"):
        self.nodes, self.w, self.D, self.Dh = nodes, weights, D, Dhat
        self.n = len(nodes)
        self.nx, self.ny = nx, ny
        self.E = nx * ny
        self.invJ = 2.0 / h
        self.a1, self.a2 = a
        self.flux_order = flux_order

    # u has shape (E, n, n) indexed [e, j, i]
    def dx(self, u, M):
        return np.einsum("ik,ejk->eji", M, u)

    def dy(self, u, M):
        return np.einsum("jk,eki->eji", M, u)

    def A(self, u, dt):
        return -dt * self.invJ * (self.a1 * self.dx(u, self.D) + self.a2 * self.dy(u, self.D))

    def nbr(self, e, face):
        ix, iy = e % self.nx, e // self.nx
        if face == 0:
            ix = (ix - 1) % self.nx
        elif face == 1:
            ix = (ix + 1) % self.nx
        elif face == 2:
            iy = (iy - 1) % self.ny
        else:
            iy = (iy + 1) % self.ny
        return ix + self.nx * iy

    def upwind(self, ul, ur, a):
        """Rusanov flux of f = a u: 0.5 a (ul + ur) - 0.5 |a| (ur - ul)."""
        return 0.5 * a * (ul + ur) - 0.5 * abs(a) * (ur - ul)

    def subcell_fv(self, u):
        """First-order subcell FV divergence on the GLL subcells of each element,
        (f_{i+1/2} - f_{i-1/2}) / w_i with zero flux at the element faces."""
        n = self.n
        out = np.zeros_like(u)
        fx = np.zeros(u.shape[:2] + (n + 1,))
        fx[:, :, 1:n] = self.upwind(u[:, :, :-1], u[:, :, 1:], self.a1)
        gy = np.zeros((u.shape[0], n + 1, n))
        gy[:, 1:n, :] = self.upwind(u[:, :-1, :], u[:, 1:, :], self.a2)
        out += (fx[:, :, 1:] - fx[:, :, :-1]) / self.w[None, None, :]
        out += (gy[:, 1:, :] - gy[:, :-1, :]) / self.w[None, :, None]
        return out

    def fr_update(self, u, Ut, Qt, dt, alpha=None):
        """u + dt * du with F̃ = a Qt, Ũ = Ut, Rusanov jump on u (tree) or Ũ (structured).

        With ``alpha`` (one blending factor per element) the update is the
        shock-capturing blend: (1 - α_e) times the LW volume term, α_e times
        the subcell FV divergence of u, and interface fluxes
        alp·fn + (1 - alp)·F* with alp = (α_L + α_R)/2 and fn the Rusanov
        flux of u (linear advection is always admissible).
        """
        n = self.n
        F, G = self.a1 * Qt, self.a2 * Qt
        vol = self.dx(F, self.Dh) + self.dy(G, self.Dh)
        if alpha is not None:
            vol = (1.0 - alpha)[:, None, None] * vol + alpha[:, None, None] * self.subcell_fv(u)
        jump = u if self.flux_order == "tree" else Ut
        surf = np.zeros_like(u)
        l1, l2 = abs(self.a1), abs(self.a2)

        def blend(fs, ul, ur, a, L, R):
            if alpha is None:
                return fs
            alp = 0.5 * (alpha[L] + alpha[R])
            return alp * self.upwind(ul, ur, a) + (1.0 - alp) * fs

        for e in range(self.E):
            L = self.nbr(e, 0)
            fs = 0.5 * (F[L, :, n - 1] + F[e, :, 0]) - 0.5 * l1 * (jump[e, :, 0] - jump[L, :, n - 1])
            surf[e, :, 0] -= blend(fs, u[L, :, n - 1], u[e, :, 0], self.a1, L, e) / self.w[0]
            R = self.nbr(e, 1)
            fs = 0.5 * (F[e, :, n - 1] + F[R, :, 0]) - 0.5 * l1 * (jump[R, :, 0] - jump[e, :, n - 1])
            surf[e, :, n - 1] += blend(fs, u[e, :, n - 1], u[R, :, 0], self.a1, e, R) / self.w[n - 1]
            B = self.nbr(e, 2)
            gs = 0.5 * (G[B, n - 1, :] + G[e, 0, :]) - 0.5 * l2 * (jump[e, 0, :] - jump[B, n - 1, :])
            surf[e, 0, :] -= blend(gs, u[B, n - 1, :], u[e, 0, :], self.a2, B, e) / self.w[0]
            T = self.nbr(e, 3)
            gs = 0.5 * (G[e, n - 1, :] + G[T, 0, :]) - 0.5 * l2 * (jump[T, 0, :] - jump[e, n - 1, :])
            surf[e, n - 1, :] += blend(gs, u[e, n - 1, :], u[T, 0, :], self.a2, e, T) / self.w[n - 1]
        du = -self.invJ * (vol + surf)
        return u + dt * du, du

    def lw_step(self, u, dt, alpha=None):
        N = self.n - 1
        terms = [u]
        for k in range(1, N + 1):
            terms.append(self.A(terms[-1], dt))
        fact = [1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120]
        Ut = sum(fact[k] * terms[k] for k in range(N + 1))
        qf = list(fact)
        if N == 4:
            qf[4] = 1.0 / 240
        Qt = sum(qf[k] * terms[k] for k in range(N + 1))
        return self.fr_update(u, Ut, Qt, dt, alpha)

    def mdrk_step(self, u, dt, alpha=None):
        Au = self.A(u, dt)
        S1 = 0.5 * u + 0.125 * Au
        us, du1 = self.fr_update(u, S1, S1, dt, alpha)
        S2 = u + Au / 6.0 + self.A(us, dt) / 3.0
        unew, du2 = self.fr_update(u, S2, S2, dt, alpha)
        return unew, du1, du2, us

    def step_matrix(self, dt, kind="lw"):
        m = self.E * self.n * self.n
        M = np.zeros((m, m))
        for k in range(m):
            e = np.zeros(m)
            e[k] = 1.0
            u = e.reshape(self.E, self.n, self.n)
            r = self.lw_step(u, dt)[0] if kind == "lw" else self.mdrk_step(u, dt)[0]
            M[:, k] = r.ravel()
        return M


class LinearParabolicLW(LinearLW):
    """Closed-form LW / MDRK step for linear advection-diffusion, f = a u - ν ∇u
    (TrixiLW's hyperbolic-parabolic kernels, src/solvers/dgsem_tree/
    dg_2d_parabolic.jl:725-1948 and multi_derivative_2d_parabolic.jl:221-711,
    with Trixi's BR1 gradient).  With a linear flux the finite-difference time
    derivatives of both parts are exact on the Taylor states, so with
      g₀ = BR1 gradient of u (weak-form D̂ volume term, central face value,
           lifting by 1/w at the end nodes, ×J⁻¹),
      u_{k+1} = -Δt J⁻¹ (Dx (a1 u_k - ν gx_k) + Dy (a2 u_k - ν gy_k)),
      g_{k+1} = J⁻¹ (Dx, Dy) u_{k+1}      (element-local gradient of u_k),
    the LW fluxes are F̃ = Σ q_k (a u_k - ν g_k) with q = 1, 1/2, 1/6, 1/24 and,
    at N=4, 1/240 (the halved f_tttt also applies to the viscous part), and the
    interface flux is ½(F̃_L + F̃_R) - ½|a| (u_R - u_L): the viscous part enters
    as the central average.  Evaluated with dense numpy operators in gather
    form, independently of the oracle's loops.
    """

    def __init__(self, nodes, weights, D, Dhat, nx, ny, h, a=(1.5, 1.0), nu=0.05):
        super().__init__(nodes, weights, D, Dhat, nx, ny, h, a=a, flux_order="tree")
        self.nu = nu

    def neighbours(self):
        E = self.E
        return [np.array([self.nbr(e, f) for e in range(E)]) for f in range(4)]

    def br1_gradient(self, u):
        n = self.n
        L, R, B, T = self.neighbours()
        gx = self.dx(u, self.Dh)
        gy = self.dy(u, self.Dh)
        gx[:, :, 0] -= 0.5 * (u[L][:, :, n - 1] + u[:, :, 0]) / self.w[0]
        gx[:, :, n - 1] += 0.5 * (u[:, :, n - 1] + u[R][:, :, 0]) / self.w[n - 1]
        gy[:, 0, :] -= 0.5 * (u[B][:, n - 1, :] + u[:, 0, :]) / self.w[0]
        gy[:, n - 1, :] += 0.5 * (u[:, n - 1, :] + u[T][:, 0, :]) / self.w[n - 1]
        return self.invJ * gx, self.invJ * gy

    def local_gradient(self, u):
        return self.invJ * self.dx(u, self.D), self.invJ * self.dy(u, self.D)

    def advance(self, u, gx, gy, dt):
        fx = self.a1 * u - self.nu * gx
        fy = self.a2 * u - self.nu * gy
        return -dt * self.invJ * (self.dx(fx, self.D) + self.dy(fy, self.D))

    def update(self, u, Fx, Fy, dt):
        """u + dt du for the flux F̃ = (Fx, Fy) (advective minus viscous) and the
        Rusanov jump on uⁿ."""
        n = self.n
        L, R, B, T = self.neighbours()
        l1, l2 = abs(self.a1), abs(self.a2)
        vol = self.dx(Fx, self.Dh) + self.dy(Fy, self.Dh)
        surf = np.zeros_like(u)
        fsL = 0.5 * (Fx[L][:, :, n - 1] + Fx[:, :, 0]) - 0.5 * l1 * (u[:, :, 0] - u[L][:, :, n - 1])
        fsR = 0.5 * (Fx[:, :, n - 1] + Fx[R][:, :, 0]) - 0.5 * l1 * (u[R][:, :, 0] - u[:, :, n - 1])
        gsB = 0.5 * (Fy[B][:, n - 1, :] + Fy[:, 0, :]) - 0.5 * l2 * (u[:, 0, :] - u[B][:, n - 1, :])
        gsT = 0.5 * (Fy[:, n - 1, :] + Fy[T][:, 0, :]) - 0.5 * l2 * (u[T][:, 0, :] - u[:, n - 1, :])
        surf[:, :, 0] -= fsL / self.w[0]
        surf[:, :, n - 1] += fsR / self.w[n - 1]
        surf[:, 0, :] -= gsB / self.w[0]
        surf[:, n - 1, :] += gsT / self.w[n - 1]
        du = -self.invJ * (vol + surf)
        return u + dt * du, du

    def lw_step(self, u, dt, alpha=None):
        N = self.n - 1
        us_, gxs, gys = [u], [], []
        gx, gy = self.br1_gradient(u)
        gxs.append(gx)
        gys.append(gy)
        for k in range(1, N + 1):
            w = self.advance(us_[-1], gxs[-1], gys[-1], dt)
            us_.append(w)
            wx, wy = self.local_gradient(w)
            gxs.append(wx)
            gys.append(wy)
        q = [1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 240 if N == 4 else 1.0 / 120]
        Fx = sum(q[k] * (self.a1 * us_[k] - self.nu * gxs[k]) for k in range(N + 1))
        Fy = sum(q[k] * (self.a2 * us_[k] - self.nu * gys[k]) for k in range(N + 1))
        return self.update(u, Fx, Fy, dt)

    def mdrk_step(self, u, dt, alpha=None):
        gx, gy = self.br1_gradient(u)
        ut = self.advance(u, gx, gy, dt)
        utx, uty = self.local_gradient(ut)
        f = (self.a1 * u - self.nu * gx, self.a2 * u - self.nu * gy)
        ft = (self.a1 * ut - self.nu * utx, self.a2 * ut - self.nu * uty)
        F1 = [0.5 * f[d] + 0.125 * ft[d] for d in range(2)]
        us, du1 = self.update(u, F1[0], F1[1], dt)
        F2 = [f[d] + ft[d] / 6.0 for d in range(2)]
        sx, sy = self.br1_gradient(us)
        ust = self.advance(us, sx, sy, dt)
        ustx, usty = self.local_gradient(ust)
        fst = (self.a1 * ust - self.nu * ustx, self.a2 * ust - self.nu * usty)
        Fs = [F2[d] + fst[d] / 3.0 for d in range(2)]
        unew, du2 = self.update(u, Fs[0], Fs[1], dt)
        return unew, du1, du2, us
