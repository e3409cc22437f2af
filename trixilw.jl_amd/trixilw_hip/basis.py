"""Gauss-Lobatto-Legendre basis on the host (Trixi's LobattoLegendreBasis).

TrixiLW reads the fields ``nodes``, ``weights``, ``inverse_weights``,
``derivative_matrix``, ``derivative_dhat`` and ``boundary_interpolation``
(src/solvers/dgsem_tree/dg_2d.jl:932, 1211-1226; Trixi's
calc_surface_integral!).  A Julia host hands Trixi's own arrays to
``tlw_create``; this Python host computes them the same way: GLL nodes by
Newton iteration on L_{N+1} - L_{N-1} (in extended precision), barycentric
weights, the negative-sum-diagonal derivative matrix and
D̂ = -W⁻¹ Dᵀ W.
"""
from __future__ import annotations

import numpy as np


def _q_and_l(N: int, x):
    one = np.longdouble(1)
    Lm2, Lm1, dLm2, dLm1 = one, x, np.longdouble(0), one
    Lk, dLk = x, one
    for k in range(2, N + 1):
        Lk = (2 * k - 1) * x * Lm1 / k - (k - 1) * Lm2 / k
        dLk = dLm2 + (2 * k - 1) * Lm1
        Lm2, Lm1, dLm2, dLm1 = Lm1, Lk, dLm1, dLk
    if N == 1:
        Lk, dLk, Lm2, dLm2 = x, one, one, np.longdouble(0)
    k = N + 1
    Lk1 = (2 * k - 1) * x * Lk / k - (k - 1) * Lm2 / k
    dLk1 = dLm2 + (2 * k - 1) * Lk
    return Lk1 - Lm2, dLk1 - dLm2, Lk


def gauss_lobatto_nodes_weights(n_nodes: int):
    N = n_nodes - 1
    x = np.zeros(n_nodes, dtype=np.longdouble)
    w = np.zeros(n_nodes, dtype=np.longdouble)
    x[0], x[N] = -1, 1
    w[0] = w[N] = np.longdouble(2) / (N * (N + 1))
    pi = np.longdouble(np.pi)
    for j in range(1, (N + 1) // 2):
        xx = -np.cos((j + np.longdouble(0.25)) * pi / N - 3 / (8 * N * pi * (j + np.longdouble(0.25))))
        for _ in range(100):
            q, dq, _LN = _q_and_l(N, xx)
            d = -q / dq
            xx = xx + d
            if abs(d) <= np.longdouble(1e-19) * abs(xx) + np.longdouble(1e-30):
                break
        q, dq, LN = _q_and_l(N, xx)
        x[j], x[N - j] = xx, -xx
        w[j] = w[N - j] = 2 / (N * (N + 1) * LN * LN)
    if N % 2 == 0:
        q, dq, LN = _q_and_l(N, np.longdouble(0))
        x[N // 2] = 0
        w[N // 2] = 2 / (N * (N + 1) * LN * LN)
    return x.astype(np.float64), w.astype(np.float64)


def barycentric_weights(nodes):
    n = len(nodes)
    wb = np.ones(n)
    for j in range(1, n):
        for k in range(j):
            wb[k] *= nodes[k] - nodes[j]
            wb[j] *= nodes[j] - nodes[k]
    return 1.0 / wb


def polynomial_derivative_matrix(nodes):
    n = len(nodes)
    wb = barycentric_weights(nodes)
    D = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if j != i:
                D[i, j] = wb[j] / wb[i] * 1.0 / (nodes[i] - nodes[j])
                D[i, i] -= D[i, j]
    return D


def calc_dhat(nodes, weights):
    D = polynomial_derivative_matrix(nodes)
    n = len(nodes)
    Dh = D.T.copy()
    for j in range(n):
        for k in range(n):
            Dh[j, k] *= -weights[k] / weights[j]
    return Dh


def legendre_normalized(m: int, x):
    """sqrt(m + 1/2) L_m(x) (Trixi legendre_polynomial_and_derivative)."""
    x = np.asarray(x, dtype=np.longdouble)
    Lm2, Lm1 = np.ones_like(x), x
    if m == 0:
        L = np.ones_like(x)
    elif m == 1:
        L = x
    else:
        for k in range(2, m + 1):
            L = ((2 * k - 1) * x * Lm1 - (k - 1) * Lm2) / k
            Lm2, Lm1 = Lm1, L
    return L * np.sqrt(np.longdouble(m) + np.longdouble(0.5))


def vandermonde_legendre(nodes):
    n = len(nodes)
    V = np.zeros((n, n), dtype=np.longdouble)
    for m in range(n):
        V[:, m] = legendre_normalized(m, np.asarray(nodes, dtype=np.longdouble))
    return V.astype(np.float64), np.linalg.inv(V.astype(np.float64))


class LobattoLegendreBasis:
    def __init__(self, polydeg: int):
        self.polydeg = int(polydeg)
        n = self.polydeg + 1
        self.nodes, self.weights = gauss_lobatto_nodes_weights(n)
        self.inverse_weights = 1.0 / self.weights
        self.derivative_matrix = polynomial_derivative_matrix(self.nodes)
        self.derivative_dhat = calc_dhat(self.nodes, self.weights)
        # boundary_interpolation[:, 1|2] = lhat(∓1) = e_{1|n} / w
        self.boundary_interpolation = np.zeros((n, 2))
        self.boundary_interpolation[0, 0] = 1.0 / self.weights[0]
        self.boundary_interpolation[n - 1, 1] = 1.0 / self.weights[n - 1]
        self.vandermonde_legendre, self.inverse_vandermonde_legendre = vandermonde_legendre(self.nodes)

    def nnodes(self) -> int:
        return self.polydeg + 1
