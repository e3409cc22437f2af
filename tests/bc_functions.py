"""Boundary value functions of the caller-defined Dirichlet tests: smooth,
time dependent, and none of them a built-in boundary kind of the library."""
import numpy as np

GAMMA = 1.4


def waves(x, t, equations=None):
    """ρ, v1, v2, p as travelling and standing waves on [0, 1]²."""
    x1, x2 = x
    rho = 1.0 + 0.2 * np.sin(np.pi * (x1 + x2) - 1.3 * t)
    v1 = 0.3 + 0.1 * np.cos(np.pi * x1) * np.cos(0.9 * t)
    v2 = -0.2 + 0.1 * np.sin(np.pi * x2 + t)
    p = 1.0 + 0.1 * np.cos(np.pi * (x1 - x2) + 0.7 * t)
    return np.stack([rho, rho * v1, rho * v2, p / (GAMMA - 1) + 0.5 * rho * (v1 * v1 + v2 * v2)], axis=-1)


def steady_shear(x, t, equations=None):
    """A steady, spatially varying inflow state (no t dependence)."""
    x1, x2 = x
    rho = 1.0 + 0.1 * np.sin(2.0 * np.pi * x2)
    v1 = 0.5 + 0.05 * np.cos(np.pi * x2)
    v2 = 0.05 * np.sin(np.pi * x1)
    p = 1.0 + 0.05 * x1
    return np.stack([rho, rho * v1, rho * v2, p / (GAMMA - 1) + 0.5 * rho * (v1 * v1 + v2 * v2)], axis=-1)
