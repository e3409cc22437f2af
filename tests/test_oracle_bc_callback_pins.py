"""Oracle pins of the caller-defined Dirichlet boundary (BoundaryConditionDirichlet
with any boundary_value_function, src/equations/equations.jl:6-37): the
callback path evaluates the same time average as the built-in functions, so
a Python restatement of the isentropic vortex through the callback gives the
oracle's built-in vortex boundary on a non-periodic TreeMesh and curvilinear
mesh, LW and both MDRK stages, to round-off."""
import numpy as np
import pytest

from conftest import rel_linf


def _vortex_py(ic):
    return lambda x, y, t: ic(np.float64(x), np.float64(y), t)


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("flux_order", [0, 1])
def test_callback_equals_builtin_vortex_tree(oracle_mod, N, flux_order):
    import trixilw_hip as tlw

    ic = tlw.IsentropicVortex(center=(0.3, 0.2), period=(1e6, 1e6))
    args = (N, 4, 4, -5.0, -5.0, 2.5)
    kw = dict(periodic=(False, False), flux_order=flux_order, nthreads=1)
    ob = oracle_mod.Oracle(*args, bc_kind=oracle_mod.BV_VORTEX, bc_par=ic.params, **kw)
    oc = oracle_mod.Oracle(*args, bc_kind=oracle_mod.BV_CALLBACK, bc_fn=_vortex_py(ic), **kw)
    x, y = ob.node_coords()
    u = ic(x, y, 0.0) * (1 + 1e-3 * np.random.default_rng(N).standard_normal(x.shape + (4,)))
    dt, t = 0.4 * ob.max_dt(u), 0.3
    for kind in ("lw", "mdrk1"):
        a, b = ob.rhs(u, t, dt, kind=kind), oc.rhs(u, t, dt, kind=kind)
        assert rel_linf(b, a) <= 1e-14, kind
    # the boundary data matters: a different function changes the boundary elements' residual
    oz = oracle_mod.Oracle(*args, bc_kind=oracle_mod.BV_CALLBACK,
                           bc_fn=lambda x, y, t: ic(np.float64(x), np.float64(y), t + 0.5), **kw)
    assert rel_linf(oz.rhs(u, t, dt), ob.rhs(u, t, dt)) > 1e-6


def test_callback_equals_builtin_vortex_curved(oracle_mod):
    import trixilw_hip as tlw

    L = 20.0
    ic = tlw.IsentropicVortex(center=(L / 2, L / 2), period=(1e9, 1e9))
    kw = dict(gamma=1.4, mapping=(oracle_mod.MAP_WARP, (L, 0.05)), periodic=(False, False), nthreads=1)
    ob = oracle_mod.Oracle(4, 4, 4, 0.0, 0.0, L / 4, bc_kind=oracle_mod.BV_VORTEX, bc_par=ic.params, **kw)
    oc = oracle_mod.Oracle(4, 4, 4, 0.0, 0.0, L / 4, bc_kind=oracle_mod.BV_CALLBACK, bc_fn=_vortex_py(ic), **kw)
    x, y = ob.node_coords()
    u = ic(x, y, 0.0)
    dt = 0.3 * ob.max_dt(u)
    assert rel_linf(oc.rhs(u, 0.1, dt), ob.rhs(u, 0.1, dt)) <= 1e-14
