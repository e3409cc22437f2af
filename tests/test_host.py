"""Host-side logic and the C ABI surface (no GPU)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    txt = (ROOT / "include" / "trixilw_hip.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tlw_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from trixilw_hip import _lib

    so = _lib.LIB_PATH
    assert so.exists(), "libtrixilw_hip.so not built"
    L = ctypes.CDLL(str(so))
    names = header_functions()
    assert len(names) >= 17
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTED)


def test_product_does_not_use_the_oracle():
    pkg = ROOT / "trixilw.jl_amd"
    for p in list(pkg.rglob("*.py")) + list(pkg.rglob("*.hip")) + list(pkg.rglob("*.hpp")):
        src = p.read_text()
        assert "oracle" not in src.lower() or p.name == "build.py", p


def test_treemesh_tables():
    from trixilw_hip import TreeMesh

    for ordering in ("morton", "lexicographic"):
        m = TreeMesh((-10, -10), (10, 10), 3, periodicity=True, ordering=ordering)
        E = m.n_elements
        assert E == 64 and m.h == 2.5
        nb = m.neighbors
        for f in range(4):
            assert np.all(nb[nb[:, f], f ^ 1] == np.arange(E))
        assert np.all(m.inverse_jacobian == 2 / 2.5)
        assert sorted(m.element_of_cell.tolist()) == list(range(E))
    m = TreeMesh((0, 0), (1, 1), 2, periodicity=False)
    assert (m.neighbors < 0).sum() == 4 * 4
    cx = m.element_center[:, 0]
    assert np.all(m.neighbors[cx < 0.25, 0] == -1)


def test_node_coordinates_match_oracle(oracle_mod):
    from trixilw_hip import LobattoLegendreBasis, TreeMesh

    m = TreeMesh((-10, -10), (10, 10), 2, ordering="lexicographic")
    b = LobattoLegendreBasis(4)
    x, y = m.node_coordinates(b.nodes)
    o = oracle_mod.Oracle(4, 4, 4, -10.0, -10.0, 5.0)
    xo, yo = o.node_coords()
    assert np.allclose(x, xo, rtol=0, atol=1e-14) and np.allclose(y, yo, rtol=0, atol=1e-14)


def test_cfl_mapping_and_isapprox():
    from trixilw_hip import DGSEM, isapprox, lw_fourier_cfl, trixi2lw

    assert lw_fourier_cfl(4) == 0.103 and lw_fourier_cfl(3) == 0.170
    assert trixi2lw(0.98, DGSEM(4)) == pytest.approx(0.98 * 5 * 0.103, rel=1e-15)
    assert isapprox(1.0, 1.0 + 1e-9) and not isapprox(1.0, 1.0 + 1e-7)
    assert isapprox(0.0, 0.0) and not isapprox(0.0, 1e-300)


def test_set_dt_truncates():
    from trixilw_hip.integrator import set_dt, set_t_and_iter

    class Semi:
        cache = {"dt": [0.0]}

    class It:
        t = 0.9
        tspan = (0.0, 1.0)
        p = Semi()
        iter = 0

        class stats:
            naccept = 0

    it = It()
    assert set_dt(it, 0.3) == pytest.approx(0.1)
    assert Semi.cache["dt"][0] == pytest.approx(0.1)
    set_t_and_iter(it, 0.1)
    assert it.t == pytest.approx(1.0) and it.iter == 1


def test_gaussian_random_field_is_seeded_periodic_and_scaled():
    """The differentiation configuration's perturbation: deterministic by seed,
    periodic in both directions, standard deviation = amplitude over the
    points, different for a different seed."""
    import numpy as np
    from trixilw_hip.adjoint import gaussian_random_field

    rng = np.random.default_rng(0)
    x, y = rng.uniform(-10, 10, 500), rng.uniform(-10, 10, 500)
    f = gaussian_random_field(x, y, (20.0, 20.0), seed=2025, amplitude=1e-2)
    assert np.array_equal(f, gaussian_random_field(x, y, (20.0, 20.0), seed=2025, amplitude=1e-2))
    assert abs(np.std(f) - 1e-2) < 1e-15
    raw = gaussian_random_field(np.concatenate([x, x + 20.0, x]), np.concatenate([y, y, y - 40.0]),
                                (20.0, 20.0), seed=2025, amplitude=1.0)
    assert np.allclose(raw[:500], raw[500:1000], atol=1e-12) and np.allclose(raw[:500], raw[1000:], atol=1e-12)
    g = gaussian_random_field(x, y, (20.0, 20.0), seed=2026, amplitude=1e-2)
    assert np.max(np.abs(f - g)) > 1e-3


def test_wall_functions_are_checked_on_the_wall_face_nodes():
    """A NoSlip/Isothermal function reaches the device as one constant: it is
    evaluated on the face nodes of the walls it is used on, and one that varies
    there (even if it agrees at a few fixed points) is rejected."""
    import pytest
    import trixilw_hip as tlw
    from trixilw_hip.semidiscretization import _wall_face_points

    mesh = tlw.TreeMesh((0.0, 0.0), (1.0, 1.0), 3, periodicity=False)
    nodes = tlw.LobattoLegendreBasis(4).nodes
    top = _wall_face_points(mesh, 3, nodes)
    assert len(top) == 8 * 5 and all(abs(y - 1.0) < 1e-15 for _, y in top)
    assert tlw.NoSlip(lambda x, t, e: (1.0, 0.0)).resolve(top) == (1.0, 0.0)
    assert tlw.NoSlip((0.5, 0.0)).resolve(top) == (0.5, 0.0)
    # agrees at (0, 0) and (0.37, -0.61) but not along the lid
    sneaky = tlw.NoSlip(lambda x, t, e: (1.0 if x[0] < 0.5 or x[0] == 0.37 else 2.0, 0.0))
    with pytest.raises(NotImplementedError):
        sneaky.resolve(top)
    with pytest.raises(NotImplementedError):
        tlw.Isothermal(lambda x, t, e: 1.0 + t).resolve(top)
    assert tlw.Isothermal(lambda x, t, e: 1.2).resolve(top) == 1.2
    eq = tlw.CompressibleEulerEquations2D(1.4)
    eqp = tlw.CompressibleNavierStokesDiffusion2D(eq, mu=1e-3, Prandtl=0.72,
                                                 gradient_variables=tlw.GradientVariablesConservative())
    rest = tlw.BoundaryConditionNavierStokesWall(tlw.NoSlip((0.0, 0.0)), tlw.Adiabatic(0.0))
    lid = tlw.BoundaryConditionNavierStokesWall(sneaky, tlw.Adiabatic(0.0))
    solver = tlw.DGSEM(4)
    with pytest.raises(NotImplementedError):  # raised on the host, before any device call
        tlw.SemidiscretizationHyperbolicParabolic(
            mesh, tlw.LW(), (eq, eqp), tlw.ConstantState(), solver,
            boundary_conditions=((tlw.boundary_condition_slip_wall_vertical,) * 2
                                 + (tlw.boundary_condition_slip_wall_horizontal,) * 2, (rest, rest, rest, lid)))
