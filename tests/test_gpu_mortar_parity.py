"""GPU parity of the TreeMesh with refined cells (L2 mortars, dg_2d.jl:1456-1646)
and of AMR against the oracle:

- rhs (LW N=3, 4) and both MDRK stages on periodic meshes with a refined box,
  and with the time-averaged Dirichlet vortex on a non-periodic mesh, within
  1e-12 relative; K-step device-loop states (LW and MDRK);
- free stream and conservation over 20 device-loop steps on a 128² mesh with
  two nested refined boxes;
- the AMR solution transfer (refine and coarsen) against the host restatement
  with the oracle's mortar operators, and an AMR step in the middle of a
  solve against the oracle doing the same;
- the configurations refused on such meshes.

The oracle takes the product's leaves in the product's element order with its
own geometric connectivity (oracle/tree_mesh.py); inputs are seeded."""
import numpy as np
import pytest

from amr_helpers import transfer
from gpu_helpers import assert_close, torch_mod

pytestmark = pytest.mark.gpu

BOX = ({"type": "box", "coordinates_min": (-5.0, -5.0), "coordinates_max": (2.6, 2.6)},)


def oracle_for(oracle_mod, mesh, N, periodic, **kw):
    import tree_mesh

    lo = mesh.coordinates_min
    nbr, mort = tree_mesh.connectivity(mesh.element_center[:, 0], mesh.element_center[:, 1], mesh.element_length,
                                       lo, mesh.length, periodic)
    assert np.array_equal(nbr, mesh.neighbors)
    return oracle_mod.Oracle(N, 1, 1, 0.0, 0.0, 1.0, gamma=1.4,
                             tree=(mesh.element_center[:, 0], mesh.element_center[:, 1], mesh.element_length, nbr,
                                   mort), **kw)


def setup_refined(oracle_mod, N, level=3, td="lw", periodic=True, patches=BOX, seed=0, amp=1e-3):
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(1.4)
    ic = tlw.IsentropicVortex(center=(0.7, -0.4), period=(20.0, 20.0) if periodic else (1e9, 1e9))
    mesh = tlw.TreeMesh((-10.0, -10.0), (10.0, 10.0), level, periodicity=periodic, refinement_patches=patches)
    assert len(mesh.mortars) > 0
    solver = tlw.DGSEM(N, volume_integral=tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW()))
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.get_time_discretization(solver), eq, ic, solver,
                                            boundary_conditions=None if periodic else ic)
    okw = {} if periodic else dict(bc_kind=ic.kind, bc_par=ic.params)
    o = oracle_for(oracle_mod, mesh, N, (periodic, periodic), **okw)
    x, y = o.node_coords()
    xm, ym = mesh.node_coordinates(semi.solver.basis.nodes)
    assert np.max(np.abs(x - xm)) < 1e-13 and np.max(np.abs(y - ym)) < 1e-13
    rng = np.random.default_rng(seed)
    u = ic(x, y, 0.0) * (1.0 + amp * rng.standard_normal(x.shape + (4,)))
    return semi, o, u, mesh


def dev(u):
    torch = torch_mod()
    return torch.from_numpy(np.ascontiguousarray(u)).cuda()


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("periodic", [True, False])
def test_refined_rhs_lw_parity(oracle_mod, N, periodic):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"refined rhs LW N={N} periodic={periodic}", flush=True)
    semi, o, u_np, mesh = setup_refined(oracle_mod, N, periodic=periodic, seed=N)
    u = dev(u_np)
    du = torch.empty_like(u)
    dt = 0.5 * o.max_dt(u_np)
    assert semi.max_dt(u) == pytest.approx(o.max_dt(u_np), rel=1e-13)
    semi.cache["dt"][0] = dt
    du_ref = o.rhs(u_np, 0.3, dt)
    tlw.rhs(du, u, semi, 0.3)
    torch.cuda.synchronize()
    assert_close(du.cpu().numpy(), du_ref, what=f"refined LW N={N}")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_refined_rhs_mdrk_parity(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"refined rhs MDRK N={N}", flush=True)
    semi, o, u_np, mesh = setup_refined(oracle_mod, N, td="mdrk", seed=10 + N)
    u = dev(u_np)
    du = torch.empty_like(u)
    dt = 0.4 * o.max_dt(u_np)
    semi.cache["dt"][0] = dt
    d1 = o.rhs(u_np, 0.0, dt, kind="mdrk1")
    us = u_np + 0.5 * dt * d1
    o.set_us(us)
    d2 = o.rhs(u_np, 0.0, dt, kind="mdrk2")
    tlw.rhs_mdrk1(du, u, semi, 0.0)
    assert_close(du.cpu().numpy(), d1, what=f"refined mdrk1 N={N}")
    semi.set_us(dev(us))
    tlw.rhs_mdrk2(du, u, semi, 0.0)
    torch.cuda.synchronize()
    assert_close(du.cpu().numpy(), d2, what=f"refined mdrk2 N={N}")
    semi.close()


@pytest.mark.parametrize("N,td,K,periodic", [(4, "lw", 25, True), (3, "mdrk", 15, True), (4, "lw", 20, False)])
def test_refined_device_loop_k_steps(oracle_mod, N, td, K, periodic):
    import trixilw_hip as tlw

    print(f"refined device loop N={N} {td} K={K} periodic={periodic}", flush=True)
    semi, o, u_np, mesh = setup_refined(oracle_mod, N, level=4, td=td, periodic=periodic)
    cfl = tlw.trixi2lw(0.9, semi.solver) * (0.5 if td == "mdrk" else 1.0)
    u_ref, t_ref, _ = o.solve(u_np, 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(dev(u_np))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(sol.u.cpu().numpy(), u_ref, what=f"refined {td} N={N} after {K} steps")
    semi.close()


def test_refined_free_stream_and_conservation_at_scale():
    """128² periodic mesh with two nested refined boxes (three levels, mortars
    of both orientations and large sides): a uniform flow stays uniform, and a
    perturbed vortex keeps Σ_e Σ_ij w_i w_j u / J⁻¹_e over 20 steps."""
    import trixilw_hip as tlw

    torch = torch_mod()
    print("refined 128^2 with nested boxes, 20 steps", flush=True)
    patches = ({"type": "box", "coordinates_min": (-6.0, -4.0), "coordinates_max": (5.0, 6.0)},
               {"type": "box", "coordinates_min": (-2.0, -1.0), "coordinates_max": (1.5, 2.5)})
    mesh = tlw.TreeMesh((-10.0, -10.0), (10.0, 10.0), 7, refinement_patches=patches)
    eq = tlw.CompressibleEulerEquations2D(1.4)
    ic = tlw.IsentropicVortex(center=(0.2, 0.3), period=(20.0, 20.0))
    solver = tlw.DGSEM(4, volume_integral=tlw.VolumeIntegralFR(tlw.LW()))
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.get_time_discretization(solver), eq, ic, solver)
    nodes, w = semi.solver.basis.nodes, semi.solver.basis.weights
    x, y = mesh.node_coordinates(nodes)
    assert len(mesh.mortars) > 100 and mesh.element_length.min() < 0.25 * mesh.element_length.max() * 1.01
    cfl = tlw.trixi2lw(0.9, semi.solver)
    for kind in ("free_stream", "vortex"):
        if kind == "free_stream":
            u0 = tlw.ConstantState(1.2, 0.3, -0.5, 0.9)(x, y, 0.0)
        else:
            rng = np.random.default_rng(8)
            u0 = ic(x, y, 0.0) * (1 + 1e-2 * rng.standard_normal(x.shape + (1,)))
        lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
        lw_update.u0.copy_(dev(u0))
        sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=20)
        torch.cuda.synchronize()
        u1 = sol.u.cpu().numpy()
        assert sol.stats.naccept == 20 and np.all(np.isfinite(u1))
        if kind == "free_stream":
            assert np.max(np.abs(u1 - u0)) < 1e-12
        else:
            W = np.outer(w, w)[None] * (0.5 * mesh.element_length)[:, None, None] ** 2
            m0, m1 = np.einsum("eji,ejiv->v", W, u0), np.einsum("eji,ejiv->v", W, u1)
            assert np.all(np.abs(m1 - m0) <= 1e-13 * np.einsum("eji,ejiv->v", W, np.abs(u0))), (m1 - m0)
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_amr_transfer_matches_host(oracle_mod, N):
    """Device refine and coarsen of a random state against the host restatement."""
    import trixilw_hip as tlw

    print(f"AMR transfer N={N}", flush=True)
    semi, o, u_np, mesh = setup_refined(oracle_mod, N)
    F, R = (lambda m: ((m[0], m[1]), (m[2], m[3])))(o.mortar_matrices())
    rng = np.random.default_rng(N)
    u_np = u_np * (1 + 0.1 * rng.standard_normal(u_np.shape))
    fine, plan = mesh.refine(np.arange(0, mesh.n_elements, 5))
    u_f = semi.adapt(dev(u_np), fine, plan)
    assert_close(u_f.cpu().numpy(), transfer(u_np, plan, F, R), tol=1e-14, what="refine")
    back, plan2 = fine.coarsen(np.nonzero(plan[:, 0] == 1)[0])
    assert np.count_nonzero(plan2[:, 0] == 2) > 0
    u_b = semi.adapt(u_f, back, plan2)
    assert_close(u_b.cpu().numpy(), transfer(u_f.cpu().numpy(), plan2, F, R), tol=1e-14, what="coarsen")
    semi.close()


@pytest.mark.parametrize("td", ["lw", "mdrk"])
def test_amr_step_inside_a_solve_matches_oracle(oracle_mod, td):
    """10 steps, an AMR refinement around the vortex core (device transfer and
    mesh re-upload), 10 more steps; the oracle runs the same with the host
    transfer and a new solver on the new mesh."""
    import trixilw_hip as tlw

    print(f"AMR inside a solve {td}", flush=True)
    N = 4 if td == "lw" else 3
    semi, o, u_np, mesh = setup_refined(oracle_mod, N, level=3, td=td)
    cfl = tlw.trixi2lw(0.9, semi.solver) * (0.5 if td == "mdrk" else 1.0)
    u1_ref, t1_ref, _ = o.solve(u_np, 0.0, 1e9, cfl, max_steps=10, mdrk=(td == "mdrk"))
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(dev(u_np))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=10)
    assert abs(sol.t - t1_ref) <= 1e-13 * t1_ref
    assert_close(sol.u.cpu().numpy(), u1_ref, what=f"{td} before AMR")
    # refine the elements whose centre is within 3 of the vortex core
    r = np.hypot(mesh.element_center[:, 0] - 0.7, mesh.element_center[:, 1] + 0.4)
    fine, plan = mesh.refine(np.nonzero(r < 3.0)[0])
    assert fine.n_elements > mesh.n_elements
    u_f = semi.adapt(sol.u, fine, plan)
    F, R = (lambda m: ((m[0], m[1]), (m[2], m[3])))(o.mortar_matrices())
    u_f_ref = transfer(u1_ref, plan, F, R)
    assert_close(u_f.cpu().numpy(), u_f_ref, what="transfer")
    o2 = oracle_for(oracle_mod, fine, N, (True, True))
    u2_ref, t2_ref, _ = o2.solve(u_f_ref, t1_ref, 1e9, cfl, max_steps=10, mdrk=(td == "mdrk"))
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (sol.t, float("inf")))
    lw_update.u0.copy_(u_f)
    sol2 = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=10)
    assert sol2.stats.naccept == 10
    assert abs(sol2.t - t2_ref) <= 1e-13 * t2_ref
    assert_close(sol2.u.cpu().numpy(), u2_ref, what=f"{td} after AMR")
    semi.close()


def test_refined_mesh_refuses_unsupported(oracle_mod):
    import trixilw_hip as tlw
    from trixilw_hip import _lib

    torch = torch_mod()
    eq = tlw.CompressibleEulerEquations2D(1.4)
    ic = tlw.IsentropicVortex()
    mesh = tlw.TreeMesh((-10.0, -10.0), (10.0, 10.0), 2, refinement_patches=BOX)
    ind = tlw.IndicatorHennemannGassner(eq, None, alpha_max=0.5, alpha_min=0.001, alpha_smooth=True)
    vi = tlw.VolumeIntegralFRShockCapturing(ind, volume_integralFR=tlw.VolumeIntegralFR(tlw.LW()))
    with pytest.raises(_lib.TLWError) as ei:
        tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, ic, tlw.DGSEM(4, volume_integral=vi))
    assert ei.value.code == _lib.TLW_ENOTSUP
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, ic, tlw.DGSEM(4, volume_integral=tlw.VolumeIntegralFR(tlw.LW())))
    x, y = mesh.node_coordinates(semi.solver.basis.nodes)
    u = dev(ic(x, y, 0.0))
    with pytest.raises(_lib.TLWError) as ei:
        tlw.adjoint_step(torch.empty_like(u), u, u.clone(), semi, 0.0, 1e-3)
    assert ei.value.code == _lib.TLW_ENOTSUP
    # a plan that points outside the old mesh is refused before any launch
    bad = np.zeros((mesh.n_elements, 5), dtype=np.int32)
    bad[0] = (0, mesh.n_elements, 0, 0, 0)
    with pytest.raises(_lib.TLWError) as ei:
        semi.adapt(u, mesh, bad)
    assert ei.value.code == _lib.TLW_EINVAL
    semi.close()
