"""Batched instances (tlw_create_batch): each instance of the batch matches the
oracle's solve of that instance alone, step for step, and the graph-replayed
loop gives the same bits as the launched one."""
import numpy as np
import pytest

from gpu_helpers import assert_close, torch_mod

pytestmark = pytest.mark.gpu


def make_batch(N, td, levels, params):
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(1.4)
    meshes = [tlw.TreeMesh((-10.0, -10.0), (10.0, 10.0), lv, ordering="morton") for lv in levels]
    ics = [tlw.IsentropicVortex(strength=s, mach=m) for s, m in params]
    solver = tlw.DGSEM(N, volume_integral=tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW()))
    batch = tlw.SemidiscretizationBatch(meshes, tlw.get_time_discretization(solver), eq, ics, solver)
    return batch, meshes, ics


@pytest.mark.parametrize("N,td", [(4, "lw"), (3, "mdrk")])
def test_batch_instances_match_single_runs(oracle_mod, N, td):
    import trixilw_hip as tlw

    torch_mod()
    levels = [3, 4, 2, 3]
    params = [(5.0, 0.5), (3.0, 0.3), (2.0, 0.8), (4.0, 0.6)]
    batch, meshes, ics = make_batch(N, td, levels, params)
    cfl = tlw.trixi2lw(0.98, batch.solver) * (0.5 if td == "mdrk" else 1.0)
    K = 15
    u0 = batch.compute_coefficients()
    u, status = batch.solve(u0, cfl, K)
    assert np.all(status["iter"] == K)
    for k, (mesh, ic) in enumerate(zip(meshes, ics)):
        nx = mesh.nx
        o = oracle_mod.Oracle(N, nx, nx, -10.0, -10.0, 20.0 / nx)
        x, y = o.node_coords()
        u_ref, t_ref, _ = o.solve(ic(x, y, 0.0), 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
        got = batch.split(u)[k][mesh.lexicographic_permutation()]
        assert abs(status["t"][k] - t_ref) <= 1e-13 * t_ref
        assert_close(got, u_ref, what=f"instance {k} ({nx}x{nx})")
    # the instances took different steps (their own CFL reductions)
    assert len(set(np.round(status["dt"], 14))) > 1
    batch.close()


def test_graph_replay_matches_launches(oracle_mod):
    import trixilw_hip as tlw

    torch = torch_mod()
    batch, meshes, ics = make_batch(4, "lw", [3, 2, 4], [(5.0, 0.5), (2.5, 0.7), (4.0, 0.4)])
    cfl = tlw.trixi2lw(0.98, batch.solver)
    u0 = batch.compute_coefficients()
    stream = torch.cuda.Stream()
    u_a, st_a = batch.solve(u0, cfl, 24, chunk=8, graphs=False, stream=stream)
    u_b, st_b = batch.solve(u0, cfl, 24, chunk=8, graphs=True, stream=stream)
    assert np.array_equal(u_a, u_b)
    assert np.array_equal(st_a["t"], st_b["t"])
    batch.close()


def test_c4_draw_every_instance_matches_oracle(oracle_mod):
    """The batching configuration's own 64-instance draw (seed 2024: meshes
    16²..128², polydeg 3/4, Mach U[0.2,0.8], strength U[2,5]), each polydeg
    group one batch context as the benchmark runs it (Morton order, graph
    replay): every instance after 10 steps matches its own oracle solve,
    state and time, ≤ 1e-12."""
    import torch
    import trixilw_hip as tlw

    torch_mod()
    K = 10
    levels, degs, machs, strengths = tlw.draw_vortex_ensemble(64, 2024)
    assert sorted(set(levels.tolist())) == [4, 5, 6, 7] and sorted(set(degs.tolist())) == [3, 4]
    eq = tlw.CompressibleEulerEquations2D(1.4)
    checked = 0
    for N in (3, 4):
        idx = np.nonzero(degs == N)[0]
        meshes = [tlw.TreeMesh((-10.0, -10.0), (10.0, 10.0), int(levels[k]), ordering="morton") for k in idx]
        ics = [tlw.IsentropicVortex(strength=float(strengths[k]), mach=float(machs[k])) for k in idx]
        solver = tlw.DGSEM(N)
        batch = tlw.SemidiscretizationBatch(meshes, tlw.LW(), eq, ics, solver)
        cfl = tlw.trixi2lw(0.98, solver)
        u, status = batch.solve(batch.compute_coefficients(), cfl, K, graphs=True, stream=torch.cuda.Stream())
        assert np.all(status["iter"] == K)
        for k, (mesh, ic) in enumerate(zip(meshes, ics)):
            nx = mesh.nx
            o = oracle_mod.Oracle(N, nx, nx, -10.0, -10.0, 20.0 / nx)
            x, y = o.node_coords()
            u_ref, t_ref, _ = o.solve(ic(x, y, 0.0), 0.0, 1e9, cfl, max_steps=K)
            got = batch.split(u)[k][mesh.lexicographic_permutation()]
            assert abs(status["t"][k] - t_ref) <= 1e-13 * t_ref, (N, k)
            assert_close(got, u_ref, what=f"N={N} instance {k} ({nx}x{nx})")
            checked += 1
        batch.close()
    assert checked == 64


def _c4_groups(tlw, n_instances=64, seed=2024, ordering="morton"):
    levels, degs, machs, strengths = tlw.draw_vortex_ensemble(n_instances, seed)
    eq = tlw.CompressibleEulerEquations2D(1.4)
    out = {}
    for N in (3, 4):
        idx = np.nonzero(degs == N)[0]
        meshes = [tlw.TreeMesh((-10.0, -10.0), (10.0, 10.0), int(levels[k]), ordering=ordering) for k in idx]
        ics = [tlw.IsentropicVortex(strength=float(strengths[k]), mach=float(machs[k])) for k in idx]
        solver = tlw.DGSEM(N)
        out[N] = (tlw.SemidiscretizationBatch(meshes, tlw.LW(), eq, ics, solver), tlw.trixi2lw(0.98, solver))
    return out


@pytest.mark.parametrize("tiles", [(16, 10), (8, 5)])
@pytest.mark.parametrize("graphs", [False, True])
def test_ensemble_one_launch_matches_separate_batches(oracle_mod, tiles, graphs):
    """The 64-instance draw as one ensemble (one volume, surface and finalize
    launch per step for both polynomial degrees) gives the same bits, states
    and times, as each polydeg group advanced by its own batch context; and
    one instance of each degree matches the oracle."""
    import torch
    import trixilw_hip as tlw

    torch_mod()
    K = 12
    groups = _c4_groups(tlw)
    stream = torch.cuda.Stream()
    ref = {}
    for N, (b, _) in groups.items():
        b.set_stream(stream)
        tlw._lib.check(tlw._lib.lib().tlw_set_elements_per_workgroup(b.handle, tiles[0] if N == 3 else tiles[1]),
                       b.handle)
        u_sep, st_sep = b.solve(b.compute_coefficients(), groups[N][1], K, graphs=False, stream=stream)
        ref[N] = (u_sep, st_sep)
    ens = tlw.SemidiscretizationEnsemble([groups[4][0], groups[3][0]])
    ens.set_tiles(*tiles)
    states, statuses = ens.solve([groups[3][0].compute_coefficients(), groups[4][0].compute_coefficients()],
                                 [groups[3][1], groups[4][1]], K, stream=stream, graphs=graphs, chunk=5)
    for N, u, st in zip((3, 4), states, statuses):
        assert np.all(st["iter"] == K)
        assert np.array_equal(u, ref[N][0]), N
        assert np.array_equal(st["t"], ref[N][1]["t"]), N
    for N, u in zip((3, 4), states):
        b = groups[N][0]
        mesh, ic = b.meshes[0], b.initial_conditions[0]
        nx = mesh.nx
        o = oracle_mod.Oracle(N, nx, nx, -10.0, -10.0, 20.0 / nx)
        x, y = o.node_coords()
        u_ref, _, _ = o.solve(ic(x, y, 0.0), 0.0, 1e9, groups[N][1], max_steps=K)
        assert_close(b.split(u)[0][mesh.lexicographic_permutation()], u_ref, what=f"ensemble N={N} instance 0")
    ens.close()
    for b, _ in groups.values():
        b.close()
