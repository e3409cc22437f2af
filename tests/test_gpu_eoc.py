"""Convergence order of the device solver on the isentropic vortex (the
reference's verification: L2 error of the density at t = 1 against the exact
translate, utils/plot_convergence.jl:66-69), 16² -> 128² elements, polydeg 3
and 4, LW and MDRK, CFL 0.98 of the LW limit (half of it for MDRK), the
device loop to t = 1 (last step truncated).  On 16²..64² the errors equal the
oracle's (whose convergence is pinned on the CPU); 128² extends the study
beyond what the oracle runs in the CPU suite."""
import numpy as np
import pytest

from gpu_helpers import torch_mod

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,td,slope_min", [(4, "lw", 3.7), (3, "lw", 3.2), (3, "mdrk", 3.2), (4, "mdrk", 3.5)])
def test_isentropic_vortex_eoc_on_device(oracle_mod, N, td, slope_min):
    import trixilw_hip as tlw

    torch = torch_mod()
    ic = tlw.IsentropicVortex()
    eq = tlw.CompressibleEulerEquations2D(1.4)
    errs, errs_o = [], []
    for level in (4, 5, 6, 7):
        mesh = tlw.TreeMesh((-10.0, -10.0), (10.0, 10.0), level, ordering="strips")
        solver = tlw.DGSEM(N, volume_integral=tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW()))
        semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.get_time_discretization(solver), eq, ic, solver)
        cfl = tlw.trixi2lw(0.98, solver) * (0.5 if td == "mdrk" else 1.0)
        upd = tlw.semidiscretize(semi, tlw.get_time_discretization(solver), (0.0, 1.0))
        sol = tlw.solve_lwfr(upd, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl))
        assert abs(sol.t - 1.0) < 1e-12
        x, y = semi.node_coordinates()
        exact = torch.from_numpy(np.ascontiguousarray(ic(x, y, sol.t)[..., 0])).cuda()
        errs.append(float(torch.sqrt(torch.mean((sol.u[..., 0] - exact) ** 2))))
        semi.close()
        if level <= 6:
            nx = 2 ** level
            o = oracle_mod.Oracle(N, nx, nx, -10.0, -10.0, 20.0 / nx)
            xo, yo = o.node_coords()
            u, t, _ = o.solve(ic(xo, yo, 0.0), 0.0, 1.0, cfl, mdrk=(td == "mdrk"))
            errs_o.append(float(np.sqrt(np.mean((u - ic(xo, yo, t))[..., 0] ** 2))))
    eocs = [np.log(errs[k] / errs[k + 1]) / np.log(2) for k in range(3)]
    slope = -np.polyfit(np.log([16, 32, 64, 128]), np.log(errs), 1)[0]
    print(f"N={N} {td}: L2 errors {errs}, EOC {eocs}, fitted order {slope:.3f}", flush=True)
    assert np.allclose(errs[:3], errs_o, rtol=1e-9, atol=0), (errs, errs_o)
    assert all(e > 3.0 for e in eocs), eocs
    assert slope > slope_min, (errs, eocs, slope)
