"""Throughput of the MI355X LW-FR step (FP64 DOF-updates/s) with roofline and CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu-baseline]

A step is one full LW step of solve_lwfr's loop on the device (CFL dt from the
current state, volume kernel, surface kernel, update, dt/t update), over the
whole mesh.  Inputs are resident in HBM before the timed region.  Prints one
JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "trixilw.jl_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

METRIC = "FP64 DOF-updates/s on 1 MI355X, % HBM roofline; adjoint/forward step ratio"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # isentropic vortex (strength 5, Mach 0.5), periodic [-10,10]^2 TreeMesh, Rusanov, CFL 0.98·LW limit
    "lw_euler_vortex_1024x1024_N4": dict(level=10, N=4, td="lw"),
    "lw_euler_vortex_256x256_N4": dict(level=8, N=4, td="lw"),
    "lw_euler_vortex_32x32_N4": dict(level=5, N=4, td="lw"),
    "mdrk_euler_vortex_1024x1024_N3": dict(level=10, N=3, td="mdrk"),
    # 512² periodic vortex mesh whose central quarter [-5,5]^2 is refined once (458752 leaves,
    # 1024 mortars): the mortar path at scale
    "lw_euler_vortex_refined_512x512_box_N4": dict(level=9, N=4, td="lw", problem="vortex_refined"),
    # four-quadrant Riemann problem (Lax-Liu configuration 3) on [0,1]^2 with a time-averaged
    # Dirichlet boundary, Hennemann-Gassner indicator on density_pressure (alpha_max 0.5) and
    # first-order subcell FV blending; CFL 0.3 (the 0.98·LW limit loses positivity within 500 steps
    # without the reference's positivity limiter, on the CPU oracle as on the device)
    "lw_euler_riemann_sc_256x256_N4": dict(level=8, N=4, td="lw", problem="riemann_sc", cfl=0.3),
    "lw_euler_riemann_sc_1024x1024_N4": dict(level=10, N=4, td="lw", problem="riemann_sc", cfl=0.3),
    # the same with MUSCLHancockReconstruction (the reconstruction config 1 names; elements on the
    # domain boundary use first order, as in the reference), CFL 0.1: without a limiter the
    # reference's MUSCL-Hancock blend loses admissibility by step 200 at CFL 0.3 and by step 300 at
    # 0.2 (128² and 256², on the oracle as on the device); first order survives 500 steps at 0.3
    "lw_euler_riemann_mh_256x256_N4": dict(level=8, N=4, td="lw", problem="riemann_sc", cfl=0.1,
                                           recon="muscl_hancock"),
    "lw_euler_riemann_mh_1024x1024_N4": dict(level=10, N=4, td="lw", problem="riemann_sc", cfl=0.1,
                                             recon="muscl_hancock"),
    # compressible Navier-Stokes (conservative gradient variables, μ = 0.01, Pr = 0.72, the
    # parameters of the reference's elixir_navierstokes_convergence.jl) on a periodic [-1,1]^2
    # TreeMesh, smooth shear/density/pressure waves, CFL 0.5
    "lw_navierstokes_shear_1024x1024_N4": dict(level=10, N=4, td="lw", problem="ns_shear", cfl=0.5),
    "mdrk_navierstokes_shear_1024x1024_N3": dict(level=10, N=3, td="mdrk", problem="ns_shear", cfl=0.5),
    "lw_navierstokes_shear_256x256_N4": dict(level=8, N=4, td="lw", problem="ns_shear", cfl=0.5),
    # isentropic vortex on a periodic curvilinear StructuredMesh: [0,20]^2 warped by
    # x, y += 20·0.05 sin(2πs) sin(2πt) (s, t the unit-square coordinates), CFL 0.98·LW limit
    "lw_euler_vortex_warped_1024x1024_N4": dict(nx=1024, N=4, td="lw", problem="vortex_warped"),
    "lw_euler_vortex_warped_256x256_N4": dict(nx=256, N=4, td="lw", problem="vortex_warped"),
    # compressible Navier-Stokes on a periodic curvilinear mesh: a Taylor-Green vortex (U = 1,
    # ρ0 = 1, Mach 0.1) on [0,2]^2 warped by x, y += 2·0.1 sin(2πs) sin(2πt), mu 0.001 (Reynolds
    # number ρ0 U (L/2) / μ = 1000), Pr 0.72, CFL 0.5; the 128² MDRK N=3 case for 300 steps is the
    # curved Navier-Stokes configuration
    "mdrk_navierstokes_warped_128x128_N3": dict(nx=128, N=3, td="mdrk", problem="ns_warped", cfl=0.5),
    "lw_navierstokes_warped_1024x1024_N4": dict(nx=1024, N=4, td="lw", problem="ns_warped", cfl=0.5),
}
WARP_L, WARP_A = 20.0, 0.05
NS_MU, NS_PR = 0.01, 0.72
CNS_L, CNS_A, CNS_MU = 2.0, 0.1, 0.001
DEFAULT_WORKLOAD = "lw_euler_vortex_1024x1024_N4"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes_per_dof(N: int, td: str, parabolic: bool = False, curved: bool = False):
    """Algorithmic HBM bytes per DOF-update: (whole step, volume kernel).

    Whole step (LW): read uⁿ + write uⁿ⁺¹ + one write and one read of the face
    traces (uⁿ, Ũ, F̃·n: 3 × 4 variables × n nodes × 4 faces):
    8 (2·4 n² + 2·3·4·4 n) / n².  The volume kernel's share is the state read,
    the state write and the trace write: 8 (2·4 n² + 3·4·4 n) / n².
    MDRK: 8 (5·4 n² + 4·48 n) / n² per step (two stages).
    Navier-Stokes adds, per stage, the read of the neighbours' face states for
    the BR1 gradient (4 variables × n nodes × 4 faces): + 8·4·4 n / n².
    A curvilinear mesh adds the node metrics (Ja: 4, J⁻¹: 1 double per node)
    and the face normals (4 doubles per face node) to the volume kernel, and
    the node metrics again to the surface kernel (Jacobian and CFL speeds):
    volume + 8 (5 n² + 4·4 n) / n², step + 8 (2·5 n² + 4·4 n) / n².
    """
    n = N + 1
    br1 = 8.0 * 4 * 4 * n / (n * n) if parabolic else 0.0
    cstep = 8.0 * (2 * 5 * n * n + 4 * 4 * n) / (n * n) if curved else 0.0
    cvol = 8.0 * (5 * n * n + 4 * 4 * n) / (n * n) if curved else 0.0
    if td == "lw":
        step = 8.0 * (2 * 4 * n * n + 2 * 3 * 4 * 4 * n) / (n * n) + br1 + cstep
        vol = 8.0 * (2 * 4 * n * n + 3 * 4 * 4 * n) / (n * n) + br1 + cvol
    else:
        step = 8.0 * (5 * 4 * n * n + 4 * 48 * n) / (n * n) + (3 if parabolic else 0) * br1 + 2 * cstep
        vol = None
    return step, vol


def volume_kernel_bytes(mesh, N, tile, record_doubles=10, parabolic=False, curved=False):
    """Compulsory HBM bytes of one volume-kernel launch in this design: the
    read of uⁿ and the write of the whole updated state (4 doubles per node
    each), the face record (`record_doubles` doubles per face node) of every
    face slot whose neighbour lies in another volume workgroup of `tile`
    consecutive elements (or on the domain boundary), and the element's
    neighbour ids and inverse Jacobian; Navier-Stokes adds the neighbour's face
    state (4 doubles per face node) its BR1 gradient reads across those same
    leaving faces, a curvilinear mesh the node metrics (Ja and J⁻¹, 5 doubles
    per node) and face normals (4 per face node).  Faces whose two elements
    share a workgroup need no HBM traffic (the neighbour's state is in the
    workgroup's own tile) and are not counted."""
    n = N + 1
    E = mesh.n_elements
    nbr = np.asarray(mesh.neighbors, dtype=np.int64).reshape(E, 4)
    tiles = np.arange(E) // tile
    leaving = (nbr < 0) | (tiles[np.clip(nbr, 0, None)] != tiles[:, None])
    b = 8.0 * (2 * 4 * n * n * E + int(leaving.sum()) * n * record_doubles) + 24.0 * E
    if parabolic:
        b += 8.0 * 4 * n * int(leaving.sum())
    if curved:
        b += 8.0 * (5 * n * n + 4 * 4 * n) * E
    return b, float(leaving.mean())


def load_traffic(workload: str):
    """HBM bytes per launch of the dominant kernel, measured with rocprofv3
    FETCH_SIZE/WRITE_SIZE scaled by a 1 GiB copy (committed summary)."""
    p = ROOT / "profiles" / "traffic.json"
    if not p.exists():
        return None, None
    try:
        d = json.loads(p.read_text()).get(workload)
        return (float(d["bytes"]), d.get("source")) if d else (None, None)
    except Exception:
        return None, None


def load_valu(workload: str, launch_s: float):
    """FP64 VALU work of the dominant kernel per launch from the committed
    rocprofv3 SQ_INSTS_VALU_*_F64 counts (profiles/valu.json, per-wave
    instructions × 64 lanes; an FMA is 2 flops), against the 78.6 TFLOP/s FP64
    peak, with the live launch time."""
    p = ROOT / "profiles" / "valu.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text()).get(workload)
    except Exception:
        return None
    if not d:
        return None
    tf = d["flops_per_launch"] / launch_s / 1e12
    return {"achieved": round(tf, 2), "peak": 78.6, "unit": "TFLOP/s", "frac": round(tf / 78.6, 4),
            "flops_per_launch": d["flops_per_launch"], "source": d.get("source")}


def cpu_baseline(o, u0, cfl, mdrk, dofs, steps, threads):
    import oracle_ctypes

    oracle_ctypes.set_threads(threads)
    t0 = time.perf_counter()
    _, _, dts = o.solve(u0, 0.0, 1e30, cfl, max_steps=steps, mdrk=mdrk)
    el = time.perf_counter() - t0
    return dofs * len(dts) / el, el


def measure_batch(steps=200, n_instances=64, seed=2024, single_value=None):
    """64 isentropic-vortex instances (mesh 16²..128², polydeg 3 or 4, Mach U[0.2, 0.8],
    strength U[2, 5], drawn with numpy default_rng(seed)): a batch context per polydeg joined
    into one ensemble, so every LW step is one volume, one surface and one finalize launch for
    all 64 instances (own CFL dt per instance, 0.98·LW limit), replayed from a captured graph;
    the tile pair is chosen by timing both.  Returns the sub-record of the JSON line."""
    import torch
    import trixilw_hip as tlw

    levels, degs, machs, strengths = tlw.draw_vortex_ensemble(n_instances, seed)
    eq = tlw.CompressibleEulerEquations2D(1.4)
    stream = torch.cuda.Stream()
    batches, cfls, u0s, counts = [], [], [], {}
    for N in (3, 4):
        idx = np.nonzero(degs == N)[0]
        meshes = [tlw.TreeMesh((-10.0, -10.0), (10.0, 10.0), int(levels[k]), ordering="morton") for k in idx]
        ics = [tlw.IsentropicVortex(strength=float(strengths[k]), mach=float(machs[k])) for k in idx]
        solver = tlw.DGSEM(int(N))
        b = tlw.SemidiscretizationBatch(meshes, tlw.LW(), eq, ics, solver)
        batches.append(b)
        cfls.append(tlw.trixi2lw(0.98, solver))
        u0s.append(b.compute_coefficients())
        counts[f"N{N}"] = {"instances": int(len(idx)), "dofs": int(b.ndofs())}
    ens = tlw.SemidiscretizationEnsemble(batches)
    total_dofs = ens.ndofs()
    tune = {}
    for tiles in ens.TILES:  # capture + warm-up, then a short timed run per tile pair
        ens.set_tiles(*tiles)
        ens.init(u0s, cfls, stream=stream, graphs=True)
        ens.advance(20)
        ens.init(u0s, cfls, stream=stream, graphs=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ens.advance(20)
        torch.cuda.synchronize()
        tune[str(tiles)] = round(1e3 * (time.perf_counter() - t0) / 20, 4)
    best = min(ens.TILES, key=lambda tl: tune[str(tl)])
    ens.set_tiles(*best)
    ens.init(u0s, cfls, stream=stream, graphs=True)
    ens.advance(steps)  # capture of the timed step count
    ens.init(u0s, cfls, stream=stream, graphs=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.advance(steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    for b in batches:
        st = b.status()
        assert np.all(st["iter"] == steps), st["iter"]
    value = total_dofs * steps / el
    out = {"value": value, "unit": "DOF-updates/s", "instances": n_instances, "steps": steps,
           "total_dofs": int(total_dofs), "ms_per_step": round(1e3 * el / steps, 4), "groups": counts,
           "tiles": list(best), "tile_ms_per_step": tune,
           "config": "64 isentropic-vortex instances, 16^2..128^2 elements, polydeg 3/4, Mach U[0.2,0.8], "
                     "strength U[2,5], LW, one volume + one surface + one finalize launch per step for all "
                     "instances (tlw_create_ensemble), graph-replayed steps"}
    ens.close()
    for b in batches:
        b.close()
    if single_value:
        # against one large instance of each polydeg (1024² vortex, LW) on the same DOF mix
        singles = {4: single_value, 3: measure_single(3)}
        mix_time = sum(counts[f"N{N}"]["dofs"] / singles[N] for N in (3, 4))
        out["single_instance_values"] = {f"N{N}": singles[N] for N in (3, 4)}
        out["vs_single_instance"] = round(value / (total_dofs / mix_time), 3)
        out["vs_single_note"] = ("the ensemble's DOF-updates/s over the single-instance rate of the same DOF mix "
                                 "(each polydeg's DOFs at its 1024² single-instance rate)")
    return out


def measure_single(N, steps=20, warmup=3):
    """DOF-updates/s of one 1024² vortex instance, LW, polydeg N (autotuned tile)."""
    wl = {"level": 10, "N": N, "td": "lw"}
    semi, mesh, ic, cfl, _, _ = make_problem(wl)
    x, y = semi.node_coordinates()
    r = run_steps(semi, np.ascontiguousarray(ic(x, y, 0.0)), cfl, steps, warmup)
    semi.close()
    return r["value"]


def ns_shear(x, y, t=0.0, gamma=1.4):
    """Smooth periodic state on [-1, 1]²: shear flow with density and pressure waves."""
    rho = 1.0 + 0.2 * np.sin(np.pi * x) * np.cos(np.pi * y)
    v1 = 0.3 + 0.1 * np.sin(np.pi * y)
    v2 = -0.2 + 0.1 * np.cos(np.pi * x)
    p = 1.0 + 0.1 * np.cos(np.pi * (x + y))
    return np.stack([rho, rho * v1, rho * v2, p / (gamma - 1) + 0.5 * rho * (v1**2 + v2**2)], axis=-1)


def warp_mapping(L, A):
    """[0, L]² warped by A sin 2πs sin 2πt in both coordinates (periodic)."""

    def mapping(xi, eta):
        s, t = 0.5 * (xi + 1.0), 0.5 * (eta + 1.0)
        w = A * np.sin(2.0 * np.pi * s) * np.sin(2.0 * np.pi * t)
        return L * (s + w), L * (t + w)

    return mapping


def make_problem(wl):
    """Semidiscretization, initial condition, CFL number and the oracle's
    constructor arguments of a workload."""
    import trixilw_hip as tlw

    N, td = wl["N"], wl["td"]
    problem = wl.get("problem")
    eq = tlw.CompressibleEulerEquations2D(1.4)
    vi = tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW())
    okw = {}
    if problem == "riemann_sc":
        ic = tlw.FourQuadrantRiemann()
        lo, hi, bc, periodic = 0.0, 1.0, ic, False
        recon = wl.get("recon", "first_order")
        rec = tlw.MUSCLHancockReconstruction() if recon == "muscl_hancock" else tlw.FirstOrderReconstruction()
        vi = tlw.VolumeIntegralFRShockCapturing(tlw.IndicatorHennemannGassner(eq, None, alpha_max=0.5), vi,
                                                reconstruction=rec)
        okw = dict(bc_kind=bc.kind, bc_par=bc.params, shock_capturing=True, alpha_max=0.5, reconstruction=recon)
    elif problem == "ns_shear":
        ic = ns_shear
        lo, hi, bc, periodic = -1.0, 1.0, None, True
        okw = dict(parabolic="navier_stokes", mu=NS_MU, prandtl=NS_PR)
    elif problem == "ns_warped":
        import oracle_ctypes

        ic = tlw.TaylorGreenVortex(U=1.0, rho0=1.0, mach=0.1, L=CNS_L)

        lo, hi, bc, periodic = 0.0, CNS_L, None, True
        okw = dict(parabolic="navier_stokes", mu=CNS_MU, prandtl=NS_PR,
                   mapping=(oracle_ctypes.MAP_WARP, (CNS_L, CNS_A)))
    elif problem == "vortex_warped":
        import oracle_ctypes

        ic = tlw.IsentropicVortex(center=(WARP_L / 2, WARP_L / 2), period=(WARP_L, WARP_L))
        lo, hi, bc, periodic = 0.0, WARP_L, None, True
        okw = dict(mapping=(oracle_ctypes.MAP_WARP, (WARP_L, WARP_A)))
    else:
        ic = tlw.IsentropicVortex()
        lo, hi, bc, periodic = -10.0, 10.0, None, True
    # two-row strips: the volume kernel's 10-element tiles are 5×2 blocks, 35 % of whose face
    # slots leave the tile (Morton runs of 10: 42.5 %, rows of 10: 55 %), so fewer face records
    # cross HBM
    if problem == "vortex_warped":
        mesh = tlw.StructuredMesh((wl["nx"], wl["nx"]), warp_mapping(WARP_L, WARP_A), ordering="strips")
    elif problem == "ns_warped":
        mesh = tlw.StructuredMesh((wl["nx"], wl["nx"]), warp_mapping(CNS_L, CNS_A), ordering="strips")
    elif problem == "vortex_refined":
        box = {"type": "box", "coordinates_min": (-5.0, -5.0), "coordinates_max": (5.0, 5.0)}
        mesh = tlw.TreeMesh((lo, lo), (hi, hi), wl["level"], periodicity=periodic, refinement_patches=(box,))
        okw["tree"] = (mesh.element_center[:, 0], mesh.element_center[:, 1], mesh.element_length, mesh.neighbors,
                       mesh.mortars)
    else:
        mesh = tlw.TreeMesh((lo, lo), (hi, hi), wl["level"], periodicity=periodic, ordering="strips")
    solver = tlw.DGSEM(N, volume_integral=vi)
    if problem in ("ns_shear", "ns_warped"):
        eqp = tlw.CompressibleNavierStokesDiffusion2D(eq, mu=NS_MU if problem == "ns_shear" else CNS_MU, Prandtl=NS_PR,
                                                     gradient_variables=tlw.GradientVariablesConservative())
        semi = tlw.SemidiscretizationHyperbolicParabolic(mesh, tlw.get_time_discretization(solver), (eq, eqp), ic,
                                                         solver)
    else:
        semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.get_time_discretization(solver), eq, ic, solver,
                                                boundary_conditions=bc)
    cfl = wl.get("cfl") or tlw.trixi2lw(0.98, solver) * (0.5 if td == "mdrk" else 1.0)
    okw.update(periodic=(periodic, periodic))
    return semi, mesh, ic, cfl, (lo, hi), okw


def run_steps(semi, u0, cfl, steps, warmup, tile=0, repeats=0):
    """Device loop: optional tile autotune, `warmup` untimed steps, then `steps`
    timed steps with HIP-event kernel timing.  With `repeats`, the same `steps`
    from the same initial state are timed that many more times (wall clock
    only), for the median.  Returns a dict of the results."""
    import torch
    from trixilw_hip import _lib

    L = _lib.lib()
    h = semi.handle
    dp = C.POINTER(C.c_double)
    semi._bind_stream()
    _lib.check(L.tlw_set_state(h, u0.ctypes.data_as(dp), u0.size), h)
    if tile:
        semi.elements_per_workgroup(tile)
        tune = None
    else:
        chosen, tune = semi.autotune(u0, reps=3)
        log("bench: tile autotune (ms per rhs): " + ", ".join(f"{k} el/wg {v:.3f}" for k, v in tune.items())
            + f" -> {chosen}")
    tile = semi.elements_per_workgroup()
    _lib.check(L.tlw_set_state(h, u0.ctypes.data_as(dp), u0.size), h)
    _lib.check(L.tlw_solve_init(h, 0.0, float("inf"), cfl), h)
    t, dt = C.c_double(), C.c_double()
    it, fin = C.c_int64(), C.c_int32()
    if warmup > 0:
        _lib.check(L.tlw_advance(h, warmup, C.byref(t), C.byref(dt), C.byref(it), C.byref(fin)), h)
    _lib.check(L.tlw_enable_kernel_timing(h, 1), h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.check(L.tlw_advance(h, steps, C.byref(t), C.byref(dt), C.byref(it), C.byref(fin)), h)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    vms, vn, sms, sn = C.c_double(), C.c_int64(), C.c_double(), C.c_int64()
    _lib.check(L.tlw_kernel_times(h, C.byref(vms), C.byref(vn), C.byref(sms), C.byref(sn)), h)
    _lib.check(L.tlw_enable_kernel_timing(h, 0), h)
    steps_done = int(it.value) - warmup
    assert steps_done == steps, (steps_done, steps)
    rep_ms = []
    for _ in range(repeats):
        _lib.check(L.tlw_set_state(h, u0.ctypes.data_as(dp), u0.size), h)
        _lib.check(L.tlw_solve_init(h, 0.0, float("inf"), cfl), h)
        if warmup > 0:
            _lib.check(L.tlw_advance(h, warmup, C.byref(t), C.byref(dt), C.byref(it), C.byref(fin)), h)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _lib.check(L.tlw_advance(h, steps, C.byref(t), C.byref(dt), C.byref(it), C.byref(fin)), h)
        torch.cuda.synchronize()
        rep_ms.append(1e3 * (time.perf_counter() - t1) / steps)
        assert int(it.value) == warmup + steps, (int(it.value), warmup, steps)
    dofs = semi.ndofs()
    return {"value": dofs * steps / el, "ms_per_step": 1e3 * el / steps, "elapsed": el,
            "vol_ms": vms.value / max(vn.value, 1), "vol_n": int(vn.value),
            "surf_ms": sms.value / max(sn.value, 1), "surf_n": int(sn.value), "tile": tile, "tune": tune,
            "repeat_ms": rep_ms}


def measure_navier_stokes(steps, cpu_steps, threads, with_cpu=True):
    """Sub-record: LW N=4 Navier-Stokes on the 1024² periodic mesh, device
    loop throughput, volume kernel bytes per launch and the CPU oracle."""
    wl = WORKLOADS["lw_navierstokes_shear_1024x1024_N4"]
    semi, mesh, ic, cfl, (lo, hi), okw = make_problem(wl)
    x, y = semi.node_coordinates()
    u0 = np.ascontiguousarray(ic(x, y, 0.0))
    r = run_steps(semi, u0, cfl, steps, 2)
    dofs = semi.ndofs()
    vol_bytes, _ = volume_kernel_bytes(mesh, 4, r["tile"], 10, parabolic=True)
    out = {"workload": "lw_navierstokes_shear_1024x1024_N4", "value": r["value"], "unit": "DOF-updates/s",
           "ms_per_step": round(r["ms_per_step"], 4), "steps": steps, "dofs": dofs,
           "volume_avg_launch_ms": round(r["vol_ms"], 4), "surface_avg_launch_ms": round(r["surf_ms"], 4),
           "volume_achieved_GBs": round(vol_bytes / (r["vol_ms"] * 1e-3) / 1e9, 2),
           "volume_alg_bytes_per_dof": round(vol_bytes / dofs, 2),
           "volume_valu": load_valu("lw_navierstokes_shear_1024x1024_N4", r["vol_ms"] * 1e-3),
           "volume_traffic": load_traffic("lw_navierstokes_shear_1024x1024_N4")[0],
           "elements_per_workgroup": r["tile"], "tile_autotune_ms": r["tune"],
           "config": f"compressible Navier-Stokes, conservative gradients, mu {NS_MU}, Pr {NS_PR}, periodic "
                     f"[-1,1]^2 {mesh.nx}x{mesh.nx} TreeMesh, polydeg 4, LW, CFL {cfl}"}
    log(f"bench: Navier-Stokes {r['value'] / 1e9:.3f} G DOF-updates/s, {r['ms_per_step']:.3f} ms/step, "
        f"volume {r['vol_ms']:.3f} ms, surface {r['surf_ms']:.3f} ms")
    semi.close()
    if with_cpu:
        import oracle_ctypes

        nx = mesh.nx
        o = oracle_ctypes.Oracle(4, nx, nx, lo, lo, (hi - lo) / nx, **okw)
        xo, yo = o.node_coords()
        v, el = cpu_baseline(o, np.ascontiguousarray(ic(xo, yo, 0.0)), cfl, False, dofs, cpu_steps, threads)
        log(f"bench: Navier-Stokes CPU {v / 1e6:.2f} M DOF-updates/s ({el:.1f} s)")
        out["cpu_baseline"] = {"value": v, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
                               "sample": f"{cpu_steps} step(s) of the same mesh, C++/OpenMP oracle"}
        out["gpu_over_cpu"] = round(r["value"] / v, 2)
        del o
    return out


def measure_shock_capturing(steps, cpu_steps, threads, with_cpu=True):
    """Sub-record: config 1 as the reference names it, the four-quadrant Riemann
    problem at 256² with HG blending and MUSCL-Hancock subcell FV: device loop
    throughput, kernel times and the CPU oracle."""
    name = "lw_euler_riemann_mh_256x256_N4"
    wl = WORKLOADS[name]
    semi, mesh, ic, cfl, (lo, hi), okw = make_problem(wl)
    x, y = semi.node_coordinates()
    u0 = np.ascontiguousarray(ic(x, y, 0.0))
    r = run_steps(semi, u0, cfl, steps, 2)
    dofs = semi.ndofs()
    out = {"workload": name, "value": r["value"], "unit": "DOF-updates/s", "ms_per_step": round(r["ms_per_step"], 4),
           "steps": steps, "dofs": dofs, "volume_avg_launch_ms": round(r["vol_ms"], 4),
           "surface_avg_launch_ms": round(r["surf_ms"], 4), "elements_per_workgroup": r["tile"],
           "tile_autotune_ms": r["tune"],
           "config": f"four-quadrant Riemann problem (Lax-Liu 3) on [0,1]^2, {mesh.nx}x{mesh.nx} TreeMesh, "
                     f"time-averaged Dirichlet boundary, polydeg 4, LW, HG indicator on density_pressure "
                     f"(alpha_max 0.5), MUSCL-Hancock subcell FV, CFL {cfl}"}
    log(f"bench: shock capturing (MUSCL-Hancock) {r['value'] / 1e9:.3f} G DOF-updates/s, {r['ms_per_step']:.3f} "
        f"ms/step, volume {r['vol_ms']:.3f} ms, surface {r['surf_ms']:.3f} ms")
    semi.close()
    if with_cpu:
        import oracle_ctypes

        nx = mesh.nx
        o = oracle_ctypes.Oracle(4, nx, nx, lo, lo, (hi - lo) / nx, **okw)
        xo, yo = o.node_coords()
        v, el = cpu_baseline(o, np.ascontiguousarray(ic(xo, yo, 0.0)), cfl, False, dofs, cpu_steps, threads)
        log(f"bench: shock capturing CPU {v / 1e6:.2f} M DOF-updates/s ({el:.1f} s)")
        out["cpu_baseline"] = {"value": v, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
                               "sample": f"{cpu_steps} steps of the same mesh, C++/OpenMP oracle"}
        out["gpu_over_cpu"] = round(r["value"] / v, 2)
        del o
    return out


def measure_config0(threads, with_cpu=True):
    """Sub-record: the reference's CPU-runnable case itself, the isentropic vortex
    on the periodic 32² TreeMesh at N=4, CFL 0.98·LW limit, solved to t = 1 by the
    device loop (launched, then replayed from a captured HIP graph), and by the
    CPU oracle on the host's cores."""
    import torch
    from trixilw_hip import _lib

    name = "lw_euler_vortex_32x32_N4"
    semi, mesh, ic, cfl, (lo, hi), okw = make_problem(WORKLOADS[name])
    x, y = semi.node_coordinates()
    u0 = np.ascontiguousarray(ic(x, y, 0.0))
    dofs = semi.ndofs()
    L, h, dp = _lib.lib(), semi.handle, C.POINTER(C.c_double)
    t, dt, it, fin = C.c_double(), C.c_double(), C.c_int64(), C.c_int32()
    runs = {}
    with torch.cuda.stream(torch.cuda.Stream()):
        semi._bind_stream()
        for mode, passes in (("launched", (False, True)), ("graph_replay", (False, True))):
            _lib.check(L.tlw_enable_graphs(h, int(mode == "graph_replay")), h)
            for timed in passes:  # the untimed pass warms up (and captures the graph)
                _lib.check(L.tlw_set_state(h, u0.ctypes.data_as(dp), u0.size), h)
                _lib.check(L.tlw_solve_init(h, 0.0, 1.0, cfl), h)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                while True:
                    _lib.check(L.tlw_advance(h, 200, C.byref(t), C.byref(dt), C.byref(it), C.byref(fin)), h)
                    if fin.value:
                        break
                torch.cuda.synchronize()
                el = time.perf_counter() - t0
            runs[mode] = (el, int(it.value), t.value)
        _lib.check(L.tlw_enable_graphs(h, 0), h)
    el, nsteps, tend = runs["launched"]
    u_dev = np.empty_like(u0)
    _lib.check(L.tlw_get_state(h, u_dev.ctypes.data_as(dp), u_dev.size), h)
    out = {"workload": name, "value": dofs * nsteps / el, "unit": "DOF-updates/s", "steps": nsteps, "t_final": tend,
           "dofs": dofs, "ms_to_t1": round(1e3 * el, 3), "ms_per_step": round(1e3 * el / nsteps, 4),
           "graph_replay": {"value": dofs * nsteps / runs["graph_replay"][0],
                            "ms_to_t1": round(1e3 * runs["graph_replay"][0], 3),
                            "ms_per_step": round(1e3 * runs["graph_replay"][0] / nsteps, 4)},
           "config": f"isentropic vortex strength 5, Mach 0.5, periodic [-10,10]^2 TreeMesh 32x32, polydeg 4, LW, "
                     f"CFL {cfl}, t = 0 .. 1"}
    log(f"bench: config 0 (32^2 to t = 1): {nsteps} steps, {1e3 * el:.2f} ms launched, "
        f"{1e3 * runs['graph_replay'][0]:.2f} ms graph-replayed")
    semi.close()
    if with_cpu:
        import oracle_ctypes

        oracle_ctypes.set_threads(threads)
        o = oracle_ctypes.Oracle(4, mesh.nx, mesh.nx, lo, lo, (hi - lo) / mesh.nx, **okw)
        xo, yo = o.node_coords()
        t0 = time.perf_counter()
        u_cpu, t_cpu, dts = o.solve(np.ascontiguousarray(ic(xo, yo, 0.0)), 0.0, 1.0, cfl)
        el_cpu = time.perf_counter() - t0
        u_lex = u_dev[mesh.lexicographic_permutation()]
        out["cpu_baseline"] = {"value": dofs * len(dts) / el_cpu, "unit": "DOF-updates/s", "cores": threads,
                               "kind": "port", "sample": f"the whole run to t = 1 ({len(dts)} steps), C++/OpenMP oracle",
                               "seconds": round(el_cpu, 4), "steps": len(dts),
                               "max_rel_diff_of_final_state": float(np.max(np.abs(u_lex - u_cpu)) /
                                                                    np.max(np.abs(u_cpu)))}
        out["gpu_over_cpu"] = round(out["value"] / out["cpu_baseline"]["value"], 2)
        log(f"bench: config 0 CPU {el_cpu:.3f} s for {len(dts)} steps")
        del o
    return out


def measure_mdrk_euler(steps, cpu_steps, threads, with_cpu=True):
    """Sub-record: MDRK N=3 Euler on the periodic 1024² vortex mesh (two
    volume/surface pairs per step): device-loop throughput, kernel times, the
    step's bytes against HBM and the CPU oracle."""
    name = "mdrk_euler_vortex_1024x1024_N3"
    wl = WORKLOADS[name]
    semi, mesh, ic, cfl, (lo, hi), okw = make_problem(wl)
    x, y = semi.node_coordinates()
    u0 = np.ascontiguousarray(ic(x, y, 0.0))
    r = run_steps(semi, u0, cfl, steps, 2)
    dofs = semi.ndofs()
    step_b, _ = alg_bytes_per_dof(3, "mdrk")
    out = {"workload": name, "value": r["value"], "unit": "DOF-updates/s", "ms_per_step": round(r["ms_per_step"], 4),
           "steps": steps, "dofs": dofs, "volume_avg_launch_ms": round(r["vol_ms"], 4), "volume_launches": r["vol_n"],
           "surface_avg_launch_ms": round(r["surf_ms"], 4), "step_alg_bytes_per_dof": step_b,
           "step_frac": round(step_b * r["value"] / 1e9 / HBM_PEAK_GBS, 4),
           "elements_per_workgroup": r["tile"], "tile_autotune_ms": r["tune"],
           "config": f"isentropic vortex strength 5, Mach 0.5, periodic [-10,10]^2 TreeMesh {mesh.nx}x{mesh.nx}, "
                     f"polydeg 3, MDRK, CFL {cfl}"}
    log(f"bench: MDRK N=3 Euler {r['value'] / 1e9:.3f} G DOF-updates/s, {r['ms_per_step']:.3f} ms/step")
    semi.close()
    if with_cpu:
        import oracle_ctypes

        o = oracle_ctypes.Oracle(3, mesh.nx, mesh.nx, lo, lo, (hi - lo) / mesh.nx, **okw)
        xo, yo = o.node_coords()
        v, el = cpu_baseline(o, np.ascontiguousarray(ic(xo, yo, 0.0)), cfl, True, dofs, cpu_steps, threads)
        out["cpu_baseline"] = {"value": v, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
                               "sample": f"{cpu_steps} step(s) of the same mesh, C++/OpenMP oracle"}
        out["gpu_over_cpu"] = round(r["value"] / v, 2)
        del o
    return out


def measure_refined(steps, cpu_steps, threads, with_cpu=True, uniform_value=None):
    """Sub-record: LW N=4 Euler on a TreeMesh with refined cells (the 512²
    vortex mesh with its central quarter refined once; L2 mortars at the
    refinement edge): device-loop throughput, kernel times (the surface time
    includes the mortar kernel) and the CPU oracle on the same leaves."""
    name = "lw_euler_vortex_refined_512x512_box_N4"
    wl = WORKLOADS[name]
    semi, mesh, ic, cfl, (lo, hi), okw = make_problem(wl)
    x, y = semi.node_coordinates()
    u0 = np.ascontiguousarray(ic(x, y, 0.0))
    r = run_steps(semi, u0, cfl, steps, 2)
    dofs = semi.ndofs()
    out = {"workload": name, "value": r["value"], "unit": "DOF-updates/s", "ms_per_step": round(r["ms_per_step"], 4),
           "steps": steps, "dofs": dofs, "elements": int(mesh.n_elements), "mortars": int(len(mesh.mortars)),
           "volume_avg_launch_ms": round(r["vol_ms"], 4), "surface_and_mortar_avg_launch_ms": round(r["surf_ms"], 4),
           "elements_per_workgroup": r["tile"], "tile_autotune_ms": r["tune"],
           "config": "isentropic vortex strength 5, Mach 0.5, periodic [-10,10]^2 TreeMesh 512x512 with the "
                     f"leaves of [-5,5]^2 refined once (Z-order), polydeg 4, LW, CFL {cfl}"}
    if uniform_value:
        out["vs_uniform_1024x1024"] = round(r["value"] / uniform_value, 4)
    log(f"bench: refined TreeMesh {r['value'] / 1e9:.3f} G DOF-updates/s, {r['ms_per_step']:.3f} ms/step, "
        f"{len(mesh.mortars)} mortars")
    semi.close()
    if with_cpu:
        import oracle_ctypes

        o = oracle_ctypes.Oracle(4, 1, 1, 0.0, 0.0, 1.0, **okw)
        xo, yo = o.node_coords()
        v, el = cpu_baseline(o, np.ascontiguousarray(ic(xo, yo, 0.0)), cfl, False, dofs, cpu_steps, threads)
        out["cpu_baseline"] = {"value": v, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
                               "sample": f"{cpu_steps} step(s) of the same leaves, C++/OpenMP oracle"}
        out["gpu_over_cpu"] = round(r["value"] / v, 2)
        del o
    return out


def measure_curved_navier_stokes(steps, cpu_steps, threads, with_cpu=True):
    """Sub-record: MDRK N=3 compressible Navier-Stokes on the warped 128²
    StructuredMesh (warp amplitude 0.1, Reynolds number 1000, Pr 0.72) for
    `steps` steps of the device loop, and the CPU oracle."""
    name = "mdrk_navierstokes_warped_128x128_N3"
    wl = WORKLOADS[name]
    semi, mesh, ic, cfl, (lo, hi), okw = make_problem(wl)
    x, y = semi.node_coordinates()
    u0 = np.ascontiguousarray(ic(x, y, 0.0))
    r = run_steps(semi, u0, cfl, steps, 2)
    dofs = semi.ndofs()
    step_b, _ = alg_bytes_per_dof(3, "mdrk", parabolic=True, curved=True)
    out = {"workload": name, "value": r["value"], "unit": "DOF-updates/s",
           "ms_per_step": round(r["ms_per_step"], 4), "steps": steps, "dofs": dofs,
           "volume_avg_launch_ms": round(r["vol_ms"], 4), "surface_avg_launch_ms": round(r["surf_ms"], 4),
           "step_alg_bytes_per_dof": step_b, "elements_per_workgroup": r["tile"], "tile_autotune_ms": r["tune"],
           "config": f"compressible Navier-Stokes, conservative gradients, mu {CNS_MU} (Re 1000), Pr {NS_PR}, "
                     f"periodic StructuredMesh [0,{CNS_L:g}]^2 warped by amplitude {CNS_A}, {mesh.nx}x{mesh.ny}, "
                     f"polydeg 3, MDRK, CFL {cfl}"}
    log(f"bench: curvilinear Navier-Stokes MDRK {r['value'] / 1e9:.3f} G DOF-updates/s, {r['ms_per_step']:.3f} "
        f"ms/step, volume {r['vol_ms']:.3f} ms x{r['vol_n']}, surface {r['surf_ms']:.3f} ms")
    # the same steps replayed from a captured HIP graph (tlw_enable_graphs): at this size the
    # launch gaps are a visible share of the step
    import torch
    from trixilw_hip import _lib

    L, h, dp = _lib.lib(), semi.handle, C.POINTER(C.c_double)
    t, dt, it, fin = C.c_double(), C.c_double(), C.c_int64(), C.c_int32()
    with torch.cuda.stream(torch.cuda.Stream()):
        semi._bind_stream()
        _lib.check(L.tlw_enable_graphs(h, 1), h)
        for timed in (False, True):  # the first pass captures the graph
            _lib.check(L.tlw_set_state(h, u0.ctypes.data_as(dp), u0.size), h)
            _lib.check(L.tlw_solve_init(h, 0.0, float("inf"), cfl), h)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _lib.check(L.tlw_advance(h, steps, C.byref(t), C.byref(dt), C.byref(it), C.byref(fin)), h)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
        assert int(it.value) == steps, (int(it.value), steps)
        _lib.check(L.tlw_enable_graphs(h, 0), h)
    out["graph_replay"] = {"value": dofs * steps / el, "ms_per_step": round(1e3 * el / steps, 4),
                           "note": "the same steps from the same initial state, replayed from one captured graph"}
    log(f"bench: curvilinear Navier-Stokes MDRK, graph replay {dofs * steps / el / 1e9:.3f} G DOF-updates/s, "
        f"{1e3 * el / steps:.3f} ms/step")
    semi.close()
    if with_cpu:
        import oracle_ctypes

        nx = mesh.nx
        o = oracle_ctypes.Oracle(3, nx, nx, lo, lo, (hi - lo) / nx, **okw)
        xo, yo = o.node_coords()
        v, el = cpu_baseline(o, np.ascontiguousarray(ic(xo, yo, 0.0)), cfl, True, dofs, cpu_steps, threads)
        log(f"bench: curvilinear Navier-Stokes CPU {v / 1e6:.2f} M DOF-updates/s ({el:.1f} s)")
        out["cpu_baseline"] = {"value": v, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
                               "sample": f"{cpu_steps} steps of the same mesh, C++/OpenMP oracle"}
        out["gpu_over_cpu"] = round(r["value"] / v, 2)
        del o
    return out


def measure_curved(steps, cpu_steps, threads, with_cpu=True):
    """Sub-record: LW N=4 isentropic vortex on the warped 1024² StructuredMesh,
    device loop throughput, volume kernel bytes per launch and the CPU oracle."""
    name = "lw_euler_vortex_warped_1024x1024_N4"
    wl = WORKLOADS[name]
    semi, mesh, ic, cfl, (lo, hi), okw = make_problem(wl)
    x, y = semi.node_coordinates()
    u0 = np.ascontiguousarray(ic(x, y, 0.0))
    r = run_steps(semi, u0, cfl, steps, 2)
    dofs = semi.ndofs()
    vol_bytes, _ = volume_kernel_bytes(mesh, 4, r["tile"], 10, curved=True)
    out = {"workload": name, "value": r["value"], "unit": "DOF-updates/s",
           "ms_per_step": round(r["ms_per_step"], 4), "steps": steps, "dofs": dofs,
           "volume_avg_launch_ms": round(r["vol_ms"], 4), "surface_avg_launch_ms": round(r["surf_ms"], 4),
           "volume_achieved_GBs": round(vol_bytes / (r["vol_ms"] * 1e-3) / 1e9, 2),
           "volume_alg_bytes_per_dof": round(vol_bytes / dofs, 2),
           "elements_per_workgroup": r["tile"], "tile_autotune_ms": r["tune"],
           "config": f"isentropic vortex, periodic StructuredMesh [0,{WARP_L:g}]^2 warped by amplitude {WARP_A}, "
                     f"{mesh.nx}x{mesh.ny}, polydeg 4, LW, CFL {cfl:.4f}"}
    log(f"bench: curvilinear {r['value'] / 1e9:.3f} G DOF-updates/s, {r['ms_per_step']:.3f} ms/step, "
        f"volume {r['vol_ms']:.3f} ms, surface {r['surf_ms']:.3f} ms")
    semi.close()
    if with_cpu:
        import oracle_ctypes

        nx = mesh.nx
        o = oracle_ctypes.Oracle(4, nx, nx, lo, lo, (hi - lo) / nx, **okw)
        xo, yo = o.node_coords()
        v, el = cpu_baseline(o, np.ascontiguousarray(ic(xo, yo, 0.0)), cfl, False, dofs, cpu_steps, threads)
        log(f"bench: curvilinear CPU {v / 1e6:.2f} M DOF-updates/s ({el:.1f} s)")
        out["cpu_baseline"] = {"value": v, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
                               "sample": f"{cpu_steps} step(s) of the same mesh, C++/OpenMP oracle"}
        out["gpu_over_cpu"] = round(r["value"] / v, 2)
        del o
    return out


def measure_gradient(nsteps=200, checkpoint_every=20, level=6, seed=2025, reps=3):
    """The differentiation configuration: d KE(u_200) / d u0 on the 64² vortex,
    N=4, 200 LW steps at 0.98 of the LW CFL limit, checkpoints every 20, the
    initial density perturbed by a seeded Gaussian random field of amplitude
    1e-2.  Times the forward device loop alone (200 steps, checkpoint copies
    included) and the whole gradient (forward, segment re-forward and adjoint
    steps); median of `reps` runs after one warm-up."""
    import torch
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(1.4)
    ic = tlw.IsentropicVortex()
    mesh = tlw.TreeMesh((-10.0, -10.0), (10.0, 10.0), level, ordering="strips")
    solver = tlw.DGSEM(4, volume_integral=tlw.VolumeIntegralFR(tlw.LW()))
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, ic, solver)
    x, y = semi.node_coordinates()
    u_host = ic(x, y, 0.0)
    u_host[..., 0] += tlw.gaussian_random_field(x, y, (20.0, 20.0), seed=seed, amplitude=1e-2)
    u0 = torch.from_numpy(np.ascontiguousarray(u_host)).to("cuda")
    cfl = tlw.trixi2lw(0.98, solver)
    from trixilw_hip.adjoint import _forward_device

    fwd, grad = [], []
    for r in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _forward_device(semi, u0, nsteps, cfl, 0.0, None, checkpoint_every)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        tlw.gradient_lw(semi, u0, nsteps, cfl, checkpoint_every=checkpoint_every)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if r > 0:
            fwd.append(t1 - t0)
            grad.append(t2 - t1)
    f, g = float(np.median(fwd)), float(np.median(grad))
    semi.close()
    return {"config": f"{mesh.nx}x{mesh.nx} vortex N=4, {nsteps} LW steps, checkpoint every {checkpoint_every}, "
                      f"GRF density perturbation 1e-2 (seed {seed})",
            "forward_ms_per_step": round(1e3 * f / nsteps, 4), "gradient_ms_per_step": round(1e3 * g / nsteps, 4),
            "gradient_over_forward": round(g / f, 3),
            "reverse_over_forward": round((g - f) / f, 3),
            "note": "wall clock of gradient_lw (device-loop forward with checkpoint copies, re-forward of each "
                    "segment by rhs! + update, adjoint steps) over the forward device loop alone; "
                    f"median of {reps}"}


def measure_adaptive(semi, u0, attempts, tol=(1e-10, 1e-10), dt0=1e-5):
    """Sub-record: the adaptive (error-controlled) LW loop on the same mesh and
    state: `attempts` device-loop attempts after 30 untimed ones (the controller
    grows dt by at most 1.5 per step from dt0)."""
    import torch
    from trixilw_hip import _lib

    L = _lib.lib()
    h = semi.handle
    dp = C.POINTER(C.c_double)
    _lib.check(L.tlw_set_state(h, u0.ctypes.data_as(dp), u0.size), h)
    _lib.check(L.tlw_solve_init_adaptive(h, 0.0, float("inf"), dt0, tol[0], tol[1]), h)
    t, dt = C.c_double(), C.c_double()
    it, fin = C.c_int64(), C.c_int32()
    _lib.check(L.tlw_advance(h, 30, C.byref(t), C.byref(dt), C.byref(it), C.byref(fin)), h)
    it0, t0s = int(it.value), t.value
    nrej0 = C.c_int64()
    _lib.check(L.tlw_rejected_steps(h, C.byref(nrej0)), h)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    _lib.check(L.tlw_advance(h, attempts, C.byref(t), C.byref(dt), C.byref(it), C.byref(fin)), h)
    torch.cuda.synchronize()
    el = time.perf_counter() - w0
    nrej = C.c_int64()
    _lib.check(L.tlw_rejected_steps(h, C.byref(nrej)), h)
    acc = int(it.value) - it0
    dofs = semi.ndofs()
    out = {"value": dofs * acc / el, "unit": "DOF-updates/s (accepted steps)", "attempts": attempts,
           "accepted": acc, "rejected": int(nrej.value - nrej0.value), "ms_per_attempt": round(1e3 * el / attempts, 4),
           "simulated_time": t.value - t0s, "last_dt": dt.value,
           "config": f"Adaptive(), abstol {tol[0]}, reltol {tol[1]}, dt_initial {dt0}, same mesh and state"}
    log(f"bench: adaptive {out['value'] / 1e9:.3f} G DOF-updates/s, {out['ms_per_attempt']:.3f} ms/attempt, "
        f"{acc} accepted / {out['rejected']} rejected")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--repeats", type=int, default=5, help="extra timed repeats of the main run (median)")
    ap.add_argument("--tile", type=int, default=0, help="elements per volume workgroup (0: autotune)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward step ratio")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-instance batch measurement")
    ap.add_argument("--no-ns", action="store_true", help="skip the Navier-Stokes sub-record")
    ap.add_argument("--no-adaptive", action="store_true", help="skip the adaptive-stepping sub-record")
    ap.add_argument("--no-curved", action="store_true", help="skip the curvilinear-mesh sub-records")
    ap.add_argument("--no-sc", action="store_true", help="skip the shock-capturing (MUSCL-Hancock) sub-record")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the config-0 (32^2 to t = 1), MDRK N=3 Euler and refined-TreeMesh sub-records")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (the reference kernel for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (element-block sharding with a "
              "face-record exchange is a design only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / 10
        log(f"copy calibration: 1 GiB read + 1 GiB write per copy, {2 * (1 << 30) / el / 1e9:.1f} GB/s")
        return

    import trixilw_hip as tlw
    from trixilw_hip import _lib

    wl = WORKLOADS[args.workload]
    N, td = wl["N"], wl["td"]
    sc = wl.get("problem") == "riemann_sc"
    par = wl.get("problem") in ("ns_shear", "ns_warped")
    curved = wl.get("problem") in ("vortex_warped", "ns_warped")
    semi, mesh, ic, cfl, (lo, hi), okw = make_problem(wl)
    dofs = semi.ndofs()
    log(f"bench: {args.workload}: {mesh.n_elements} elements, {dofs} DOFs")
    x, y = semi.node_coordinates()
    u0 = np.ascontiguousarray(ic(x, y, 0.0))
    L = _lib.lib()
    h = semi.handle
    dp = C.POINTER(C.c_double)
    r = run_steps(semi, u0, cfl, args.steps, args.warmup, args.tile, repeats=args.repeats)
    tile, tune = r["tile"], r["tune"]
    repeats = None
    if r["repeat_ms"]:
        ms = sorted(r["repeat_ms"])
        med = float(np.median(ms))
        repeats = {"n": len(ms), "steps_each": args.steps, "median_ms_per_step": round(med, 4),
                   "min_ms_per_step": round(ms[0], 4), "max_ms_per_step": round(ms[-1], 4),
                   "median_value": dofs / (med * 1e-3),
                   "note": "the same warm-up and timed steps from the same initial state, re-timed after the main run"}
        log(f"bench: {len(ms)} repeats, median {med:.3f} ms/step ({dofs / med / 1e6:.3f} G DOF-updates/s)")
    value, ms_per_step = r["value"], r["ms_per_step"]
    vms_avg, sms_avg, vn = r["vol_ms"], r["surf_ms"], r["vol_n"]
    log(f"bench: {value / 1e9:.3f} G DOF-updates/s, {ms_per_step:.3f} ms/step, "
        f"volume {vms_avg:.3f} ms x{vn}, surface {sms_avg:.3f} ms")

    step_b, _ = alg_bytes_per_dof(N, td, parabolic=par, curved=curved)
    vol_avg_s = vms_avg * 1e-3
    rec = 16 if sc else 10
    rl_unit_bytes, leave_frac = volume_kernel_bytes(semi.mesh, N, tile, rec, parabolic=par, curved=curved)
    achieved = rl_unit_bytes / vol_avg_s / 1e9
    traffic, traffic_src = load_traffic(args.workload)
    valu = load_valu(args.workload, vol_avg_s)
    roofline = {
        "bound": "hbm",
        "kernel": (("k_volume_lw4" if (td == "lw" and N == 4) else ("k_volume_lw3" if td == "lw" else "k_volume_mdrk*"))
                   + ("_navier_stokes" if par else "") + ("_curved" if curved else "")),
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "traffic_source": traffic_src,
        "alg_bytes_per_launch": rl_unit_bytes,
        "alg_basis": ("uⁿ read + updated state write (32 + 32 B per DOF) + the face records of the "
                      f"{100 * leave_frac:.1f}% of face slots that leave a {tile}-element workgroup "
                      f"({rec} doubles per face node) + element ids and J⁻¹"),
        "avg_launch_ms": round(vol_avg_s * 1e3, 4),
        "step_alg_bytes_per_dof": step_b,
        "step_frac": round(step_b * value / 1e9 / HBM_PEAK_GBS, 4),
        "step_frac_state_only": round(64.0 * (2 if td == "mdrk" else 1) * value / 1e9 / HBM_PEAK_GBS, 4),
        "surface_avg_launch_ms": round(sms_avg, 4),
        "valu": valu,
        "limiter": ("FP64 VALU issue and LDS waits (rocprof SQ counters in profiles/): the kernel moves "
                    "the bytes above at a fraction of HBM bandwidth"),
    }

    adjoint = None
    if td == "lw" and not sc and not par and not curved and not args.no_adjoint:
        # adjoint/forward step ratio: one adjoint step (tlw_adjoint_lw: face records + reverse
        # sweep) against one forward step of the loop above, same state and dt
        _lib.check(L.tlw_get_state(h, u0.ctypes.data_as(dp), u0.size), h)  # the state after the timed steps
        u = torch.from_numpy(u0).to("cuda")
        lam_a = torch.randn_like(u)
        lam_b = torch.empty_like(u)
        dt_a = cfl * semi.max_dt(u)
        for _ in range(2):
            tlw.adjoint_step(lam_b, u, lam_a, semi, 0.0, dt_a)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = max(3, min(args.steps, 10))
        e0.record()
        for r in range(reps):
            if r % 2 == 0:
                tlw.adjoint_step(lam_b, u, lam_a, semi, 0.0, dt_a)
            else:
                tlw.adjoint_step(lam_a, u, lam_b, semi, 0.0, dt_a)
        e1.record()
        torch.cuda.synchronize()
        adj_ms = e0.elapsed_time(e1) / reps
        adjoint = {"step_ms": round(adj_ms, 4), "forward_step_ms": round(ms_per_step, 4),
                   "ratio": round(adj_ms / ms_per_step, 3), "reps": reps,
                   "note": "kernel-only: one discrete-adjoint LW step (k_adj_records + k_adjoint) on the same "
                           "mesh and state as the forward step"}
        log(f"bench: adjoint step {adj_ms:.3f} ms = {adj_ms / ms_per_step:.2f} x forward step")
        del u, lam_a, lam_b
        if args.workload == DEFAULT_WORKLOAD:
            adjoint["gradient"] = measure_gradient()
            log(f"bench: gradient {adjoint['gradient']}")

    adaptive = None
    if args.workload == DEFAULT_WORKLOAD and not args.no_adaptive:
        adaptive = measure_adaptive(semi, u0, max(5, min(args.steps, 20)))  # tile chosen above

    batch = None
    if args.workload == DEFAULT_WORKLOAD and not args.no_batch:
        log("bench: 64-instance batch")
        batch = measure_batch(single_value=value)
        log(f"bench: batch {batch['value'] / 1e9:.3f} G DOF-updates/s "
            f"({batch.get('vs_single_instance')} x the single 1024^2 instance)")

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    threads = max(1, min(threads, os.cpu_count() or threads))
    navier_stokes = None
    if args.workload == DEFAULT_WORKLOAD and not args.no_ns:
        navier_stokes = measure_navier_stokes(max(3, min(args.steps, 10)), 1, threads,
                                              with_cpu=not args.no_cpu_baseline)

    sc_rec = None
    if args.workload == DEFAULT_WORKLOAD and not args.no_sc:
        # the configuration's own 500 steps (0.3 ms each at 256²)
        sc_rec = measure_shock_capturing(500, 5, threads, with_cpu=not args.no_cpu_baseline)

    curved_rec = None
    curved_ns = None
    if args.workload == DEFAULT_WORKLOAD and not args.no_curved:
        curved_rec = measure_curved(max(3, min(args.steps, 10)), 1, threads, with_cpu=not args.no_cpu_baseline)
        curved_ns = measure_curved_navier_stokes(300, 20, threads, with_cpu=not args.no_cpu_baseline)

    config0 = mdrk_euler = refined = None
    if args.workload == DEFAULT_WORKLOAD and not args.no_extra:
        config0 = measure_config0(threads, with_cpu=not args.no_cpu_baseline)
        mdrk_euler = measure_mdrk_euler(max(3, min(args.steps, 10)), 1, threads, with_cpu=not args.no_cpu_baseline)
        refined = measure_refined(max(3, min(args.steps, 10)), 1, threads, with_cpu=not args.no_cpu_baseline,
                                  uniform_value=value)

    cpu = None
    if not args.no_cpu_baseline:
        import oracle_ctypes

        nx = mesh.nx or 1
        o = (oracle_ctypes.Oracle(N, 1, 1, 0.0, 0.0, 1.0, **okw) if "tree" in okw else
             oracle_ctypes.Oracle(N, nx, nx, lo, lo, (hi - lo) / nx, **okw))
        xo, yo = o.node_coords()
        uo = np.ascontiguousarray(ic(xo, yo, 0.0))
        log(f"bench: CPU oracle baseline, {args.cpu_steps} steps on {threads} threads")
        v_all, el_all = cpu_baseline(o, uo, cfl, td == "mdrk", dofs, args.cpu_steps, threads)
        log(f"bench: CPU {v_all / 1e6:.2f} M DOF-updates/s ({el_all:.1f} s)")
        v_one, el_one = cpu_baseline(o, uo, cfl, td == "mdrk", dofs, 1, 1)
        log(f"bench: CPU single thread {v_one / 1e6:.2f} M DOF-updates/s ({el_one:.1f} s)")
        cpu = {"value": v_all, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the same N={N} mesh ({mesh.n_elements} elements, {dofs} DOFs), "
                         f"C++/OpenMP oracle, g++ -O3 -march=x86-64-v3",
               "gpu_over_cpu": round(value / v_all, 2),
               "single_thread": {"value": v_one, "cores": 1, "sample": "1 step of the same mesh",
                                 "gpu_over_cpu": round(value / v_one, 2)}}
        del o

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload,
                   "mesh": (f"{mesh.nx}x{mesh.nx} TreeMesh [0,1]^2, time-averaged Dirichlet boundary" if sc else
                            f"{mesh.nx}x{mesh.nx} periodic StructuredMesh [0,{hi:g}]^2 warped by "
                            f"{CNS_A if par else WARP_A}" if curved
                            else f"{mesh.n_elements}-leaf periodic TreeMesh [{lo:g},{hi:g}]^2 (512^2, [-5,5]^2 "
                                 "refined once)" if "tree" in okw
                            else f"{mesh.nx}x{mesh.nx} periodic TreeMesh [{lo:g},{hi:g}]^2"),
                   "polydeg": N, "time_discretization": td.upper(), "dofs": dofs,
                   "initial_condition": ("four-quadrant Riemann problem (Lax-Liu 3)" if sc else
                                         "shear flow with density/pressure waves" if par else
                                         "isentropic vortex strength 5, Mach 0.5"),
                   "parabolic": (f"CompressibleNavierStokesDiffusion2D mu {CNS_MU if curved else NS_MU}, "
                                 f"Pr {NS_PR}, conservative gradients" if par else None),
                   "shock_capturing": (("Hennemann-Gassner on density_pressure, alpha_max 0.5, "
                                        + ("MUSCL-Hancock" if wl.get("recon") == "muscl_hancock" else "first-order")
                                        + " subcell FV") if sc else None),
                   "cfl": cfl, "elements_per_workgroup": tile, "tile_autotune_ms": tune},
        "roofline": roofline,
        "repeats": repeats,
        "cpu_baseline": cpu,
        "adjoint": adjoint,
        "batch": batch,
        "navier_stokes": navier_stokes,
        "adaptive": adaptive,
        "curved": curved_rec,
        "curved_navier_stokes": curved_ns,
        "shock_capturing": sc_rec,
        "config0": config0,
        "mdrk_euler": mdrk_euler,
        "refined": refined,
    }
    print(json.dumps(out), flush=True)
    semi.close()


if __name__ == "__main__":
    main()
