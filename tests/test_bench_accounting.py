"""The benchmark's accounting on the CPU: compulsory volume-kernel bytes on
meshes where the leaving-face count is known by hand, the committed rocprof
figures it reads (HBM bytes, FP64 instruction counts) and the one-GPU rule."""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


@pytest.fixture(scope="module")
def bench():
    import bench as b
    return b


def tree(level, ordering="morton"):
    import trixilw_hip as tlw
    return tlw.TreeMesh((-1.0, -1.0), (1.0, 1.0), level, ordering=ordering)


@pytest.mark.parametrize("N", [3, 4])
def test_volume_bytes_one_tile_and_single_elements(bench, N):
    """A periodic 4×4 mesh in one 16-element tile has no leaving face: only the
    state read/write and the per-element ids and J⁻¹ remain.  With 1-element
    tiles every face slot leaves, and Navier-Stokes adds 4 doubles per face node."""
    n = N + 1
    m = tree(2)
    E = 16
    b, frac = bench.volume_kernel_bytes(m, N, 16)
    assert frac == 0.0
    assert b == 8.0 * 2 * 4 * n * n * E + 24.0 * E
    b1, frac1 = bench.volume_kernel_bytes(m, N, 1)
    assert frac1 == 1.0
    assert b1 == b + 8.0 * 4 * E * n * 10
    bp, _ = bench.volume_kernel_bytes(m, N, 1, parabolic=True)
    assert bp == b1 + 8.0 * 4 * E * n * 4
    bp0, _ = bench.volume_kernel_bytes(m, N, 16, parabolic=True)
    assert bp0 == b


def test_strips_leave_fewer_faces_than_morton(bench):
    """Ten-element tiles: two-row strips (5×2 blocks, 14 of 40 slots leave) beat
    Morton runs on a 1024² mesh; the strips fraction is 0.35 away from the
    strip seams."""
    _, f_strips = bench.volume_kernel_bytes(tree(10, "strips"), 4, 10)
    _, f_morton = bench.volume_kernel_bytes(tree(10, "morton"), 4, 10)
    assert abs(f_strips - 0.35) < 0.002
    assert f_strips < f_morton


def test_committed_profile_figures(bench):
    """valu.json's flops are 64 lanes × (2 FMA + MUL + ADD + TRANS) of its own
    instruction counts; traffic.json's bytes are the scaled FETCH/WRITE sizes of
    its summary; both name summaries that exist."""
    valu = json.loads((ROOT / "profiles" / "valu.json").read_text())
    traffic = json.loads((ROOT / "profiles" / "traffic.json").read_text())
    for wl in ("lw_euler_vortex_1024x1024_N4", "lw_navierstokes_shear_1024x1024_N4"):
        v = valu[wl]
        c = v["f64_insts_per_launch"]
        assert v["flops_per_launch"] == 64.0 * (2 * c["fma"] + c["mul"] + c["add"] + c["trans"])
        assert (ROOT / v["source"]).exists()
        t = traffic[wl]
        assert (ROOT / t["source"]).exists()
        assert t["kernel"] == v["kernel"]
        r = bench.load_valu(wl, 1e-3)
        assert r["peak"] == 78.6 and abs(r["achieved"] - v["flops_per_launch"] / 1e-3 / 1e12) < 0.01
        assert bench.load_traffic(wl)[0] == t["bytes"]
    # the measured bytes are within 20 % above the compulsory ones on the 1024² strips mesh
    comp, _ = bench.volume_kernel_bytes(tree(10, "strips"), 4, 10)
    assert 1.0 <= traffic["lw_euler_vortex_1024x1024_N4"]["bytes"] / comp < 1.2
    comp_ns, _ = bench.volume_kernel_bytes(tree(10, "strips"), 4, 10, parabolic=True)
    assert 1.0 <= traffic["lw_navierstokes_shear_1024x1024_N4"]["bytes"] / comp_ns < 1.2


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "--gpus 1" in r.stderr
