"""The source terms' sin/cos (Cody-Waite reduction and Taylor polynomials,
trixilw.jl_amd/csrc/tlw_sincos.hpp) compiled for the host with g++ against the
C library's sin and cos on 2·10⁶ seeded arguments in [-200, 200]: within
1.2e-16 absolute, about one ulp of values of size one."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

PROGRAM = r"""
#include <cmath>
#include <cstdio>
#include <random>
using std::fma; using std::rint; using std::floor;
#define __device__
#define __forceinline__ inline
#include "tlw_sincos.hpp"
int main() {
    std::mt19937_64 g(1);
    std::uniform_real_distribution<double> d(-200.0, 200.0);
    double es = 0.0, ec = 0.0;
    for (int i = 0; i < 2000000; ++i) {
        const double a = i < 9 ? (i - 4) * 0.7853981633974483 : d(g);  // multiples of π/4 first
        double s, c;
        tlw::sincos_cw(a, &s, &c);
        es = std::fmax(es, std::fabs(s - std::sin(a)));
        ec = std::fmax(ec, std::fabs(c - std::cos(a)));
    }
    std::printf("%.17g %.17g\n", es, ec);
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_sincos_cw_matches_libm(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "t"
    subprocess.run(["g++", "-O2", "-I", str(ROOT / "trixilw.jl_amd" / "csrc"), "-o", str(exe), str(src)], check=True)
    es, ec = (float(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split())
    assert es <= 1.2e-16 and ec <= 1.2e-16, (es, ec)
