import os
import sys
from pathlib import Path

# Under pytest-xdist each worker gets its share of the cores for the oracle's
# OpenMP loops: 8 workers with 8 spinning threads each oversubscribe the host
# many times over.  Set before any OpenMP runtime is loaded.
_workers = int(os.environ.get("PYTEST_XDIST_WORKER_COUNT", "1") or 1)
if _workers > 1:
    os.environ.setdefault("OMP_NUM_THREADS", str(max(1, (os.cpu_count() or 1) // _workers)))
    os.environ.setdefault("OMP_WAIT_POLICY", "passive")

import pytest  # noqa: E402

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "trixilw.jl_amd", ROOT / "oracle", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtrixilw_hip.so)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle_ctypes

    oracle_ctypes.lib()
    return oracle_ctypes


def rel_l2(a, b):
    import numpy as np

    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def rel_linf(a, b):
    import numpy as np

    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
