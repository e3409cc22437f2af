"""tlw_max_dt_noalloc (the allocation-free, context-free max_dt) against
tlw_max_dt through the C ABI: the same bits on two calls in a row, the second
on a smaller mesh, for polydeg 3 and 4, on a side stream; NaN for a
non-admissible state; sizes beyond the scratch capacity are refused before any
launch."""
import ctypes as C

import numpy as np
import pytest

from gpu_helpers import setup, to_dev, torch_mod

pytestmark = pytest.mark.gpu


def _inv_jac(torch, semi):
    return torch.from_numpy(np.ascontiguousarray(semi.mesh.inverse_jacobian, dtype=np.float64)).cuda()


@pytest.mark.parametrize("N", [3, 4])
def test_noalloc_matches_max_dt_twice(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    cases = []
    for level, seed in ((5, 11), (3, 12)):  # 32² first, then the smaller 8²
        semi, o, u_lex, mesh = setup(oracle_mod, N, level, ordering="morton", seed=seed + N, amp=5e-2,
                                     keep_aux=False)
        u = to_dev(mesh, u_lex)
        ref = semi.max_dt(u)
        assert ref == pytest.approx(o.max_dt(u_lex), rel=1e-14)
        cases.append((semi, u, ref))
    stream = torch.cuda.Stream()
    outs = []
    with torch.cuda.stream(stream):
        for semi, u, _ in cases:
            out = torch.full((1,), -7.0, dtype=torch.float64, device="cuda")
            tlw.max_dt_noalloc(u, _inv_jac(torch, semi), N, 1.4, out, stream=stream)
            outs.append(out)
    stream.synchronize()
    for (semi, _, ref), out in zip(cases, outs):
        got = float(out.cpu()[0])
        assert got == ref, (got, ref)  # bit-identical to tlw_max_dt
        semi.close()


def test_noalloc_flags_non_admissible(oracle_mod):
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = setup(oracle_mod, 4, 3, seed=3, keep_aux=False)
    u_bad = u_lex.copy()
    u_bad[5, 2, 3, 0] = -1.0  # negative density at one node
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    tlw.max_dt_noalloc(to_dev(mesh, u_bad), _inv_jac(torch, semi), 4, 1.4, out)
    torch.cuda.synchronize()
    assert np.isnan(float(out.cpu()[0]))
    # an admissible state afterwards is unaffected by the flagged call
    u = to_dev(mesh, u_lex)
    tlw.max_dt_noalloc(u, _inv_jac(torch, semi), 4, 1.4, out)
    torch.cuda.synchronize()
    assert float(out.cpu()[0]) == semi.max_dt(u)
    semi.close()


def test_noalloc_refuses_oversize_and_bad_arguments():
    import trixilw_hip as tlw
    from trixilw_hip import _lib

    torch = torch_mod()
    L = _lib.lib()
    buf = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = C.c_void_p(buf.data_ptr())
    # checked on the host before any launch: the kernels never see these sizes
    assert L.tlw_max_dt_noalloc(p, p, 1_310_721, 4, 1.4, p, None) == _lib.TLW_EINVAL
    assert b"scratch capacity" in L.tlw_last_error(None)
    assert L.tlw_max_dt_noalloc(p, p, 2_097_153, 3, 1.4, p, None) == _lib.TLW_EINVAL
    assert L.tlw_max_dt_noalloc(p, p, 16, 5, 1.4, p, None) == _lib.TLW_EINVAL
    assert L.tlw_max_dt_noalloc(None, p, 16, 4, 1.4, p, None) == _lib.TLW_EINVAL
    with pytest.raises(tlw._lib.TLWError):
        _lib.check(L.tlw_max_dt_noalloc(p, p, 0, 4, 1.4, p, None))
