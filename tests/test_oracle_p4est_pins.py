"""Pins of the oracle's P4estMesh path (p4est/dg_2d.jl:8-393: interfaces with
each side's own normal, the primary's normal for λ, boundaries in the
conditions' P4est methods) on the warped mapping's cells:

- element frames turned by random multiples of 90° (the orientations of an
  unstructured quadrilateral mesh) give the aligned mesh's residual at every
  physical node (LW, both MDRK stages, periodic and with boundaries), so face
  orientation, flips and normal signs are all consistent;
- the aligned P4est mesh equals the StructuredMesh path up to round-off (the two
  differ only in whose contravariant vector a shared face node uses), with
  outflow and the approximate slip wall too;
- the connectivity: every face is an interface side or a boundary face once,
  flips occur, and partners' face nodes coincide in space;
- the Dirichlet flux against a numpy evaluation of its P4est method (λ from uⁿ
  and the time-averaged outer state, jump on Ũ, outward normal), free stream
  with a constant Dirichlet state in LW and in both MDRK stages (stage 1's half
  time average), conservation, and the vortex's convergence order on a turned
  mesh."""
import numpy as np
import pytest

from conftest import rel_linf

L, A = 20.0, 0.08


def mk(oracle_mod, N, nx, p4est=True, periodic=(True, True), **kw):
    return oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, L / nx, mapping=(oracle_mod.MAP_WARP, (L, A)), periodic=periodic,
                             p4est=p4est, **kw)


def turns(E, seed):
    return np.random.default_rng(seed).integers(0, 4, E)


def to_frame(a, r):
    """Per element: values at the aligned frame's nodes [j, i] -> the turned frame's.
    A node (i, j) of a frame turned by r·90° sits at the aligned node that turn(r)
    maps it to."""
    n = a.shape[1]
    out = np.empty_like(a)
    idx = np.arange(n)
    for e in range(a.shape[0]):
        i, j = np.meshgrid(idx, idx)  # i varies along axis 1
        k = r[e] % 4
        if k == 0:
            ai, aj = i, j
        elif k == 1:
            ai, aj = n - 1 - j, i
        elif k == 2:
            ai, aj = n - 1 - i, n - 1 - j
        else:
            ai, aj = j, n - 1 - i
        out[e] = a[e][aj, ai]
    return out


def state(x, y, seed=0):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.2 * np.sin(0.3 * x) * np.cos(0.25 * y) + 0.01 * rng.standard_normal(x.shape)
    v1 = 0.4 + 0.1 * np.cos(0.2 * y)
    v2 = -0.3 + 0.1 * np.sin(0.15 * x)
    p = 1.0 + 0.1 * np.cos(0.2 * (x + y))
    return np.stack([rho, rho * v1, rho * v2, p / 0.4 + 0.5 * rho * (v1**2 + v2**2)], axis=-1)


def residuals(o, u, dt, kind):
    if kind == "lw":
        return o.rhs(u, 0.3, dt)
    d1 = o.rhs(u, 0.3, dt, kind="mdrk1")
    if kind == "mdrk1":
        return d1
    o.set_us(u + 0.5 * dt * d1)
    return o.rhs(u, 0.3, dt, kind="mdrk2")


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("kind", ["lw", "mdrk1", "mdrk2"])
@pytest.mark.parametrize("bc", ["periodic", "vortex", "outflow"])
def test_turned_frames_give_the_aligned_residual(oracle_mod, N, kind, bc):
    from trixilw_hip.equations import IsentropicVortex

    nx = 5
    kw = {}
    if bc == "vortex":
        ic = IsentropicVortex(center=(L / 2, L / 2), period=(1e9, 1e9))
        kw = dict(periodic=(False, False), bc_kind=ic.kind, bc_par=ic.params)
    elif bc == "outflow":
        kw = dict(periodic=(False, False), bc_kind=oracle_mod.BV_OUTFLOW)
    r = turns(nx * nx, 7 + N)
    o0 = mk(oracle_mod, N, nx, **kw)
    o1 = mk(oracle_mod, N, nx, p4est=r, **kw)
    itf, _ = o1.p4est_connectivity()
    assert np.any(itf[:, 4] == 1)  # flipped interfaces occur
    x, y = o0.node_coords()
    x1, y1 = o1.node_coords()
    assert np.max(np.abs(to_frame(x[..., None], r)[..., 0] - x1)) < 1e-12
    assert np.max(np.abs(to_frame(y[..., None], r)[..., 0] - y1)) < 1e-12
    u = state(x, y)
    dt = 0.4 * o0.max_dt(u)
    assert o1.max_dt(to_frame(u, r)) == pytest.approx(dt / 0.4, rel=1e-13)
    d0 = residuals(o0, u, dt, kind)
    d1 = residuals(o1, np.ascontiguousarray(to_frame(u, r)), dt, kind)
    assert rel_linf(d1, to_frame(d0, r)) < 1e-12


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("bc", ["periodic", "outflow", "slip"])
def test_aligned_p4est_equals_structured_up_to_round_off(oracle_mod, N, bc):
    kw = {}
    if bc != "periodic":
        kw = dict(periodic=(False, False),
                  bc_kind=oracle_mod.BV_OUTFLOW if bc == "outflow" else oracle_mod.BV_SLIP_WALL_APPROX)
    o_s = mk(oracle_mod, N, 6, p4est=None, **kw)
    o_p = mk(oracle_mod, N, 6, **kw)
    x, y = o_s.node_coords()
    u = state(x, y, seed=3)
    dt = 0.4 * o_s.max_dt(u)
    for kind in ("lw", "mdrk1", "mdrk2"):
        assert rel_linf(residuals(o_p, u, dt, kind), residuals(o_s, u, dt, kind)) < 1e-12


@pytest.mark.parametrize("periodic", [True, False])
def test_connectivity_covers_every_face_once(oracle_mod, periodic):
    nx, N = 6, 4
    r = turns(nx * nx, 1)
    o = mk(oracle_mod, N, nx, p4est=r, periodic=(periodic, periodic), bc_kind=oracle_mod.BV_OUTFLOW)
    itf, bd = o.p4est_connectivity()
    sides = [(a, f) for a, f in zip(itf[:, 0], itf[:, 1])] + [(a, f) for a, f in zip(itf[:, 2], itf[:, 3])]
    sides += [tuple(b) for b in bd]
    assert sorted(sides) == sorted((e, f) for e in range(nx * nx) for f in range(4))
    assert (len(bd) == 0) == periodic and (periodic or len(bd) == 4 * nx)
    x, y = o.node_coords()
    n = N + 1

    def face_xy(e, f):
        i = {0: 0, 1: n - 1}.get(f, None)
        j = {2: 0, 3: n - 1}.get(f, None)
        if i is not None:
            return x[e, :, i], y[e, :, i]
        return x[e, j, :], y[e, j, :]

    for p, pf, s, sf, flip in itf:
        xp, yp = face_xy(p, pf)
        xs, ys = face_xy(s, sf)
        if flip:
            xs, ys = xs[::-1], ys[::-1]
        # periodic partners differ by a period in one coordinate
        dx = np.abs(xp - xs) % L
        dy = np.abs(yp - ys) % L
        assert np.all(np.minimum(dx, L - dx) < 1e-9) and np.all(np.minimum(dy, L - dy) < 1e-9)


@pytest.mark.parametrize("N", [3, 4])
def test_dirichlet_flux_follows_its_p4est_method(oracle_mod, N):
    """sfv at boundary face nodes against numpy: the time averages of the vortex
    and of its normal flux (outward normal), then ½(F̃·n + F_o) - ½λ(uⁿ, U_o)(U_o - Ũ),
    stored times ±1 for +/- faces."""
    from trixilw_hip.equations import IsentropicVortex

    ic = IsentropicVortex(center=(L / 2, L / 2), period=(1e9, 1e9))
    nx, n = 4, N + 1
    r = turns(nx * nx, 4)
    o = mk(oracle_mod, N, nx, p4est=r, periodic=(False, False), bc_kind=ic.kind, bc_par=ic.params)
    x, y = o.node_coords()
    u = state(x, y, seed=9)
    t, dt = 0.3, 0.4 * o.max_dt(u)
    o.rhs(u, t, dt)
    U = o.aux("U").reshape(u.shape)
    F = o.aux("F").reshape(u.shape[:-1] + (2, 4))
    Ja = o.aux("contravariant_vectors").reshape(u.shape[:-1] + (4,))
    sfv = o.aux("surface_flux_values").reshape(nx * nx, 4, n, 4)
    nodes, w, _, _ = o.basis()
    _, bd = o.p4est_connectivity()
    g = 1.4

    def nflux(q, nrm):
        rho = q[..., 0]
        v1, v2 = q[..., 1] / rho, q[..., 2] / rho
        p = (g - 1) * (q[..., 3] - 0.5 * rho * (v1**2 + v2**2))
        vn = v1 * nrm[0] + v2 * nrm[1]
        return np.stack([rho * vn, rho * vn * v1 + p * nrm[0], rho * vn * v2 + p * nrm[1], (q[..., 3] + p) * vn], -1)

    def speed(q, nrm):
        rho = q[0]
        v1, v2 = q[1] / rho, q[2] / rho
        p = (g - 1) * (q[3] - 0.5 * rho * (v1**2 + v2**2))
        return abs(v1 * nrm[0] + v2 * nrm[1]), np.sqrt(g * p / rho) * np.hypot(*nrm)

    worst = 0.0
    for e, f in bd:
        s = 1.0 if f % 2 else -1.0
        for l in range(n):
            i, j = {0: (0, l), 1: (n - 1, l), 2: (l, 0), 3: (l, n - 1)}[f]
            J = Ja[e, j, i]
            nrm = s * (J[:2] if f < 2 else J[2:])
            ts = t + 0.5 * dt * (nodes + 1.0)
            ub = np.array([ic(np.array([x[e, j, i]]), np.array([y[e, j, i]]), tq)[0] for tq in ts])
            Uo = np.einsum("q,qv->v", w / 2.0, ub)
            Fo = np.einsum("q,qv->v", w / 2.0, nflux(ub, nrm))
            fin = F[e, j, i, 0] * nrm[0] + F[e, j, i, 1] * nrm[1]
            a1, c1 = speed(u[e, j, i], nrm)
            a2, c2 = speed(Uo, nrm)
            want = s * (0.5 * (fin + Fo) - 0.5 * (max(a1, a2) + max(c1, c2)) * (Uo - U[e, j, i]))
            worst = max(worst, np.max(np.abs(sfv[e, f, l] - want)) / np.max(np.abs(want)))
    assert worst < 1e-13


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("kind", ["lw", "mdrk"])
def test_free_stream_with_constant_dirichlet(oracle_mod, N, kind):
    """A uniform flow held by its own constant Dirichlet state on a turned,
    non-periodic mesh: the LW step and both MDRK stages (stage 1 with the
    boundary's half time average) leave it unchanged."""
    from trixilw_hip.equations import ConstantState

    cs = ConstantState(1.2, 0.3, -0.5, 0.9)
    nx = 5
    o = mk(oracle_mod, N, nx, p4est=turns(nx * nx, 2), periodic=(False, False), bc_kind=cs.kind, bc_par=cs.params)
    x, y = o.node_coords()
    u = cs(x, y, 0.0)
    dt = 0.4 * o.max_dt(u)
    if kind == "lw":
        assert np.max(np.abs(o.rhs(u, 0.0, dt))) < 1e-12
    else:
        d1 = o.rhs(u, 0.0, dt, kind="mdrk1")
        assert np.max(np.abs(d1)) < 1e-12
        o.set_us(u)
        assert np.max(np.abs(o.rhs(u, 0.0, dt, kind="mdrk2"))) < 1e-12


@pytest.mark.parametrize("N", [3, 4])
def test_turned_mesh_conserves(oracle_mod, N):
    nx = 6
    o = mk(oracle_mod, N, nx, p4est=turns(nx * nx, 5))
    x, y = o.node_coords()
    u = state(x, y, seed=1)
    dt = 0.4 * o.max_dt(u)
    _, w, _, _ = o.basis()
    Wq = np.outer(w, w)[None] / o.aux("inverse_jacobian").reshape(x.shape)
    for kind in ("lw", "mdrk1", "mdrk2"):
        du = residuals(o, u, dt, kind)
        tot = np.einsum("eji,ejiv->v", Wq, du)
        assert np.all(np.abs(tot) <= 1e-13 * np.einsum("eji,ejiv->v", Wq, np.abs(du))), (kind, tot)


@pytest.mark.parametrize("N,mdrk,eoc_min", [(4, False, 3.5), (3, True, 3.2)])
def test_turned_mesh_vortex_convergence(oracle_mod, N, mdrk, eoc_min):
    from trixilw_hip.equations import IsentropicVortex

    ic = IsentropicVortex(center=(L / 2, L / 2), period=(L, L))
    fc = {3: 0.170, 4: 0.103}[N]
    cfl = 0.9 * (N + 1) * fc * (0.5 if mdrk else 1.0)
    errs = []
    for nx in (8, 16, 32):
        o = mk(oracle_mod, N, nx, p4est=turns(nx * nx, nx))
        x, y = o.node_coords()
        u, t, _ = o.solve(ic(x, y, 0.0), 0.0, 1.0, cfl, mdrk=mdrk)
        wj = 1.0 / o.aux("inverse_jacobian").reshape(x.shape)
        errs.append(np.sqrt(np.sum(wj * (u - ic(x, y, t))[..., 0] ** 2) / np.sum(wj)))
    eoc = np.log(errs[1] / errs[2]) / np.log(2)
    assert eoc > eoc_min, (errs, eoc)


def test_p4est_refuses_what_it_does_not_run(oracle_mod):
    for kw in (dict(shock_capturing=True), dict(parabolic="navier_stokes", mu=0.01),
               dict(periodic=(False, False), bc_kind=oracle_mod.BV_SLIP_WALL)):
        with pytest.raises(ValueError):
            mk(oracle_mod, 3, 4, **kw)
