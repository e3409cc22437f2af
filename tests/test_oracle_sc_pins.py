"""Pins of the oracle's shock-capturing path (no GPU).

VolumeIntegralFRShockCapturing with IndicatorHennemannGassner on
density_pressure and first-order subcell FV (src/solvers/fr.jl:63-86,
src/solvers/dgsem_tree/dg_2d.jl:251-282, 1648-1804, 1873-1991).  The
reference ships no golden vectors for it, so each piece is pinned against
mathematics the test computes independently:
  * the normalized Legendre Vandermonde matrix against numpy's Legendre
    polynomials;
  * the indicator against states whose modal content is chosen by
    construction (both energy ratios, the clipping, the alpha_max cap and
    the interface smoothing);
  * alpha = 1 against a textbook first-order Rusanov FV scheme on the GLL
    subcell grid (the whole residual, surface terms included);
  * the blend against the dense linear closed form with per-element alpha,
    and for Euler against affinity in a constant alpha.
The admissibility fallback of compute_alp (alp = 1 when a low-order update
is not admissible) has no independent pin here; the GPU parity tests cover it
on the four-quadrant Riemann problem.
"""
import numpy as np
import pytest

from conftest import rel_l2, rel_linf
from linear_closed_form import LinearLW

GAMMA = 1.4


def euler_flux(u, orientation, g=GAMMA):
    rho, m1, m2, E = u[..., 0], u[..., 1], u[..., 2], u[..., 3]
    v1, v2 = m1 / rho, m2 / rho
    p = (g - 1) * (E - 0.5 * (m1 * v1 + m2 * v2))
    if orientation == 1:
        return np.stack([m1, m1 * v1 + p, m1 * v2, (E + p) * v1], -1)
    return np.stack([m2, m2 * v1, m2 * v2 + p, (E + p) * v2], -1)


def rusanov(ul, ur, orientation, g=GAMMA):
    def parts(u):
        rho = u[..., 0]
        v = u[..., orientation] / rho
        p = (g - 1) * (u[..., 3] - 0.5 * (u[..., 1] ** 2 + u[..., 2] ** 2) / rho)
        return np.abs(v), np.sqrt(g * p / rho)

    al, cl = parts(ul)
    ar, cr = parts(ur)
    lam = np.maximum(al, ar) + np.maximum(cl, cr)
    return 0.5 * (euler_flux(ul, orientation, g) + euler_flux(ur, orientation, g)) - 0.5 * lam[..., None] * (ur - ul)


def subcell_fv_global(u, nx, ny, w, h):
    """First-order Rusanov FV on the grid of all GLL nodes (periodic), cell
    width w_i h / 2: du = -(F_{k+1/2} - F_{k-1/2}) / (w_i h / 2) per direction."""
    E, n, _, nv = u.shape
    # global array [gy, gx, v] with gx = ix*n + i, gy = iy*n + j
    G = u.reshape(ny, nx, n, n, nv).transpose(0, 2, 1, 3, 4).reshape(ny * n, nx * n, nv)
    wx = np.tile(w, nx)
    wy = np.tile(w, ny)
    fx = rusanov(G, np.roll(G, -1, axis=1), 1)  # flux between column k and k+1
    gy = rusanov(G, np.roll(G, -1, axis=0), 2)
    div = (fx - np.roll(fx, 1, axis=1)) / (0.5 * h * wx[None, :, None])
    div += (gy - np.roll(gy, 1, axis=0)) / (0.5 * h * wy[:, None, None])
    du = -div
    return du.reshape(ny, n, nx, n, nv).transpose(0, 2, 1, 3, 4).reshape(E, n, n, nv)


def euler_state(rho, v1, v2, p, g=GAMMA):
    return np.stack([rho, rho * v1, rho * v2, p / (g - 1) + 0.5 * rho * (v1 ** 2 + v2 ** 2)], -1)


def hg_alpha(energy, n, alpha_min, alpha_max):
    """Hennemann et al. (JCP 426, 2021, eq. 46-47) with Trixi's defaults."""
    threshold = 0.5 * 10 ** (-1.8 * n ** 0.25)
    s = np.log((1 - 1e-4) / 1e-4)
    a = 1 / (1 + np.exp(-s / threshold * (energy - threshold)))
    if a < alpha_min:
        a = 0.0
    if a > 1 - alpha_min:
        a = 1.0
    return min(alpha_max, a)


def sc_oracle(oracle_mod, N, nx, ny=None, L=2.0, x0=-1.0, **kw):
    ny = nx if ny is None else ny
    return oracle_mod.Oracle(N, nx, ny, x0, x0, L / nx, shock_capturing=True, **kw)


# ---------------------------------------------------------------- Vandermonde
@pytest.mark.parametrize("N", [3, 4])
def test_inverse_vandermonde_maps_normalized_legendre_to_unit_modes(oracle_mod, N):
    from numpy.polynomial import legendre
    from trixilw_hip.basis import LobattoLegendreBasis

    o = sc_oracle(oracle_mod, N, 2)
    Vinv = o.inverse_vandermonde()
    nodes = o.basis()[0]
    for m in range(N + 1):
        c = np.zeros(m + 1)
        c[m] = np.sqrt(m + 0.5)
        vals = legendre.legval(nodes, c)
        e = np.zeros(N + 1)
        e[m] = 1
        assert np.max(np.abs(Vinv @ vals - e)) < 1e-13
    b = LobattoLegendreBasis(N)
    assert np.max(np.abs(b.inverse_vandermonde_legendre - Vinv)) < 1e-13


# ---------------------------------------------------------------- indicator
def _modal_pressure(N, nodes, modes, p0):
    """p(ξ, η) = p0 + Σ c_ab φ_a(ξ) φ_b(η), φ_m = sqrt(m+½) P_m; returns [j, i]."""
    from numpy.polynomial import legendre

    def phi(m, x):
        c = np.zeros(m + 1)
        c[m] = np.sqrt(m + 0.5)
        return legendre.legval(x, c)

    p = np.full((N + 1, N + 1), p0)
    for (a, b), c in modes.items():
        p += c * np.outer(phi(b, nodes), phi(a, nodes))
    return p


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("branch", ["top", "second"])
@pytest.mark.parametrize("delta", [-0.3, 0.0, 0.25, 2.0])
def test_indicator_on_constructed_modes(oracle_mod, N, branch, delta):
    """ρ = 1, v = 0 makes density_pressure equal p.  With modal coefficient
    A on the constant mode and B on mode (N, 0) ('top': energy ratio
    (total - clip1)/total) or (N-1, 0) ('second': (clip1 - clip2)/clip1) the
    energy is B²/(A² + B²); B is chosen so it sits at threshold·(1 + delta)."""
    n = N + 1
    o = sc_oracle(oracle_mod, N, 3, alpha_max=1.0, alpha_smooth=False)
    nodes = o.basis()[0]
    threshold = 0.5 * 10 ** (-1.8 * n ** 0.25)
    energy = threshold * (1 + delta)
    p0 = 2.0
    A = p0 * 2  # p0 = A φ_0 φ_0 = A / 2
    B = A * np.sqrt(energy / (1 - energy))
    mode = (N, 0) if branch == "top" else (N - 1, 0)
    hot = 4
    p = np.full((o.E, n, n), p0)
    p[hot] = _modal_pressure(N, nodes, {mode: B}, p0)
    assert np.all(p > 0)
    u = euler_state(np.ones_like(p), 0 * p, 0 * p, p)
    o.rhs(u, 0.0, 1e-4)
    alpha = o.aux("alpha")
    expect = hg_alpha(energy, n, 0.001, 1.0)
    assert abs(alpha[hot] - expect) < 1e-9, (alpha[hot], expect)
    assert np.all(np.delete(alpha, hot) == 0.0)


def test_indicator_cap_and_smoothing(oracle_mod):
    """A pure top mode has energy 1: α = α_max; smoothing lifts the four face
    neighbours to α_max/2 and leaves corner neighbours at 0."""
    N, n = 4, 5
    o = sc_oracle(oracle_mod, N, 5, alpha_max=0.5, alpha_smooth=True)
    nodes = o.basis()[0]
    p = np.full((o.E, n, n), 3.0)
    hot = 12  # (2, 2)
    p[hot] = _modal_pressure(N, nodes, {(N, N): 0.5}, 3.0)
    u = euler_state(np.ones_like(p), 0 * p, 0 * p, p)
    o.rhs(u, 0.0, 1e-4)
    alpha = o.aux("alpha")
    expect = np.zeros(o.E)
    expect[hot] = 0.5
    for nb in (11, 13, 7, 17):
        expect[nb] = 0.25
    assert np.array_equal(alpha, expect), alpha.reshape(5, 5)


# ---------------------------------------------------------------- FV limit
@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("kind", ["lw", "mdrk1", "mdrk2"])
def test_alpha_one_is_first_order_subcell_fv(oracle_mod, N, kind):
    """With α = 1 on every element the LW term vanishes and every interface
    takes the Rusanov flux of uⁿ: the residual is first-order FV on the GLL
    subcell grid, in every stage."""
    from trixilw_hip.equations import IsentropicVortex

    nx = 4
    o = sc_oracle(oracle_mod, N, nx, L=20.0, x0=-10.0)
    x, y = o.node_coords()
    rng = np.random.default_rng(5 + N)
    u = IsentropicVortex()(x, y, 0.0) * (1 + 1e-2 * rng.standard_normal(x.shape + (1,)))
    o.set_alpha(np.ones(o.E))
    dt = 0.3 * o.max_dt(u)
    if kind == "mdrk2":  # stage 2 reads the F2/U2 cache that stage 1 writes
        o.rhs(u, 0.0, dt, kind="mdrk1")
        o.set_us(u + 0.01 * rng.standard_normal(u.shape))
    du = o.rhs(u, 0.0, dt, kind=kind)
    _, w, _, _ = o.basis()
    ref = subcell_fv_global(u, nx, nx, w, 20.0 / nx)
    assert rel_linf(du, ref) < 1e-13


# ---------------------------------------------------------------- blend
@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("flux_order", [0, 1])
def test_linear_blend_matches_closed_form(oracle_mod, N, flux_order):
    nx, ny, h = 3, 3, 0.5
    a = (1.0, -0.6)
    o = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, equation="linadv", adv=a, flux_order=flux_order,
                          shock_capturing=True)
    nodes, w, D, Dh = o.basis()
    ref = LinearLW(nodes, w, D, Dh, nx, ny, h, a=a, flux_order="tree" if flux_order == 0 else "structured")
    rng = np.random.default_rng(77 + N)
    u = rng.standard_normal((nx * ny, N + 1, N + 1))
    alpha = rng.uniform(0, 1, nx * ny)
    alpha[[1, 5]] = 0.0  # pure DG elements take the fn_low-only path
    o.set_alpha(alpha)
    dt = 0.011
    du = o.rhs(u[..., None], 0.0, dt)[..., 0]
    _, du_ref = ref.lw_step(u, dt, alpha)
    assert rel_linf(du, du_ref) < 1e-14
    du1 = o.rhs(u[..., None], 0.0, dt, kind="mdrk1")[..., 0]
    _, du1_ref, du2_ref, us_ref = ref.mdrk_step(u, dt, alpha)
    assert rel_linf(du1, du1_ref) < 1e-14
    o.set_us((u + dt * du1)[..., None])
    du2 = o.rhs(u[..., None], 0.0, dt, kind="mdrk2")[..., 0]
    assert rel_linf(du2, du2_ref) < 1e-14
    # the blend matters: the pure LW residual is far from it
    o.set_alpha(np.zeros(nx * ny))
    assert rel_linf(o.rhs(u[..., None], 0.0, dt)[..., 0], du_ref) > 1e-3


@pytest.mark.parametrize("N", [3, 4])
def test_euler_blend_is_affine_in_constant_alpha(oracle_mod, N):
    from trixilw_hip.equations import IsentropicVortex

    o = sc_oracle(oracle_mod, N, 4, L=20.0, x0=-10.0)
    x, y = o.node_coords()
    u = IsentropicVortex()(x, y, 0.0)
    dt = 0.2 * o.max_dt(u)
    res = {}
    for c in (0.0, 0.35, 1.0):
        o.set_alpha(np.full(o.E, c))
        res[c] = o.rhs(u, 0.0, dt)
    assert rel_linf(res[0.35], 0.65 * res[0.0] + 0.35 * res[1.0]) < 1e-13
    assert rel_linf(res[0.0], res[1.0]) > 1e-3


# ---------------------------------------------------------------- invariants
@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("kind", ["lw", "mdrk"])
def test_sc_free_stream_and_conservation(oracle_mod, N, kind):
    """Free stream: LW with any α; MDRK only with a uniform α, because stage 1
    blends the half-weighted F̃ = f/2 + f_t/8 with full-weight subcell and
    interface FV fluxes (tree/multi_derivative_2d.jl:151-204), so α_e ≠ α_nbr
    leaves (α_e - alp)·f/2 at face nodes, as in the reference.  Conservation
    holds for both with the indicator's α."""
    from trixilw_hip.equations import ConstantState, IsentropicVortex

    o = sc_oracle(oracle_mod, N, 4)
    x, y = o.node_coords()
    rng = np.random.default_rng(21)
    o.set_alpha(rng.uniform(0, 0.5, o.E) if kind == "lw" else np.full(o.E, 0.3))
    u = ConstantState(1.2, 0.3, -0.5, 0.9)(x, y, 0.0)
    dt = 0.3 * o.max_dt(u)

    def residual(u):
        if kind == "lw":
            return o.rhs(u, 0.0, dt)
        du = o.rhs(u, 0.0, dt, kind="mdrk1")
        o.set_us(u + dt * du)
        return o.rhs(u, 0.0, dt, kind="mdrk2")

    assert np.max(np.abs(residual(u))) < 1e-12
    if kind == "mdrk":
        o.set_alpha(rng.uniform(0, 0.5, o.E))
        assert np.max(np.abs(residual(u))) > 1e-3
    o.set_alpha(None)  # indicator from the state
    u = IsentropicVortex(center=(0.2, -0.1))(x * 10, y * 10, 0.0)
    u = u * (1 + 0.05 * rng.standard_normal(x.shape + (1,)))
    dt = 0.3 * o.max_dt(u)
    du = residual(u)
    assert np.any(o.aux("alpha") > 0)
    _, w, _, _ = o.basis()
    W = np.outer(w, w)
    total = np.einsum("ji,ejiv->v", W, du)
    scale = np.einsum("ji,ejiv->v", W, np.abs(du))
    assert np.all(np.abs(total) <= 1e-13 * scale)


# ---------------------------------------------------------------- admissibility fallback
def test_admissibility_fallback_switches_to_rusanov(oracle_mod):
    """compute_alp (dg_2d.jl:1879-1901) with α = 0 everywhere: an interface
    keeps the LW flux F* unless the low-order update of either adjacent node,
    u - Δt J⁻¹/(w λx) (F* - f_low) with λx = ½, leaves the admissible set;
    then it takes the Rusanov flux fn of uⁿ.  The test recomputes fn, the
    subcell fluxes f_low next to the face and the decision in numpy, rebuilds
    the residual from the plain LW one and counts a mix of both outcomes."""
    N, n, nx = 4, 5, 6
    h = 1.0 / nx
    plain = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, h)
    sc = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, h, shock_capturing=True)
    sc.set_alpha(np.zeros(sc.E))
    rng = np.random.default_rng(8)
    E = nx * nx
    rho = rng.uniform(0.2, 1.0, E)[:, None, None] * (1 + 0.05 * rng.standard_normal((E, n, n)))
    p = rng.uniform(0.1, 1.0, E)[:, None, None] * (1 + 0.05 * rng.standard_normal((E, n, n)))
    v1 = rng.uniform(-1, 1, E)[:, None, None] + 0.05 * rng.standard_normal((E, n, n))
    v2 = rng.uniform(-1, 1, E)[:, None, None] + 0.05 * rng.standard_normal((E, n, n))
    u = euler_state(rho, v1, v2, p)
    dt = 1.5 * plain.max_dt(u)
    du_plain = plain.rhs(u, 0.0, dt)
    du_sc = sc.rhs(u, 0.0, dt)
    U = plain.aux("U").reshape(E, n, n, 4)
    F = plain.aux("F").reshape(E, n, n, 2, 4)
    _, w, _, _ = plain.basis()
    invJ = 2.0 / h

    def admissible(q):
        r = q[0]
        pr = 0.4 * (q[3] - 0.5 * (q[1] ** 2 + q[2] ** 2) / r)
        return r > 0 and pr > 0

    def lam(a, b, o):
        def parts(q):
            return abs(q[o] / q[0]), np.sqrt(1.4 * 0.4 * (q[3] - 0.5 * (q[1] ** 2 + q[2] ** 2) / q[0]) / q[0])

        al, cl = parts(a)
        ar, cr = parts(b)
        return max(al, ar) + max(cl, cr)

    expected = du_plain.copy()
    flagged = total = 0
    for L in range(E):
        ix, iy = L % nx, L // nx
        for o in (1, 2):
            R = ((ix + 1) % nx + nx * iy) if o == 1 else (ix + nx * ((iy + 1) % nx))
            for l in range(n):
                # [e, j, i] indices of the face node and its inner neighbour
                if o == 1:
                    iL, iLi, iR, iRi = (l, n - 1), (l, n - 2), (l, 0), (l, 1)
                else:
                    iL, iLi, iR, iRi = (n - 1, l), (n - 2, l), (0, l), (1, l)
                uL, uR = u[(L,) + iL], u[(R,) + iR]
                Fstar = 0.5 * (F[(L,) + iL + (o - 1,)] + F[(R,) + iR + (o - 1,)]) - 0.5 * lam(
                    U[(L,) + iL], U[(R,) + iR], o) * (uR - uL)
                fn = rusanov(uL, uR, o)
                f_ll = rusanov(u[(L,) + iLi], uL, o)
                f_rr = rusanov(uR, u[(R,) + iRi], o)
                low_l = uL - dt * invJ / (w[n - 1] * 0.5) * (Fstar - f_ll)
                low_r = uR - dt * invJ / (w[0] * 0.5) * (f_rr - Fstar)
                total += 1
                if not (admissible(low_l) and admissible(low_r)):
                    flagged += 1
                    expected[(L,) + iL] += -invJ * (fn - Fstar) / w[n - 1]
                    expected[(R,) + iR] += invJ * (fn - Fstar) / w[0]
    assert 0 < flagged < total, (flagged, total)
    assert rel_linf(du_sc, expected) < 1e-13
