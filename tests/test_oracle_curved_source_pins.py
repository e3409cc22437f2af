"""Pins of the oracle's source terms on the curvilinear StructuredMesh
(lw_volume_kernel_3!/4! of src/solvers/dgsem_structured/dg_2d.jl:1379-2105 with
their source lines; the reference's elixir_euler_source_terms_parallelogram.jl
and _nonperiodic.jl; no GPU).

- On a uniform state the warped mesh keeps the free stream, so with a source
  of t alone the residual is the source part S itself: kernel 4 (five-point
  s_tt, calc_source_tt_N4, as the structured kernel calls it) gives the exact
  average of s over [t, t + dt] for quartics; kernel 3 for cubics, a quartic
  then exceeding it by dt⁴ s⁽⁴⁾/180 (three-point s_tt, no s_tttt level).
- Trixi's convergence test on the parallelogram x = ξ + η, y = η (periodic)
  converges at order ≈ N + 1 in the density's L2 error at t = 0.5.
- The curvilinear MDRK stages (structured/multi_derivative_2d.jl:146-200,
  326-385): on the same uniform state stage 1 gives s/2 + dt s'/8 and the step
  s + dt s'/3 for a linear source, because the reference's stage 2 loads stage
  1's S2 = s + s_t/6 into S and then overwrites S with s (the exact average
  would be s + dt s'/2; the TreeMesh stages give it).
"""
import numpy as np
import pytest

W4 = np.array([1.0, 0.5, -0.3, 2.0])


def exact(x, y, t):
    phi = 2.0 + 0.1 * np.sin(np.pi * (x + y - t))
    return np.stack([phi, phi, phi, phi * phi], -1)


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("p4est", [False, True])
def test_source_weights_on_a_uniform_state_of_a_warped_mesh(oracle_mod, N, p4est):
    """(also with P4est element frames turned by random multiples of 90°)"""
    t, dt = 0.3, 0.01
    turns = np.random.default_rng(4).integers(0, 4, 36) if p4est else None
    for deg in (1, 2, 3, 4):
        P = np.polynomial.Polynomial([0.3, -0.7, 0.5, 0.9, -0.4][:deg + 1])
        o = oracle_mod.Oracle(N, 6, 6, 0.0, 0.0, 1.0, gamma=1.4, mapping=(oracle_mod.MAP_WARP, (2.0, 0.08)),
                              periodic=(True, True), source_terms=lambda u, x, y, tt, P=P: P(tt) * W4, p4est=turns)
        x, _ = o.node_coords()
        u = np.empty(x.shape + (4,))
        u[:] = [1.0, 0.1, 0.2, 2.6]
        I = P.integ()
        avg = (I(t + dt) - I(t)) / dt
        excess = dt ** 4 * P.deriv(4)(t) / 180.0 if N == 3 else 0.0
        np.testing.assert_allclose(o.rhs(u, t, dt), np.broadcast_to((avg + excess) * W4, u.shape), rtol=0,
                                   atol=1e-11)


@pytest.mark.parametrize("N,orders", [(3, (3.0, 4.3)), (4, (4.4, 4.0))])
def test_parallelogram_converges(oracle_mod, N, orders):
    errs = []
    for nx in (4, 8, 16):
        o = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, 1.0 / nx, gamma=1.4,
                              mapping=(oracle_mod.MAP_PARALLELOGRAM, (0.0, 0.0, 0.0, 0.0)), periodic=(True, True),
                              source_terms="convergence_test")
        x, y = o.node_coords()
        u1, t, _ = o.solve(exact(x, y, 0.0), 0.0, 0.5, 0.2)
        assert t == 0.5
        W = np.outer(*[o.basis()[1]] * 2)[None] / o.aux("inverse_jacobian").reshape(x.shape)
        errs.append(np.sqrt(np.einsum("eji,eji->", W, (u1[..., 0] - exact(x, y, t)[..., 0]) ** 2)))
    eoc = [np.log2(errs[k] / errs[k + 1]) for k in range(2)]
    assert eoc[0] > orders[0] and eoc[1] > orders[1], (errs, eoc)


@pytest.mark.parametrize("N", [3, 4])
def test_curved_mdrk_source_weights_on_a_uniform_state(oracle_mod, N):
    t, dt = 0.3, 0.01
    for a in ([0.3, -0.7], [0.3, -0.7, 0.5]):
        P = np.polynomial.Polynomial(a)
        o = oracle_mod.Oracle(N, 6, 6, 0.0, 0.0, 1.0, gamma=1.4, mapping=(oracle_mod.MAP_WARP, (2.0, 0.08)),
                              periodic=(True, True), source_terms=lambda u, x, y, tt, P=P: P(tt) * W4)
        x, _ = o.node_coords()
        u = np.empty(x.shape + (4,))
        u[:] = [1.0, 0.1, 0.2, 2.6]
        d1 = o.rhs(u, t, dt, kind="mdrk1")
        # s_t of N34 is dt s' for a polynomial of degree <= 4
        np.testing.assert_allclose(d1, np.broadcast_to((P(t) / 2 + dt * P.deriv()(t) / 8) * W4, u.shape), rtol=0,
                                   atol=1e-11)
        o.set_us(u + dt * d1)
        d2 = o.rhs(u, t, dt, kind="mdrk2")
        np.testing.assert_allclose(d2, np.broadcast_to((P(t) + dt * P.deriv()(t) / 3) * W4, u.shape), rtol=0,
                                   atol=1e-11)
