"""Pin the CPU oracle before trusting it: reference known answers, the reference-generated
Brusselator code, closed forms and invariants.  Each check is chosen so that a plausible
mistake (dropped term, wrong sign or index, transposed operand) fails it."""
import json
import os

import numpy as np
import pytest

import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- Fornberg known answers
def test_fornberg_known_answers():
    # test/Components/MOLfornberg_weights.jl:5-30 (exact equality there)
    assert np.array_equal(orc.calculate_weights(2, 0.0, [-1, 0, 1.0]), [1, -2, 1])
    assert np.array_equal(orc.calculate_weights(1, 0.0, [-1.0, 1.0]), [-0.5, 0.5])
    assert np.array_equal(orc.calculate_weights(1, 0.0, [0, 1]), [-1, 1])
    assert np.array_equal(orc.calculate_weights(1, 1.0, [0, 1]), [-1, 1])
    assert np.array_equal(orc.calculate_weights(3, 0.0, [0, 1, 2, 3, 4, 5]),
                          [-17 / 4, 71 / 4, -59 / 2, 49 / 2, -41 / 4, 7 / 4])


@pytest.mark.parametrize("p,expected", [
    (2, ([-0.5, 0, 0.5], [1.0, -2.0, 1.0], [-1 / 2, 1.0, 0.0, -1.0, 1 / 2])),
    (4, ([1 / 12, -2 / 3, 0, 2 / 3, -1 / 12], [-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12],
         [1 / 8, -1.0, 13 / 8, 0.0, -13 / 8, 1.0, -1 / 8])),
])
def test_centered_tables(p, expected):
    # test/shared/finite_diff_schemes.jl:22-30 (dx = 1)
    for d in (1, 2, 3):
        st, low, high = orc.centered_weights(d, p, 1.0)
        np.testing.assert_allclose(st, expected[d - 1], rtol=1e-12, atol=1e-14)
        # boundary rows are derivatives at nodes 0.. of one-sided stencils: exact on polynomials
        for r in range(low.shape[0]):
            xs = np.arange(low.shape[1], dtype=float)
            for deg in range(d, d + p):
                poly = xs ** deg
                exact = np.prod(range(deg - d + 1, deg + 1)) * r ** (deg - d) if deg >= d else 0.0
                assert abs(low[r] @ poly - exact) <= 1e-9 * max(1.0, abs(exact)) * (xs.max() ** deg)
            # high rows mirror the low rows with sign (-1)^d (centered_diff_weights.jl:49-55)
            np.testing.assert_array_equal(high[r], low[r][::-1] * (-1.0) ** d)


def test_neumann_boundary_row():
    st, low, high = orc.centered_weights(1, 2, 0.5)
    np.testing.assert_array_equal(low[0], np.array([-1.5, 2.0, -0.5]) / 0.5)


def test_half_offset_tables():
    # interpolation to the half point with 4 nodes: [-1, 9, 9, -1]/16 and boundary rows
    st, low, high = orc.half_weights(0, 4, 0.1)
    np.testing.assert_allclose(st, np.array([-1, 9, 9, -1]) / 16, rtol=0, atol=1e-16)
    np.testing.assert_allclose(low[0], np.array([5, 15, -5, 1]) / 16, rtol=0, atol=1e-16)
    np.testing.assert_allclose(high[0], np.array([1, -5, 15, 5]) / 16, rtol=0, atol=1e-16)
    st, low, high = orc.half_weights(1, 2, 0.25)
    np.testing.assert_allclose(st, np.array([-1.0, 1.0]) / 0.25, rtol=1e-15)


# ---------------------------------------------------------------- reference-generated Brusselator
def test_brusselator_matches_generated_code():
    fx = json.load(open(os.path.join(GOLD, "bruss_dx025.json")))
    for s in fx["samples"]:
        du = orc.bruss2d_rhs(5, np.array(s["u"]), t=0.0)
        np.testing.assert_allclose(du, s["du"], rtol=1e-14, atol=1e-12)


def test_brusselator_structure_matches_generated_code():
    fx = json.load(open(os.path.join(GOLD, "bruss_dx025.json")))
    z = np.zeros(32)
    d0 = orc.bruss2d_rhs(5, z, 0.0)
    for k in range(32):
        assert d0[k] == fx["structure"][k]["const"]
    for j in range(32):
        e = np.zeros(32)
        e[j] = 1.0
        col = orc.bruss2d_rhs(5, e, 0.0) - d0
        for k in range(32):
            lin = dict((i, c) for i, c in fx["structure"][k]["linear"])
            assert abs(col[k] - lin.get(j, 0.0)) <= 1e-12 * 650, (k, j, col[k], lin.get(j))


def test_brusselator_laplacian_is_kronecker_sum():
    import scipy.sparse as sps
    n = 17
    m = n - 1
    h = 1.0 / (n - 1)
    one = sps.diags([np.ones(m - 1), -2 * np.ones(m), np.ones(m - 1)], [-1, 0, 1]).tolil()
    one[0, m - 1] = 1
    one[m - 1, 0] = 1
    L1 = one.tocsr() / h ** 2
    I = sps.identity(m)
    lap = sps.kron(I, L1) + sps.kron(L1, I)
    rng = np.random.default_rng(3)
    u = rng.uniform(0.5, 1.5, m * m)
    v = np.zeros(m * m)
    du = orc.bruss2d_rhs(n, np.concatenate([u, v]), 0.0)
    expected_u = 1.0 - 4.4 * u + 10.0 * (lap @ u)
    expected_v = 3.4 * u + 10.0 * (lap @ v)
    np.testing.assert_allclose(du[:m * m], expected_u, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(du[m * m:], expected_v, rtol=1e-12, atol=1e-9)


def test_brusselator_forcing_switches_on():
    n = 11
    m = n - 1
    u = np.ones(2 * m * m)
    a = orc.bruss2d_rhs(n, u, 1.0)
    b = orc.bruss2d_rhs(n, u, 1.2)
    diff = (b - a).reshape(2, m, m)
    xs = np.arange(1, n) / (n - 1)   # the reference grid 0:0.1:1 holds correctly rounded k/10
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    mask = ((X - 0.3) ** 2 + (Y - 0.6) ** 2) <= 0.1 ** 2
    np.testing.assert_array_equal(diff[0].T, np.where(mask, 5.0, 0.0))  # diff[f][y][x]
    assert np.all(diff[1] == 0)


# ---------------------------------------------------------------- heat: closed forms
def test_heat_rk4_closed_form():
    n, D, dt, nsteps = 101, 1.0, 1e-5, 10_000
    x = np.linspace(0, 1, n)
    dx = 0.01
    u0 = np.sin(np.pi * x[1:-1])
    u = orc.diffusion1d_integrate(n, D, [1, 0, 0], [1, 0, 0], u0, 0.0, dt, nsteps)
    lam = -(4 * D / dx ** 2) * np.sin(np.pi * dx / 2) ** 2
    z = dt * lam
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    exact_semi = R ** nsteps * u0
    assert np.max(np.abs(u - exact_semi)) < 1e-12
    analytic = np.exp(-np.pi ** 2 * D * 0.1) * u0
    assert 1e-6 < np.max(np.abs(u - analytic)) < 5e-5


@pytest.mark.parametrize("kind", ["neumann", "robin"])
def test_neumann_robin_ghost_closed_form(kind):
    n, D, dx = 21, 0.7, 0.05
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, n - 2)
    if kind == "neumann":
        bcl, bch = [0.0, 1.0, 0.0], [0.0, 1.0, 0.0]
        u1 = (4 * u[0] - u[1]) / 3
        un = (4 * u[-1] - u[-2]) / 3
    else:
        a0, a1, g = 2.0, 0.5, 0.3
        bcl, bch = [a0, a1, g], [a0, a1, g]
        # a0 u1 + a1 (-3/2 u1 + 2 u2 - 1/2 u3)/dx = g
        u1 = (g - a1 * (2 * u[0] - 0.5 * u[1]) / dx) / (a0 - 1.5 * a1 / dx)
        # upper row mirrored with sign: (1/2 u_{n-2} - 2 u_{n-1} + 3/2 u_n)/dx
        un = (g - a1 * (0.5 * u[-2] - 2 * u[-1]) / dx) / (a0 + 1.5 * a1 / dx)
    full = np.concatenate([[u1], u, [un]])
    expected = D * (full[:-2] - 2 * full[1:-1] + full[2:]) / dx ** 2
    du = orc.diffusion1d_rhs(n, D, bcl, bch, u)
    np.testing.assert_allclose(du, expected, rtol=1e-11, atol=1e-9)


# ---------------------------------------------------------------- WENO invariants
def test_weno_constant_linear_quadratic():
    rng = np.random.default_rng(11)
    for _ in range(20):
        a, b, c = rng.uniform(-2, 2, 3)
        dx = rng.uniform(0.01, 0.5)
        i = np.arange(-2, 3)
        assert abs(orc.weno_f(np.full(5, a), 1e-6, dx)) <= 1e-14 * abs(a) / dx
        lin = a + b * i * dx
        assert abs(orc.weno_f(lin, 1e-6, dx) - b) < 1e-11 * max(1, abs(b)) / dx * 10
        quad = a + b * i * dx + c * (i * dx) ** 2
        # every candidate reconstruction is exact on quadratics: the node derivative is exact
        assert abs(orc.weno_f(quad, 1e-6, dx) - b) < 1e-9 / dx


def test_weno_not_exact_on_cubics():
    dx = 0.1
    i = np.arange(-2, 3) * dx
    assert abs(orc.weno_f(i ** 3 + 1.0, 1e-6, dx) - 0.0) > 1e-6


# ---------------------------------------------------------------- nonlinear Laplacian / spherical / mixed / upwind
def _interp_half(u):
    return (-u[:-3] + 9 * u[1:-2] + 9 * u[2:-1] - u[3:]) / 16


@pytest.mark.parametrize("periodic", [True, False])
def test_nonlinlap_second_order_closed_form(periodic):
    n = 33
    dx = 1.0 / (n - 1)
    rng = np.random.default_rng(7)
    if periodic:
        u = rng.uniform(-1, 1, n - 1)        # nodes 2..n
        ring = u
        ext = np.concatenate([ring[-2:], ring, ring[:2]])
        half = _interp_half(ext)              # half points between nodes k and k+1
        a = 1 + half ** 2
        d = (ext[2:-1] - ext[1:-2])
        flux = a * d / dx
        du = (flux[1:] - flux[:-1]) / dx
        got = orc.nonlinlap1d_rhs(n, u, True)
        np.testing.assert_allclose(got, du, rtol=1e-11, atol=1e-9)
    else:
        gl, gh = 0.2, -0.4
        u = rng.uniform(-1, 1, n - 2)
        full = np.concatenate([[gl], u, [gh]])
        hp = np.empty(n - 1)
        hp[0] = (5 * full[0] + 15 * full[1] - 5 * full[2] + full[3]) / 16
        hp[-1] = (full[-4] - 5 * full[-3] + 15 * full[-2] + 5 * full[-1]) / 16
        hp[1:-1] = _interp_half(full)
        a = 1 + hp ** 2
        flux = a * (full[1:] - full[:-1]) / dx
        du = (flux[1:] - flux[:-1]) / dx
        got = orc.nonlinlap1d_rhs(n, u, False, gl, gh)
        np.testing.assert_allclose(got, du, rtol=1e-11, atol=1e-9)


def test_spherical_interior_closed_form():
    n = 21
    dx = 0.05
    r = np.linspace(0, 1, n)
    rng = np.random.default_rng(9)
    u = rng.uniform(0, 1, n - 2)
    g1 = 0.25
    u1 = (4 * u[0] - u[1]) / 3
    full = np.concatenate([[u1], u, [g1]])
    rr = r[1:-1]
    expected = (full[2:] - full[:-2]) / (2 * dx) / rr + (full[2:] - 2 * full[1:-1] + full[:-2]) / dx ** 2
    got = orc.spherical1d_rhs(n, g1, u)
    np.testing.assert_allclose(got, expected, rtol=1e-11, atol=1e-8)


def test_mixed_derivative_brute_force():
    n = 9
    h = 1.0 / (n - 1)
    x = np.linspace(0, 1, n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    full = X + 2 * Y
    rng = np.random.default_rng(13)
    inner = rng.uniform(-1, 1, (n - 2, n - 2))
    full[1:-1, 1:-1] = inner
    for c in [(0, 0), (0, -1), (-1, 0), (-1, -1)]:
        full[c] = 0.0   # corner nodes ~ 0 (generate_bc_eqs.jl:415)
    lap = (full[2:, 1:-1] - 2 * full[1:-1, 1:-1] + full[:-2, 1:-1]) / h ** 2 + \
          (full[1:-1, 2:] - 2 * full[1:-1, 1:-1] + full[1:-1, :-2]) / h ** 2
    mix = (full[2:, 2:] - full[2:, :-2] - full[:-2, 2:] + full[:-2, :-2]) / (4 * h * h)
    expected = lap + 0.7 * mix
    got = orc.mixed2d_rhs(n, 0.7, inner.reshape(-1, order="F"))
    np.testing.assert_allclose(got, expected.reshape(-1, order="F"), rtol=1e-11, atol=1e-8)


def test_upwind_sign_selection_closed_form():
    n = 41
    dx = 1.0 / (n - 1)
    x = np.linspace(0, 1, n)
    rng = np.random.default_rng(17)
    u = rng.uniform(-1, 1, n - 2)
    gl, gh = 0.3, -0.2
    full = np.concatenate([[gl], u, [gh]])
    c = 0.1 + np.sin(2 * np.pi * x[1:-1])
    back = (full[1:-1] - full[:-2]) / dx
    fwd = (full[2:] - full[1:-1]) / dx
    expected = -np.where(c > 0, c * back, c * fwd)
    got = orc.upwind1d_rhs(n, 1, 0.1, 1.0, gl, gh, u)
    np.testing.assert_allclose(got, expected, rtol=1e-11, atol=1e-9)


def test_weno_dirichlet_clipped_interior_shape():
    # WENO on a non-periodic axis clips the unknowns by its extent 2 (WENO.jl:94, interior_map.jl:117-139)
    n = 41
    u = np.linspace(0.1, 0.9, n - 4)
    du = orc.weno1d_dirichlet_rhs(n, 1.0, 1.0, 1e-6, 0.1, 0.9, u)
    assert du.shape == (n - 4,)
    assert np.all(np.isfinite(du))


def test_nonlindiff3d_constant_state_gives_source():
    n = 10
    G = n ** 3
    d0 = np.ones(G)
    rng = np.random.default_rng(2)
    src = rng.uniform(-1, 1, G)
    u = np.full((n - 2) ** 3, 0.37)
    du = orc.nonlindiff3d_rhs(n, d0, src, u)
    s_int = src.reshape(n, n, n, order="F")[1:-1, 1:-1, 1:-1].reshape(-1, order="F")
    np.testing.assert_allclose(du, s_int, rtol=0, atol=1e-11)


def test_nonlindiff3d_reduces_to_laplacian_for_constant_coefficient():
    # a = d0 (1 + u^2) with u tiny and d0 = 1: equals the Neumann 7-point Laplacian to O(u^3)
    n = 10
    h = 1.0 / (n - 1)
    G = n ** 3
    rng = np.random.default_rng(4)
    ui = 1e-4 * rng.uniform(-1, 1, (n - 2,) * 3)
    du = orc.nonlindiff3d_rhs(n, np.ones(G), np.zeros(G), ui.reshape(-1, order="F"))
    full = np.zeros((n, n, n))
    full[1:-1, 1:-1, 1:-1] = ui
    for ax in range(3):
        sl = [slice(1, -1)] * 3
        for side, (g, a1, a2) in enumerate([(0, 1, 2), (-1, -2, -3)]):
            s_g, s_1, s_2 = list(sl), list(sl), list(sl)
            s_g[ax], s_1[ax], s_2[ax] = g, a1, a2
            full[tuple(s_g)] = (4 * full[tuple(s_1)] - full[tuple(s_2)]) / 3
    lap = np.zeros_like(ui)
    for ax in range(3):
        lap += (np.roll(full, -1, ax) - 2 * full + np.roll(full, 1, ax))[1:-1, 1:-1, 1:-1] / h ** 2
    np.testing.assert_allclose(du.reshape((n - 2,) * 3, order="F"), lap, rtol=1e-6, atol=1e-9)


def test_oracle_rejects_a_state_of_the_wrong_length():
    """The oracle's right-hand sides read as many values as the system has unknowns; a state of
    another length is refused instead of read past its end."""
    import numpy as np
    import oracle as orc
    import pytest
    n = 21
    orc.spherical1d_rhs(n, 0.4, np.zeros(n - 2))
    for bad in (n - 1, n - 3):
        with pytest.raises(RuntimeError):
            orc.spherical1d_rhs(n, 0.4, np.zeros(bad))
    with pytest.raises(RuntimeError):
        orc.diffusion1d_rhs(n, 1.0, (1, 0, 0), (1, 0, 0), np.zeros(n))


# ---------------------------------------------------------------- WENO5 weights and smoothness indicators
def test_weno_linear_limit_is_the_fourth_order_central_difference():
    """With eps far above every smoothness indicator the nonlinear weights equal the ideal weights,
    and (u-_{i+1/2} - u+_{i-1/2})/dx of the two fifth-order reconstructions collapses to the
    five-point central difference (u_{-2} - 8u_{-1} + 8u_1 - u_2)/(12 dx) for ANY data.  Swapping the
    two ideal-weight sets of WENO.jl:14-43 (or any other choice) breaks this identity."""
    rng = np.random.default_rng(21)
    for _ in range(200):
        u = rng.uniform(-3, 3, 5)
        dx = rng.uniform(1e-3, 1.0)
        exact = (u[0] - 8 * u[1] + 8 * u[3] - u[4]) / (12 * dx)
        got = orc.weno_f(u, 1e30, dx)
        assert abs(got - exact) <= 1e-12 * np.abs(u).max() / dx, (u, dx, got, exact)


def _jiang_shu_from_definitions(u, eps, dx):
    """Jiang & Shu (J. Comput. Phys. 126, 1996), built from its definitions rather than from
    its closed-form constants: the five values are cell averages over cells of width dx centred
    at j*dx (j = -2..2); each candidate quadratic p_k matches the averages of three adjacent cells;
    the smoothness indicator is beta_k = sum_{l=1,2} dx^(2l-1) * integral over cell 0 of
    (d^l p_k / dx^l)^2; the ideal weights are the ones whose combination of the candidates equals
    the quartic that matches all five averages, at x = +dx/2 and at x = -dx/2 separately.
    Returns (u^-_{+1/2} - u^+_{-1/2}) / dx with weights gamma_k / (eps + beta_k)^2, normalised."""
    from numpy.polynomial import polynomial as P

    def avg_row(j, deg):
        # cell average of x^m over [(j - 1/2)dx, (j + 1/2)dx], m = 0..deg
        a, b = (j - 0.5) * dx, (j + 0.5) * dx
        return np.array([(b ** (m + 1) - a ** (m + 1)) / ((m + 1) * dx) for m in range(deg + 1)])

    def fit(cells, vals):
        A = np.array([avg_row(j, len(cells) - 1) for j in cells])
        return np.linalg.solve(A, vals)            # monomial coefficients

    stencils = [(-2, -1, 0), (-1, 0, 1), (0, 1, 2)]
    polys = [fit(s, u[[j + 2 for j in s]]) for s in stencils]
    beta = []
    for c in polys:
        tot = 0.0
        for l in (1, 2):
            d = P.polyder(c, l)
            sq = P.polyint(P.polymul(d, d))
            tot += dx ** (2 * l - 1) * (P.polyval(dx / 2, sq) - P.polyval(-dx / 2, sq))
        beta.append(tot)
    beta = np.array(beta)

    def ideal(xe):
        # coefficients of u_{-2..2} in each candidate's value at xe, and in the quartic's value
        def row(cells):
            A = np.array([avg_row(j, len(cells) - 1) for j in cells])
            mon = np.array([xe ** m for m in range(len(cells))])
            coef = np.zeros(5)
            coef[[j + 2 for j in cells]] = np.linalg.solve(A.T, mon)
            return coef
        M = np.array([row(s) for s in stencils]).T   # 5 x 3
        g, *_ = np.linalg.lstsq(M, row((-2, -1, 0, 1, 2)), rcond=None)
        return g

    def recon(xe):
        g = ideal(xe)
        w = g / (eps + beta) ** 2
        w = w / w.sum()
        return sum(w[k] * P.polyval(xe, polys[k]) for k in range(3))

    return (recon(dx / 2) - recon(-dx / 2)) / dx


def test_weno_ideal_weights_from_the_definition():
    """The derived ideal weights are the published (1/10, 6/10, 3/10) at the right face and their
    mirror at the left face (a check of the independent build itself)."""
    from numpy.polynomial import polynomial as P  # noqa: F401
    rng = np.random.default_rng(0)
    u = rng.uniform(-1, 1, 5)
    # eps -> infinity: the independent build must reduce to the five-point central difference too
    got = _jiang_shu_from_definitions(u, 1e30, 0.25)
    exact = (u[0] - 8 * u[1] + 8 * u[3] - u[4]) / (12 * 0.25)
    assert abs(got - exact) < 1e-12 * np.abs(u).max() / 0.25


@pytest.mark.parametrize("eps", [1e-6, 1e-2, 1.0])
def test_weno_matches_an_independent_jiang_shu_build(eps):
    """The oracle's WENO5 (WENO.jl:6-57) against the independent build above on non-smooth random
    data, where the nonlinear weights depend on every coefficient of the smoothness indicators
    (the 13/12 and 1/4 of WENO.jl:19-21) and on which ideal-weight set goes with which face."""
    rng = np.random.default_rng(31)
    for trial in range(300):
        dx = float(rng.choice([1.0, 0.1, 1 / 64]))
        kind = trial % 3
        if kind == 0:
            u = rng.uniform(-1, 1, 5)                           # rough
        elif kind == 1:
            u = np.where(np.arange(5) >= rng.integers(1, 5), 1.0, 0.0) + 0.01 * rng.standard_normal(5)  # a jump
        else:
            x = (np.arange(5) - 2) * dx
            u = np.sin(3 * x + rng.uniform(0, 6)) + rng.uniform(-1, 1) * x ** 3   # smooth
        ref = _jiang_shu_from_definitions(u, eps, dx)
        got = orc.weno_f(u, eps, dx)
        assert abs(got - ref) <= 1e-11 * np.abs(u).max() / dx, (trial, u, dx, got, ref)


def test_weno_coefficient_mistakes_are_caught():
    """The independent build separates the published indicator from plausible misprints: with
    13/12 -> 1 or 1/4 -> 1/2 the nonlinear result moves far beyond the tolerance above."""
    u = np.array([0.0, 0.1, 1.0, 1.05, 0.2])
    dx, eps = 0.1, 1e-6
    good = _jiang_shu_from_definitions(u, eps, dx)

    def variant(c1, c2):
        b1 = c1 * (u[2] - 2 * u[3] + u[4]) ** 2 + c2 * (3 * u[2] - 4 * u[3] + u[4]) ** 2
        b2 = c1 * (u[1] - 2 * u[2] + u[3]) ** 2 + c2 * (u[1] - u[3]) ** 2
        b3 = c1 * (u[0] - 2 * u[1] + u[2]) ** 2 + c2 * (u[0] - 4 * u[1] + 3 * u[2]) ** 2
        b = np.array([b3, b2, b1])                       # stencils left, centre, right
        hp = np.array([(2 * u[0] - 7 * u[1] + 11 * u[2]) / 6, (-u[1] + 5 * u[2] + 2 * u[3]) / 6,
                       (2 * u[2] + 5 * u[3] - u[4]) / 6])
        hm = np.array([(-u[0] + 5 * u[1] + 2 * u[2]) / 6, (2 * u[1] + 5 * u[2] - u[3]) / 6,
                       (11 * u[2] - 7 * u[3] + 2 * u[4]) / 6])
        wp = np.array([0.1, 0.6, 0.3]) / (eps + b) ** 2
        wm = np.array([0.3, 0.6, 0.1]) / (eps + b) ** 2
        return ((wp / wp.sum()) @ hp - (wm / wm.sum()) @ hm) / dx

    assert abs(variant(13 / 12, 1 / 4) - good) < 1e-11 / dx
    assert abs(variant(1.0, 1 / 4) - good) > 1e-6
    assert abs(variant(13 / 12, 1 / 2) - good) > 1e-6


# ---------------------------------------------------------------- SSPRK33: closed form on eigenvectors
@pytest.mark.parametrize("mode,z", [(19, -2.3), (10, -1.2), (5, -0.35)])
def test_ssprk33_closed_form(mode, z):
    """Shu-Osher SSPRK(3,3) on a linear system advances an eigenvector by the stability polynomial
    R(z) = 1 + z + z^2/2 + z^3/6 per step.  Dirichlet-0 heat on n nodes has eigenvectors
    sin(k pi x_j) with eigenvalue -(4D/dx^2) sin^2(k pi dx/2); dt is chosen so that dt * lambda = z,
    where a dropped or misplaced stage coefficient changes R by O(z^2) per step.  Every other mode
    has |dt * lambda| below 2.5, inside both schemes' stability intervals, so rounding in the
    other modes stays bounded."""
    n, D = 21, 1.0
    dx = 1.0 / (n - 1)
    x = np.linspace(0, 1, n)
    lam = -(4 * D / dx ** 2) * np.sin(mode * np.pi * dx / 2) ** 2
    dt = z / lam
    nsteps = 40
    u0 = np.sin(mode * np.pi * x[1:-1])
    got = orc.diffusion1d_integrate(n, D, [1, 0, 0], [1, 0, 0], u0, 0.0, dt, nsteps, scheme=1)
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6
    np.testing.assert_allclose(got, R ** nsteps * u0, rtol=0, atol=1e-12 * max(abs(R) ** nsteps, 1e-3))
    # and RK4 on the same data is the fourth-order polynomial: the two schemes differ by z^4/24
    got4 = orc.diffusion1d_integrate(n, D, [1, 0, 0], [1, 0, 0], u0, 0.0, dt, nsteps, scheme=0)
    R4 = R + z ** 4 / 24
    np.testing.assert_allclose(got4, R4 ** nsteps * u0, rtol=0, atol=1e-12 * max(abs(R4) ** nsteps, 1e-3))


# ---------------------------------------------------------------- nonuniform first-order upwind rows
@pytest.mark.parametrize("c", [1.3, -0.8])
def test_nonuniform_first_order_upwind_rows_brute_force(c):
    """CompleteUpwindDifference on nodes (upwind_diff_weights.jl:116-127) stores the backward
    operator's rows with offside 0, and _upwind_difference (upwind_difference.jl:154-158) reads
    row II for node II.  At first order that row was computed at x[II+1] on x[II..II+1] and is
    applied to u[II-1], u[II]: du = -c (u[II] - u[II-1]) / (x[II+1] - x[II]) for c > 0; the
    forward operator (c <= 0) gives -c (u[II+1] - u[II]) / (x[II+1] - x[II]).  Written out node by
    node on a random 7-node grid with Dirichlet ends."""
    rng = np.random.default_rng(41)
    x = np.sort(np.concatenate([[0.0, 1.0], rng.uniform(0.05, 0.95, 5)]))
    n = len(x)
    gl, gh = 0.35, -0.6
    u = rng.uniform(-1, 1, n - 2)
    full = np.concatenate([[gl], u, [gh]])
    expected = np.empty(n - 2)
    for i in range(1, n - 1):          # 0-based node i is the reference's II = i + 1
        h = x[i + 1] - x[i]
        if c > 0:
            expected[i - 1] = -c * (full[i] - full[i - 1]) / h
        else:
            expected[i - 1] = -c * (full[i + 1] - full[i]) / h
    got = orc.nonuniform1d_rhs(x, 0.0, c, 2, 1, (1, 0, gl), (1, 0, gh), u)
    np.testing.assert_allclose(got, expected, rtol=1e-13, atol=1e-13)
