"""Centered and upwind operators on nonuniform grids (CompleteCenteredDifference /
CompleteUpwindDifference on nodes: src/discretization/schemes/centered_difference/
centered_diff_weights.jl:77-153, schemes/upwind_difference/upwind_diff_weights.jl:93-171,
row selection centered_difference.jl:32-57 and upwind_difference.jl:131-162).

The oracle's nonuniform path is pinned by: equality with the uniform path (itself pinned to the
reference's tables and fixtures) when the nodes are uniform; exactness of the discrete
operator and of the Neumann / Robin closures for polynomials the stencils reproduce; the
order of convergence on a smoothly stretched grid.  The generated kernels (compiled for the
CPU) are checked against the oracle, and their reverse mode against finite differences.
"""
import os
import sys

import numpy as np
import pytest

import oracle as orc
from molgpu import systems
from molgpu.discretization import lower
from molgpu import fornberg

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "genhost"))
import genhost  # noqa: E402

BCS = [((1, 0, 0.3), (1, 0, -0.2)), ((0, 1, 0.5), (1, 0, 0.0)), ((1, 2, 0.1), (2, -1, 0.4))]


@pytest.mark.parametrize("p", [2, 4, 6])
@pytest.mark.parametrize("bcl,bch", BCS)
def test_uniform_nodes_reproduce_the_uniform_operator(p, bcl, bch):
    n = 41
    x = np.linspace(0.0, 1.0, n)
    u = np.random.default_rng(p).uniform(-1, 1, n - 2)
    a = orc.nonuniform1d_rhs(x, 0.7, 0.0, p, 1, bcl, bch, u)
    b = orc.diffusion1d_rhs(n, 0.7, bcl, bch, u, p=p)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))


@pytest.mark.parametrize("c", [0.9, -0.9])
def test_uniform_nodes_reproduce_the_uniform_upwind_operator(c):
    n = 41
    x = np.linspace(0.0, 1.0, n)
    u = np.random.default_rng(3).uniform(-1, 1, n - 2)
    a = orc.nonuniform1d_rhs(x, 0.0, c, 2, 1, (1, 0, 0.3), (1, 0, -0.2), u)
    b = orc.upwind1d_rhs(n, 1, c, 0.0, 0.3, -0.2, u)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))


@pytest.mark.parametrize("p,deg", [(2, 2), (4, 4), (6, 6)])
@pytest.mark.parametrize("kind", ["dirichlet", "neumann", "robin"])
def test_second_derivative_and_closures_exact_for_polynomials(p, deg, kind):
    """Interior rows use p+1 nodes (exact to degree p), boundary rows and the closures' first-
    derivative rows p+2 and p+1 nodes: for a degree-p polynomial with boundary data taken from
    it, every unknown gets D u'' exactly on a random nonuniform grid."""
    rng = np.random.default_rng(p + len(kind))
    n = 31
    x = systems.stretched_grid(n, 1.3, seed=p, spread=0.6) + 0.4
    coef = rng.uniform(-1, 1, deg + 1)
    P = np.polynomial.Polynomial(coef)
    d1, d2 = P.deriv(), P.deriv(2)
    a, b = x[0], x[-1]
    if kind == "dirichlet":
        bcl, bch = (1, 0, P(a)), (1, 0, P(b))
    elif kind == "neumann":
        bcl, bch = (0, 1, d1(a)), (0, 1, d1(b))
    else:
        bcl, bch = (1, 0.5, P(a) + 0.5 * d1(a)), (2, -1, 2 * P(b) - d1(b))
    u = P(x[1:-1])
    du = orc.nonuniform1d_rhs(x, 0.7, 0.0, p, 1, bcl, bch, u)
    ref = 0.7 * d2(x[1:-1])
    assert np.max(np.abs(du - ref)) <= 1e-8 * (1 + np.max(np.abs(ref)))


@pytest.mark.parametrize("p", [2, 4])
def test_convergence_order_on_a_smoothly_stretched_grid(p):
    errs = []
    ns = [41, 81, 161]
    for n in ns:
        xi = np.linspace(0, 1, n)
        x = xi + 0.15 * np.sin(np.pi * xi) / np.pi
        f = lambda z: np.sin(3 * z) + z * z
        u = f(x[1:-1])
        du = orc.nonuniform1d_rhs(x, 1.0, 0.0, p, 1, (1, 0, f(x[0])), (1, 0, f(x[-1])), u)
        errs.append(np.max(np.abs(du - (-9 * np.sin(3 * x[1:-1]) + 2))))
    orders = [np.log2(errs[k] / errs[k + 1]) for k in range(len(ns) - 1)]
    assert orders[-1] > p - 0.3, (errs, orders)


def test_forward_upwind_exact_for_linear_data():
    """Negative wind: node i takes its own forward row (Fornberg at x[i] on x[i], x[i+1])."""
    n = 25
    x = systems.stretched_grid(n, 2.0, seed=9, spread=0.7)
    u = 0.4 + 1.7 * x[1:-1]
    du = orc.nonuniform1d_rhs(x, 0.0, -1.1, 2, 1, (1, 0, 0.4 + 1.7 * x[0]), (1, 0, 0.4 + 1.7 * x[-1]), u)
    assert np.max(np.abs(du - 1.1 * 1.7)) <= 1e-12


def test_operator_row_bookkeeping():
    """Row ranges of the per-node operators: centered rows for bpc+1..n-bpc, forward rows for
    1..n-hc with hc high rows, backward rows for L..n-L+1 (the reference builds no row past it)."""
    x = systems.stretched_grid(20, 1.0, seed=1)
    c = fornberg.centered_difference_nonuniform(2, 4, x)
    assert (c.pn_lo, c.pn_hi, c.nlow, c.nhigh, c.stencil_length, c.boundary_stencil_length) == (3, 18, 2, 2, 5, 6)
    assert c.stencil_coefs.shape == (16, 5)
    f = fornberg.upwind_difference_nonuniform(1, 1, x, 0)
    assert (f.pn_lo, f.pn_hi, f.nhigh, f.tap_offset) == (1, 19, 1, 0)
    b = fornberg.upwind_difference_nonuniform(1, 1, x, 1)
    assert (b.pn_lo, b.pn_hi, b.nlow, b.tap_offset) == (2, 19, 0, -1)
    # row sums vanish (first and second derivatives)
    assert np.max(np.abs(c.stencil_coefs.sum(axis=1))) <= 1e-9 * np.max(np.abs(c.stencil_coefs))


def test_lowering_rejects_what_the_reference_does_not_build():
    x = systems.stretched_grid(21, 1.0, seed=1)
    with pytest.raises(NotImplementedError):
        lower(*systems.nonuniform1d(x, 0.0, 1.0, 2, 2))          # upwind order 2
    with pytest.raises(NotImplementedError):
        lower(*systems.nonlinlap1d(x, True))                     # periodic nonuniform half-point operators


@pytest.mark.parametrize("p", [2, 4])
def test_uniform_nodes_reproduce_the_uniform_nonlinear_and_spherical_laplacians(p):
    for n in (21, 41):
        x = np.linspace(0.0, 1.0, n)
        u = np.random.default_rng(n + p).uniform(-1, 1, n - 2)
        a = orc.nonlinlap1d_nu_rhs(x, p, 0.2, -0.3, u)
        b = orc.nonlinlap1d_rhs(n, u, False, 0.2, -0.3, p=p)
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
        a = orc.spherical1d_nu_rhs(x, p, 0.4, u)
        b = orc.spherical1d_rhs(n, 0.4, u, p=p)
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))


def test_nonlinear_laplacian_closed_form_on_a_nonuniform_grid():
    """p = 2, a = 1 + u^2, u = al + be x: the inner two-point difference is exact (be), the
    4-point interpolation reproduces u at the half points, and the outer two-point difference
    over the midpoints differentiates the quadratic a(u) be exactly at the midpoint of the two
    midpoints, xb = (x[i-1] + 2 x[i] + x[i+1]) / 4, where the reference's outer operator, built
    on the midpoints (differential_discretizer.jl:44-51), evaluates it."""
    x = systems.stretched_grid(41, 1.0, seed=6, spread=0.6)
    al, be = 0.3, 0.7
    du = orc.nonlinlap1d_nu_rhs(x, 2, al, al + be, al + be * x[1:-1])
    xb = (x[:-2] + 2 * x[1:-1] + x[2:]) / 4
    ref = 2 * be ** 2 * (al + be * xb)
    assert np.max(np.abs(du - ref)) <= 1e-10
    # and not at the node itself
    assert np.max(np.abs(du - 2 * be ** 2 * (al + be * x[1:-1]))) > 1e-4


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("kind", ["nonlinlap", "spherical"])
def test_generated_nonlinear_and_spherical_laplacians_match_oracle(p, kind):
    x = systems.stretched_grid(41, 1.0, seed=6)
    if kind == "nonlinlap":
        low = lower(*systems.nonlinlap1d(x, False, 0.2, -0.3, p))
        ref_fn = lambda w: orc.nonlinlap1d_nu_rhs(x, p, 0.2, -0.3, w)
    else:
        low = lower(*systems.spherical1d(x, 0.4, p))
        ref_fn = lambda w: orc.spherical1d_nu_rhs(x, p, 0.4, w)
    assert low.n_unknowns == len(x) - 2
    u = np.random.default_rng(p).uniform(-1, 1, low.n_unknowns)
    ref = ref_fn(u)
    for tile in [(8, 1, 1), None]:
        got = genhost.rhs(low, u, tile=tile)
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    rng = np.random.default_rng(9)
    lam, v = rng.uniform(-1, 1, low.n_unknowns), rng.uniform(-1, 1, low.n_unknowns)
    g, _ = genhost.vjp(low, u, lam, tile=(8, 1, 1))
    F = lambda w: lam @ ref_fn(w)
    h = 1e-5
    d1 = (F(u + h * v) - F(u - h * v)) / (2 * h)
    d2 = (F(u + h / 2 * v) - F(u - h / 2 * v)) / h
    fd = (4 * d2 - d1) / 3
    assert abs(g @ v - fd) <= 1e-8 * (abs(fd) + 1)


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("c", [0.0, 0.8, -0.8])
@pytest.mark.parametrize("bcl,bch", BCS)
def test_generated_kernel_matches_oracle(p, c, bcl, bch):
    n = 41
    x = systems.stretched_grid(n, 1.5, seed=2)
    low = lower(*systems.nonuniform1d(x, 0.7, c, p, 1, bcl, bch))
    u = np.random.default_rng(p).uniform(-1, 1, low.n_unknowns)
    ref = orc.nonuniform1d_rhs(x, 0.7, c, p, 1, bcl, bch, u)
    for tile in [(8, 1, 1), None]:
        got = genhost.rhs(low, u, tile=tile)
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("c", [0.8, -0.8])
def test_generated_reverse_mode_matches_finite_differences(c):
    n = 41
    x = systems.stretched_grid(n, 1.5, seed=2)
    bcl, bch = (0, 1, 0.5), (1, 2, 0.1)
    low = lower(*systems.nonuniform1d(x, 0.7, c, 4, 1, bcl, bch))
    rng = np.random.default_rng(5)
    N = low.n_unknowns
    u, lam, v = rng.uniform(-1, 1, N), rng.uniform(-1, 1, N), rng.uniform(-1, 1, N)
    g, _ = genhost.vjp(low, u, lam, tile=(8, 1, 1))
    # the map is affine in u: the central difference is exact up to rounding
    F = lambda w: lam @ orc.nonuniform1d_rhs(x, 0.7, c, 4, 1, bcl, bch, w)
    fd = (F(u + v) - F(u - v)) / 2
    assert abs(g @ v - fd) <= 1e-10 * (abs(fd) + np.linalg.norm(g) * np.linalg.norm(v))


@pytest.mark.parametrize("bct", [(0, 0, 0, 0), (0, 1, 1, 0), (1, 1, 1, 1)])
def test_uniform_counts_reproduce_the_2d_oracle(bct):
    n = 24
    u = np.random.default_rng(1).uniform(0, 1, (n - 2) ** 2)
    a = orc.diffreact_nd_rhs([n, n], 2, 0.05, 2.5, bct, u)
    b = orc.diffreact2d_rhs(n, 0.05, 2.5, bct, u)
    assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b))


@pytest.mark.parametrize("which", ["x", "y", "xy", "xz3d"])
@pytest.mark.parametrize("p", [2, 4])
def test_generated_multidimensional_nonuniform_matches_oracle(which, p):
    """Per-node rows on one or more axes of a 2D / 3D grid, with the register-blocked axis among
    them (tiles 8x4, 16x8, 8x4x4 and the default), Dirichlet / Neumann sides."""
    xg = systems.stretched_grid(29 if "3d" not in which else 14, 1.0, seed=4)
    yg = systems.stretched_grid(23, 1.0, seed=5)
    zg = systems.stretched_grid(11, 1.0, seed=7)
    grids = {"x": [xg, 23], "y": [31, yg], "xy": [xg, yg], "xz3d": [xg, 12, zg]}[which]
    bct = (0, 1, 1, 0) + ((1, 0) if len(grids) == 3 else ())
    low = lower(*systems.diffreact_grids(grids, 0.05, 2.5, bct, p))
    u = np.random.default_rng(p).uniform(0, 1, low.n_unknowns)
    ref = orc.diffreact_nd_rhs(grids, p, 0.05, 2.5, bct, u)
    tiles = [(8, 4, 4), None] if len(grids) == 3 else [(8, 4, 1), (16, 8, 1), None]
    for tile in tiles:
        got = genhost.rhs(low, u, tile=tile)
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


def _rk4(F, u, t, dt):
    k1 = F(u, t)
    k2 = F(u + dt / 2 * k1, t + dt / 2)
    k3 = F(u + dt / 2 * k2, t + dt / 2)
    k4 = F(u + dt * k3, t + dt)
    return u + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)


@pytest.mark.parametrize("c", [0.8, -0.8])
def test_fused_rk4_step_on_a_nonuniform_grid(c):
    """The fused step reads per-node rows from the pool in its core positions and refills the
    Neumann / Robin ghosts each stage: equal to RK4 over the oracle."""
    x = systems.stretched_grid(301, 1.5, seed=2)
    bcl, bch = (0, 1, 0.5), (1, 2, 0.1)
    low = lower(*systems.nonuniform1d(x, 0.7, c, 2, 1, bcl, bch))
    u = np.random.default_rng(3).uniform(-1, 1, low.n_unknowns)
    got, _ = genhost.rk4_fused(low, u, 0.1, 1e-6, tile=(64, 1, 1))
    ref = _rk4(lambda w, t: orc.nonuniform1d_rhs(x, 0.7, c, 2, 1, bcl, bch, w, t), u, 0.1, 1e-6)
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))


# ---------------------------------------------------------------- mixed derivatives on nodes
@pytest.mark.parametrize("p", [2, 4])
def test_mixed_derivative_oracle_on_uniform_nodes_equals_the_uniform_oracle(p):
    """Node arrays of a uniform grid reproduce the uniform mixed-derivative system."""
    n = 17
    x = np.linspace(0.0, 1.0, n)
    u = np.random.default_rng(p).uniform(-1, 1, (n - 2) ** 2)
    a = orc.mixed2d_grid_rhs([x, x], 0.7, u, p=p)
    b = orc.mixed2d_rhs(n, 0.7, u, p=p)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("which", ["x", "xy"])
def test_mixed_derivative_exact_for_quadratics_on_random_nodes(p, which):
    """With u = q(x, y) a quadratic on the nodes and on the Dirichlet sides, the right-hand side
    is q_xx + q_yy + c q_xy = 2 q3 + 2 q5 + c q4 at every node whose mixed stencil reads no
    corner (corner nodes are 0, generate_bc_eqs.jl:396-416): the per-node centered rows (first
    and second derivative, p + 1 or more nodes) are exact for degree 2.  A row taken at the
    wrong node, or the uniform row on a nonuniform axis, breaks it."""
    rng = np.random.default_rng(11 + p)
    gx = np.concatenate([[0.0], np.sort(rng.uniform(0.02, 0.98, 13)), [1.0]])
    gy = np.concatenate([[0.0], np.sort(rng.uniform(0.02, 0.98, 11)), [1.0]]) if which == "xy" else 13
    q = (0.3, -0.7, 1.1, 0.9, -1.3, 0.4)
    c = 0.6
    xs = gx
    ys = gy if which == "xy" else np.linspace(0.0, 1.0, gy)
    X, Y = np.meshgrid(xs[1:-1], ys[1:-1], indexing="ij")
    u = (q[0] + q[1] * X + q[2] * Y + q[3] * X ** 2 + q[4] * X * Y + q[5] * Y ** 2).reshape(-1, order="F")
    du = orc.mixed2d_grid_rhs([gx, gy], c, u, p=p, q=q).reshape(X.shape, order="F")
    exact = 2 * q[3] + 2 * q[5] + c * q[4]
    # node i (1-based) reads a boundary node along an axis of n nodes when i <= p/2 + 1 or
    # i >= n - p/2; the mixed term reads a corner where that holds along both axes
    touch = lambda m: np.array([i <= p // 2 + 1 or i >= m - p // 2 for i in range(2, m)])
    keep = ~np.logical_and.outer(touch(len(xs)), touch(len(ys)))
    assert keep.sum() > 0.5 * keep.size
    err = np.abs(du - exact)[keep]
    assert np.max(err) <= 1e-9 * (1 + abs(exact)), np.max(err)


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("which", ["x", "y", "xy"])
def test_generated_mixed_derivative_on_nonuniform_axes_matches_oracle(p, which):
    """The generated kernel's mixed slot reads per-node rows of nonuniform axes (core tiles from
    the pool, boundary tiles as literals): equal to the oracle on random data."""
    gx = systems.stretched_grid(27, 1.0, seed=8)
    gy = systems.stretched_grid(21, 1.0, seed=9)
    grids = {"x": (gx, 21), "y": (27, gy), "xy": (gx, gy)}[which]
    q = (0.3, -0.7, 1.1, 0.9, -1.3, 0.4)
    low = lower(*systems.mixed2d(grids, 0.6, approx_order=p, q=q))
    u = np.random.default_rng(p).uniform(-1, 1, low.n_unknowns)
    ref = orc.mixed2d_grid_rhs(list(grids), 0.6, u, p=p, q=q)
    for tile in [(8, 4, 1), (16, 8, 1), None]:
        got = genhost.rhs(low, u, tile=tile)
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)), tile


def test_generated_mixed_derivative_reverse_mode_matches_finite_differences():
    """The reverse sweep of the mixed slot with per-node rows on both axes: the map is affine in
    u, so <J^T lam, v> equals the central difference of <lam, f> up to rounding."""
    gx = systems.stretched_grid(23, 1.0, seed=8)
    gy = systems.stretched_grid(19, 1.0, seed=9)
    q = (0.3, -0.7, 1.1, 0.9, -1.3, 0.4)
    low = lower(*systems.mixed2d((gx, gy), 0.6, approx_order=2, q=q))
    rng = np.random.default_rng(6)
    N = low.n_unknowns
    u, lam, v = rng.uniform(-1, 1, N), rng.uniform(-1, 1, N), rng.uniform(-1, 1, N)
    for tile in [(8, 4, 1), None]:
        g, _ = genhost.vjp(low, u, lam, tile=tile)
        F = lambda w: lam @ orc.mixed2d_grid_rhs([gx, gy], 0.6, w, p=2, q=q)
        fd = (F(u + v) - F(u - v)) / 2
        assert abs(g @ v - fd) <= 1e-10 * (abs(fd) + np.linalg.norm(g) * np.linalg.norm(v)), tile
