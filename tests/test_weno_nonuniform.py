"""Nonuniform node-centred WENO-5 (weno_f_nonuniform, src/discretization/schemes/WENO/
nonuniform_weno.jl) and the discretization of u_t = -c u_x on a nonuniform grid.

The oracle's reconstruction is pinned to what the reference's own tests check
(test/Components/weno_nonuniform_core.jl, weno_nonuniform_boundary.jl): exactness for
polynomials of degree <= 2 at every target node, the linear (ideal-weight) limit equal to the
derivative of the 5-point interpolating polynomial (computed here by numpy's polynomial fit, an
independent routine), fourth order under self-similar refinement, finite output under 1:1e6
stretching.  The generated kernel (compiled for the CPU) is checked against the oracle, and
its reverse mode against finite differences.
"""
import os
import sys

import numpy as np
import pytest

import oracle as orc
from molgpu import systems
from molgpu.discretization import lower

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "genhost"))
import genhost  # noqa: E402

EPS = 1e-6
GRIDS = ([0.0, 0.3, 0.9, 1.7, 2.2], [-0.3, 0.4, 0.55, 1.9, 2.4], [0.0, 0.6, 1.4, 2.1, 3.3])


def lagrange_derivative(xs, us, xt):
    """Derivative at xt of the degree-4 polynomial through the five points."""
    xs = np.asarray(xs)
    s = xs - xs.mean()
    c = np.polyfit(s, us, 4)
    return np.polyval(np.polyder(c), xt - xs.mean())


@pytest.mark.parametrize("target", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("xs", GRIDS)
def test_exact_for_quadratics_at_every_target(xs, target):
    xs = np.asarray(xs)
    xt = xs[target - 1]
    cases = [(lambda x: 1.7 + 0 * x, lambda x: 0.0), (lambda x: 1.7 + 0.9 * x, lambda x: 0.9),
             (lambda x: 1.7 + 0.9 * x - 0.4 * x * x, lambda x: 0.9 - 0.8 * x)]
    for f, df in cases:
        assert abs(orc.weno_nu(f(xs), xs, EPS, target) - df(xt)) <= 1e-12


def test_not_exact_for_cubics():
    xs = np.asarray(GRIDS[2])
    assert abs(orc.weno_nu(xs ** 3, xs, EPS, 3) - 3 * xs[2] ** 2) > 1e-6


@pytest.mark.parametrize("target", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("xs", GRIDS)
def test_linear_limit_is_five_point_interpolant_derivative(xs, target):
    """With eps >> beta the nonlinear weights equal the ideal ones, and sum_k d_k p_k'(x_t) is the
    derivative of the 5-point interpolating polynomial: exact for degree <= 4, and equal to it
    for any data (a wrong ideal weight or sub-stencil weight fails this)."""
    xs = np.asarray(xs)
    big = 1e30
    rng = np.random.default_rng(target)
    for f in (lambda x: x ** 3 - 2 * x, lambda x: x ** 4 - 0.5 * x ** 2 + x, lambda x: 2.3 * x ** 4 - 1.1 * x ** 3,
              lambda x: rng.uniform(-1, 1, 5)):
        us = f(xs)
        ref = lagrange_derivative(xs, us, xs[target - 1])
        got = orc.weno_nu(us, xs, big, target)
        assert abs(got - ref) <= 1e-10 * (1 + abs(ref) + np.max(np.abs(us))), (got, ref)


def test_fourth_order_under_self_similar_refinement():
    f = lambda x: np.sin(1.3 * x) + 0.5 * x
    df = lambda x: 1.3 * np.cos(1.3 * x) + 0.5
    o = np.array([-2.0, -1.13, 0.08, 0.91, 2.0])
    hs = [0.2, 0.1, 0.05, 0.025, 0.0125]
    errs = []
    for h in hs:
        xs = 1.0 + h * o
        errs.append(abs(orc.weno_nu(f(xs), xs, EPS, 3) - df(xs[2])))
    orders = [np.log2(errs[k] / errs[k + 1]) for k in range(len(hs) - 1)]
    assert orders[-1] > 3.85 and orders[-2] > 3.7 and min(orders) > 3.0, orders
    slope = np.polyfit(np.log(hs[2:]), np.log(errs[2:]), 1)[0]
    assert slope > 3.8


def test_extreme_stretching_stays_finite():
    lin = lambda x: 2.0 + 3.0 * x
    for g in ([0.0, 1e-6, 2e-6, 1.0, 2.0], [0.0, 1.0, 1.0 + 1e-6, 2.0 + 1e-6, 3.0 + 1e-6],
              [0.0, 1.0, 2.0, 2.0 + 1e-6, 2.0 + 2e-6]):
        g = np.asarray(g)
        for target in range(1, 6):
            assert np.isfinite(orc.weno_nu(g ** 2, g, EPS, target))
            assert abs(orc.weno_nu(lin(g), g, EPS, target) - 3.0) <= 3e-3
    for k in range(61):
        s = 10.0 ** (k / 10 - 3)
        g = np.cumsum([1.0, s, 1.0, s, 1.0]) - 1.0
        for target in range(1, 6):
            assert np.isfinite(orc.weno_nu(np.sin(g), g, EPS, target))


def _periodic_grid(n, L, seed):
    return systems.stretched_grid(n, L, seed=seed, spread=0.4)


def test_periodic_rhs_is_shift_equivariant():
    """Rotating the spacing sequence and the data of a periodic grid rotates the derivative:
    pins the coordinate continuation across the periodic end."""
    n, L, c = 33, 3.0, 1.1
    x = _periodic_grid(n, L, 5)
    h = np.diff(x)
    u = np.random.default_rng(3).uniform(-1, 1, n - 1)
    ref = orc.weno_nu1d_rhs(x, c, EPS, True, 0, 0, u)
    for s in (1, 2, 7, n - 3):
        xs = np.concatenate([[0.0], np.cumsum(np.roll(h, -s))])
        got = orc.weno_nu1d_rhs(xs, c, EPS, True, 0, 0, np.roll(u, -s))
        assert np.max(np.abs(got - np.roll(ref, -s))) <= 1e-12 * np.max(np.abs(ref))


def test_periodic_rhs_converges_on_a_smoothly_stretched_grid():
    L = 1.0
    errs = []
    ns = [41, 81, 161, 321]
    for n in ns:
        xi = np.linspace(0, L, n)
        x = xi + 0.08 * np.sin(2 * np.pi * xi) / (2 * np.pi)
        # periodic unknowns are nodes 2..n (node 1 is identified with node n)
        u = np.sin(2 * np.pi * x[1:])
        du = orc.weno_nu1d_rhs(x, 1.0, 1e-12, True, 0, 0, u)
        errs.append(np.max(np.abs(du + 2 * np.pi * np.cos(2 * np.pi * x[1:]))))
    orders = [np.log2(errs[k] / errs[k + 1]) for k in range(len(ns) - 1)]
    assert orders[-1] > 3.5, (errs, orders)


def test_dirichlet_rhs_end_nodes_use_one_sided_targets():
    """Near a Dirichlet end the reconstruction uses the first/last five nodes with targets 2 and
    n-1 (node 2 is the first unknown)."""
    n, c = 21, 0.7
    x = systems.stretched_grid(n, 1.0, seed=8)
    u = np.random.default_rng(9).uniform(-1, 1, n - 2)
    du = orc.weno_nu1d_rhs(x, c, EPS, False, 0.3, -0.2, u)
    full = np.concatenate([[0.3], u, [-0.2]])
    assert du[0] == pytest.approx(-c * orc.weno_nu(full[:5], x[:5], EPS, 2), rel=1e-14)
    assert du[1] == pytest.approx(-c * orc.weno_nu(full[:5], x[:5], EPS, 3), rel=1e-14)
    assert du[-1] == pytest.approx(-c * orc.weno_nu(full[-5:], x[-5:], EPS, 4), rel=1e-14)
    assert du[5] == pytest.approx(-c * orc.weno_nu(full[4:9], x[4:9], EPS, 3), rel=1e-14)


def test_lowering_validates_nonuniform_nodes():
    with pytest.raises(ValueError):
        lower(*systems.weno1d_nonuniform([0.0, 0.5, 0.4, 1.0, 2.0, 3.0], 1.0))
    with pytest.raises(ValueError):
        lower(*systems.weno1d_nonuniform([0.0, 0.5, 1.0, 2.0], 1.0))


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("tile", [(8, 1, 1), (16, 1, 1), None])
def test_generated_kernel_matches_oracle(periodic, tile):
    n, c = 41, 1.3
    x = systems.stretched_grid(n, 2.0, seed=3)
    low = lower(*systems.weno1d_nonuniform(x, c, 0.2, -0.1, periodic=periodic))
    assert low.slots[0]["kind"] == 6
    u = np.random.default_rng(1).uniform(-1, 1, low.n_unknowns)
    ref = orc.weno_nu1d_rhs(x, c, EPS, periodic, 0.2, -0.1, u)
    got = genhost.rhs(low, u, tile=tile)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("periodic", [False, True])
def test_generated_reverse_mode_matches_finite_differences(periodic):
    n, c = 41, 1.3
    x = systems.stretched_grid(n, 2.0, seed=3)
    low = lower(*systems.weno1d_nonuniform(x, c, 0.2, -0.1, periodic=periodic))
    rng = np.random.default_rng(2)
    N = low.n_unknowns
    u, lam, v = rng.uniform(-1, 1, N), rng.uniform(-1, 1, N), rng.uniform(-1, 1, N)
    g, _ = genhost.vjp(low, u, lam, tile=(16, 1, 1))
    F = lambda w: lam @ orc.weno_nu1d_rhs(x, c, EPS, periodic, 0.2, -0.1, w)
    h = 1e-5
    d1 = (F(u + h * v) - F(u - h * v)) / (2 * h)
    d2 = (F(u + h / 2 * v) - F(u - h / 2 * v)) / h
    fd = (4 * d2 - d1) / 3
    assert abs(g @ v - fd) <= 1e-8 * abs(fd)
