"""Edge-aligned grids (grid_align = edge_align): nodes at the cell centres plus a ghost past each
end (generate_grid, src/discretization/discretize_vars.jl:373-390), conditions at the cell edges
through half-point interpolation and differences (boundary_value_maps,
src/discretization/generate_bc_eqs.jl:79-161).  The oracle's restatement is pinned by
polynomial exactness (the half-point rows are exact on polynomials of their order, so the
right-hand side of u = polynomial is D u'' at every cell centre), then the lowering and the
generated kernel (on the CPU) are checked against it."""
import os
import sys
from fractions import Fraction

import numpy as np
import pytest
import sympy as sp

import oracle as orc
import molgpu
from molgpu import systems
from molgpu.discretization import MOLFiniteDifference, edge_align, edge_grid, lower
from molgpu.pdesys import Differential, Eq, Interval, PDESystem, parameters, variables
from conftest import assert_parity

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "genhost"))
import genhost  # noqa: E402


def test_edge_grid_is_the_shifted_range():
    """test/Components/DiscreteSpace.jl:79-96: grid == (xmin - dx/2):dx:(xmax + dx/2)."""
    for lo, hi, dx in ((0.0, 1.0, 0.1), (-1.0, 2.0, 0.25), (0.0, float(np.pi), float(np.pi) / 29)):
        g = edge_grid(lo, hi, dx)
        ncell = round((hi - lo) / dx)
        assert len(g) == ncell + 2
        want = [float(Fraction(lo - dx / 2) + k * Fraction(dx)) for k in range(ncell + 2)]
        assert np.array_equal(g, np.array(want))
        assert abs(g[0] - (lo - dx / 2)) < 1e-15 and abs(g[-1] - (hi + dx / 2)) < 1e-12
    # an integer step is a cell count (prepare_dx, discretize_vars.jl:284)
    t, x = parameters("t x")
    u = variables("u")
    sys_, disc = systems.heat1d(n=20, grid_align="edge")
    low = lower(sys_, disc)
    assert low.n == [22] and low.lo == [[2]] and low.hi == [[21]]
    assert abs(low.coords[0][1] - 0.025) < 1e-16 and abs(low.coords[0][2] - low.coords[0][1] - 0.05) < 1e-15


BCS = {"dirichlet": ((1.0, 0.0), (1.0, 0.0)), "neumann": ((0.0, 1.0), (0.0, 1.0)),
       "robin": ((2.0, 0.5), (1.5, -0.25))}


@pytest.mark.parametrize("p,deg", [(2, 2), (4, 3)])
@pytest.mark.parametrize("bc", sorted(BCS))
def test_oracle_exact_on_polynomials(bc, p, deg):
    ncell, D = 16, 0.7
    h = 1.0 / ncell
    xc = (np.arange(ncell) + 0.5) * h
    c = np.array([0.3, -1.1, 0.8, 0.45])[:deg + 1]
    poly = np.polynomial.Polynomial(c)
    d1, d2 = poly.deriv(1), poly.deriv(2)
    (a0l, a1l), (a0h, a1h) = BCS[bc]
    bcl = (a0l, a1l, a0l * poly(0.0) + a1l * d1(0.0))
    bch = (a0h, a1h, a0h * poly(1.0) + a1h * d1(1.0))
    got = orc.diffusion1d_edge_rhs(ncell, D, bcl, bch, poly(xc), p=p)
    want = D * d2(xc)
    assert np.max(np.abs(got - want)) < 1e-9 * max(1.0, np.max(np.abs(want)))


def test_oracle_not_exact_beyond_the_order():
    """A wrong weight set that is still exact on low degrees would pass the test above only by
    luck: a quartic is not reproduced at p = 2 next to a Dirichlet edge."""
    ncell, D = 16, 1.0
    xc = (np.arange(ncell) + 0.5) / ncell
    poly = np.polynomial.Polynomial([0.0, 0.0, 0.0, 0.0, 1.0])
    got = orc.diffusion1d_edge_rhs(ncell, D, (1, 0, 0.0), (1, 0, 1.0), poly(xc), p=2)
    assert np.max(np.abs(got - 12 * xc ** 2)) > 1e-3


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("bc", sorted(BCS))
def test_generated_kernel_matches_oracle(bc, p):
    ncell, D = 24, 0.6
    (a0l, a1l), (a0h, a1h) = BCS[bc]
    bcl, bch = (a0l, a1l, 0.3), (a0h, a1h, -0.2)
    sys_, disc = systems.heat1d(n=ncell, D=D, bcl=bcl, bch=bch, approx_order=p, grid_align="edge")
    low = lower(sys_, disc)
    assert low.n_unknowns == ncell
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, ncell)
    got = genhost.rhs(low, u, tile=(8, 1, 1))
    ref = orc.diffusion1d_edge_rhs(ncell, D, bcl, bch, u, p=p)
    assert_parity(got, ref, np.max(np.abs(u)) * D * 16 * ncell ** 2 + 1, what="edge %s p=%d" % (bc, p))


def test_initial_values_at_cell_centres():
    sys_, disc = systems.heat1d(n=10, grid_align="edge")
    low = lower(sys_, disc)
    xc = (np.arange(10) + 0.5) / 10
    assert np.allclose(low.u0, np.sin(np.pi * xc), rtol=0, atol=1e-15)


def test_edge_grid_refusals():
    t, x = parameters("t x")
    u = variables("u")
    Dx = Differential(x)
    eq = Eq(Differential(t)(u(t, x)), (Dx ** 2)(u(t, x)))
    per = PDESystem([eq], [Eq(u(0, x), sp.sin(2 * sp.pi * x)), Eq(u(t, 0), u(t, 1))],
                    [Interval(t, 0.0, 1.0), Interval(x, 0.0, 1.0)], [t, x], [u(t, x)])
    with pytest.raises(NotImplementedError):
        lower(per, MOLFiniteDifference({x: 0.1}, t, grid_align=edge_align))
    sys_, _ = systems.heat1d(n=10)
    with pytest.raises(NotImplementedError):
        lower(sys_, MOLFiniteDifference({sys_.ivs[1]: np.linspace(0, 1, 11)}, t, grid_align=edge_align))
    with pytest.raises(NotImplementedError):
        MOLFiniteDifference({x: 0.1}, t, grid_align="stagger")


@pytest.mark.parametrize("p", [2, 4])
def test_generated_2d_kernel_exact_on_polynomials(p):
    """u_t = u_xx + u_yy on [0,1]^2 with edge-aligned axes and Dirichlet conditions taken from
    u = polynomial (degree 2 at p = 2, 3 at p = 4, in each variable): the generated kernel's
    right-hand side is the exact Laplacian at every cell centre, so the 2D closures (tangential
    coordinates of the edge ghosts at cell centres) are right."""
    t, x, y = parameters("t x y")
    u = variables("u")
    Dx, Dy = Differential(x), Differential(y)
    deg = 2 if p == 2 else 3
    P = (0.3 + 0.7 * x - 0.4 * x ** 2 + (0.2 * x ** 3 if deg == 3 else 0)) * (1.1 - 0.5 * y + 0.6 * y ** 2 +
                                                                              (-0.3 * y ** 3 if deg == 3 else 0))
    eq = Eq(Differential(t)(u(t, x, y)), (Dx ** 2)(u(t, x, y)) + (Dy ** 2)(u(t, x, y)))
    bcs = [Eq(u(0, x, y), P), Eq(u(t, 0, y), P.subs(x, 0)), Eq(u(t, 1, y), P.subs(x, 1)),
           Eq(u(t, x, 0), P.subs(y, 0)), Eq(u(t, x, 1), P.subs(y, 1))]
    sys_ = PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, 0.0, 1.0), Interval(y, 0.0, 1.0)],
                     [t, x, y], [u(t, x, y)])
    low = lower(sys_, MOLFiniteDifference({x: 12, y: 10}, t, approx_order=p, grid_align=edge_align))
    assert low.n == [14, 12]
    xc, yc = low.coords[0][1:-1], low.coords[1][1:-1]
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    f = sp.lambdify((x, y), P)
    lap = sp.lambdify((x, y), sp.diff(P, x, 2) + sp.diff(P, y, 2))
    u0 = f(X, Y).reshape(-1, order="F")
    assert np.allclose(low.u0, u0, rtol=0, atol=1e-14)
    got = genhost.rhs(low, u0, tile=(8, 4, 1))
    want = lap(X, Y).reshape(-1, order="F")
    assert np.max(np.abs(got - want)) < 1e-9 * max(1.0, np.max(np.abs(want)))
