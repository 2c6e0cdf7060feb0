"""Interface boundaries between variables on adjacent intervals (interface_boundary.jl:79-111,
generate_bc_eqs.jl:35-58, interior_map.jl:1-10): lowering, oracle pins and the generated
kernel on the CPU."""
import os
import sys

import numpy as np
import pytest
import sympy as sp

import oracle as orc
from molgpu.discretization import MOLFiniteDifference, lower
from molgpu.pdesys import Differential, Eq, Interval, PDESystem, parameters, variables

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "genhost"))
import genhost  # noqa: E402


def interface_system(n1, n2, D1=1.0, D2=1.0, gl=0.0, gr=0.0, p=2, L1=0.5, dx=None):
    """c1 on [0, L1] (n1 nodes), c2 on [L1, L1 + (n2 - 1) dx] with the same spacing, Dirichlet
    gl / gr at the outer ends, c1(t, L1) ~ c2(t, L1) (the reference's interface test, Test 14 of
    test/Diffusion/MOL_1D_Linear_Diffusion.jl, with general sizes and coefficients)."""
    t, x1, x2 = parameters("t x1 x2")
    c1, c2 = variables("c1 c2")
    dx = L1 / (n1 - 1) if dx is None else dx
    L2 = L1 + (n2 - 1) * dx
    Dx1, Dx2 = Differential(x1), Differential(x2)
    eqs = [Eq(Differential(t)(c1(t, x1)), D1 * (Dx1 ** 2)(c1(t, x1))),
           Eq(Differential(t)(c2(t, x2)), D2 * (Dx2 ** 2)(c2(t, x2)))]
    bcs = [Eq(c1(0, x1), -x1 * (x1 - 1) * sp.sin(x1)), Eq(c2(0, x2), x2 * (x2 - 1) * sp.sin(x2)),
           Eq(c1(t, 0), gl), Eq(c1(t, L1), c2(t, L1)), Eq(c2(t, L2), gr)]
    doms = [Interval(t, 0.0, 1.0), Interval(x1, 0.0, L1), Interval(x2, L1, L2)]
    return (PDESystem(eqs, bcs, doms, [t, x1, x2], [c1(t, x1), c2(t, x2)]),
            MOLFiniteDifference({x1: n1, x2: n2}, t, approx_order=p))


@pytest.mark.parametrize("p", [2, 4])
def test_interface_oracle_equals_one_domain(p):
    """With one coefficient on both sides the joined system is the one-variable diffusion on the
    union grid: the wrapped reads, the alias of the upper variable's first node and the interior
    rows next to the interface give the same right-hand side, unknown by unknown."""
    n1, n2, dx = 11, 15, 0.05
    N = n1 + n2 - 1
    u = np.random.default_rng(p).uniform(-1, 1, N - 2)
    got = orc.interface1d_rhs(n1, n2, dx, 0.7, 0.7, u, p=p, gl=0.3, gr=-0.2)
    ref = orc.diffusion1d_rhs(N, 0.7, (1.0, 0.0, 0.3), (1.0, 0.0, -0.2), u, p=p, xlo=0.0, xhi=(N - 1) * dx)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_interface_lowering_boxes_and_closures():
    """The lower variable keeps the interface node, the upper one takes it as an alias; the
    positions past each side of the interface read the other variable."""
    n1, n2 = 11, 15
    low = lower(*interface_system(n1, n2))
    N = n1 + n2 - 1
    assert low.n == [N]
    assert (low.lo[0][0], low.hi[0][0]) == (2, n1)           # c1: nodes 2..n1
    assert (low.lo[1][0], low.hi[1][0]) == (n1 + 1, N - 1)   # c2: nodes 2..n2-1
    src = {(c.field, c.side): c.src for c in low.closures}
    assert src[(0, 1)] == 1 and src[(1, 0)] == 0
    assert src[(0, 0)] == -1 and src[(1, 1)] == -1
    np.testing.assert_allclose(low.coords[0], np.arange(N) * 0.05, atol=1e-14)
    # initial values of each variable at its own nodes
    x = low.coords[0]
    np.testing.assert_allclose(low.u0[:n1 - 1], -x[1:n1] * (x[1:n1] - 1) * np.sin(x[1:n1]), rtol=1e-14)


def test_interface_lowering_refuses_what_it_does_not_build():
    t, x1, x2 = parameters("t x1 x2")
    c1, c2 = variables("c1 c2")
    eqs = [Eq(Differential(t)(c1(t, x1)), (Differential(x1) ** 2)(c1(t, x1))),
           Eq(Differential(t)(c2(t, x2)), (Differential(x2) ** 2)(c2(t, x2)))]
    base = [Eq(c1(0, x1), 0), Eq(c2(0, x2), 0), Eq(c1(t, 0), 0), Eq(c2(t, 1), 0)]
    doms = [Interval(t, 0.0, 1.0), Interval(x1, 0.0, 0.5), Interval(x2, 0.5, 1.0)]
    # no interface condition where the intervals meet
    with pytest.raises(NotImplementedError):
        lower(PDESystem(eqs, base, doms, [t, x1, x2], [c1(t, x1), c2(t, x2)]), MOLFiniteDifference({x1: 11, x2: 11}, t))
    # grids of different spacing across the interface
    with pytest.raises(NotImplementedError):
        lower(PDESystem(eqs, base + [Eq(c1(t, 0.5), c2(t, 0.5))], doms, [t, x1, x2], [c1(t, x1), c2(t, x2)]),
              MOLFiniteDifference({x1: 11, x2: 21}, t))


@pytest.mark.parametrize("p", [2, 4])
def test_generated_interface_rhs_matches_oracle(p):
    """Different coefficients on the two sides: the generated kernel (interface ghosts staged
    from the other variable) against the oracle, with tiles that split the interface."""
    n1, n2, dx = 17, 23, 0.5 / 16
    low = lower(*interface_system(n1, n2, D1=0.4, D2=1.3, gl=0.25, gr=-0.5, p=p))
    u = np.random.default_rng(p).uniform(-1, 1, low.n_unknowns)
    ref = orc.interface1d_rhs(n1, n2, dx, 0.4, 1.3, u, p=p, gl=0.25, gr=-0.5)
    for tile in [(8, 1, 1), (16, 1, 1), None]:
        got = genhost.rhs(low, u, tile=tile)
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)), tile
