"""GPU: edge-aligned grids (grid_align = edge_align) against the oracle's restatement, and the
reference's edge-aligned heat test (test/Diffusion/MOL_1D_Linear_Diffusion.jl:59-82): u_t = u_xx on
[0, pi], u(t, 0) = exp(-t), u(t, pi) = -exp(-t), u0 = cos(x), cell centres within 0.01 of
exp(-t) cos(x) at t = 0.1, 0.2, ..., 1."""
import numpy as np
import pytest
import sympy as sp

import oracle as orc
import molgpu
from molgpu import systems
from molgpu.discretization import MOLFiniteDifference, edge_align
from molgpu.pdesys import Differential, Eq, Interval, PDESystem, parameters, variables
from conftest import assert_parity

pytestmark = pytest.mark.gpu

BCS = {"dirichlet": ((1.0, 0.0), (1.0, 0.0)), "neumann": ((0.0, 1.0), (0.0, 1.0)),
       "robin": ((2.0, 0.5), (1.5, -0.25))}


def _torch():
    import torch
    return torch


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("bc", sorted(BCS))
def test_edge_rhs_matches_oracle(bc, p):
    torch = _torch()
    ncell, D = 200, 0.6
    (a0l, a1l), (a0h, a1h) = BCS[bc]
    bcl, bch = (a0l, a1l, 0.3), (a0h, a1h, -0.2)
    prob = molgpu.discretize(*systems.heat1d(n=ncell, D=D, bcl=bcl, bch=bch, approx_order=p, grid_align="edge"))
    rng = np.random.default_rng(7)
    u = rng.uniform(-1, 1, ncell)
    print("  edge rhs %s p=%d" % (bc, p), flush=True)
    ug = torch.tensor(u, device="cuda")
    du = torch.empty_like(ug)
    prob.f(du, ug, None, 0.0)
    ref = orc.diffusion1d_edge_rhs(ncell, D, bcl, bch, u, p=p)
    assert_parity(du.cpu().numpy(), ref, np.max(np.abs(u)) * D * 16 * ncell ** 2 + 1, what="edge %s p=%d" % (bc, p))


def test_reference_edge_aligned_heat():
    torch = _torch()
    t, x = parameters("t x")
    u = variables("u")
    Dxx = Differential(x) ** 2
    eq = Eq(Differential(t)(u(t, x)), Dxx(u(t, x)))
    bcs = [Eq(u(0, x), sp.cos(x)), Eq(u(t, 0), sp.exp(-t)), Eq(u(t, sp.pi), -sp.exp(-t))]
    pdesys = PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, 0.0, float(np.pi))], [t, x], [u(t, x)])
    dx = float(np.pi) / 29           # the step of range(0, pi, length = 30)
    prob = molgpu.discretize(pdesys, MOLFiniteDifference({x: dx}, t, grid_align=edge_align))
    xc = prob.lowered.coords[0][1:-1]
    assert len(xc) == 29 and abs(xc[0] - dx / 2) < 1e-15
    ug = torch.tensor(prob.u0, dtype=torch.float64, device="cuda")
    dt, per = 1e-3, 100
    print("  reference edge-aligned heat", flush=True)
    worst = 0.0
    for k in range(10):
        prob.integrate(ug, k * per * dt, dt, per, molgpu.RK4)
        tk = (k + 1) * per * dt
        err = np.max(np.abs(ug.cpu().numpy() - np.exp(-tk) * np.cos(xc)))
        worst = max(worst, err)
        assert err < 0.01, (tk, err)
    print("  max error %.2e" % worst)
    # second order in space: the error is far inside the reference's tolerance
    assert worst < 2e-3
