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


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_edge_2d_exact_laplacian_on_gpu(dtype):
    """Edge-aligned 2D: Dirichlet data from u = p(x) q(y) (quadratics, p = 2); the device
    right-hand side is the exact Laplacian at every cell centre (FP32: to FP32 rounding)."""
    torch = _torch()
    t, x, y = parameters("t x y")
    u = variables("u")
    Dx, Dy = Differential(x), Differential(y)
    P = (0.3 + 0.7 * x - 0.4 * x ** 2) * (1.1 - 0.5 * y + 0.6 * y ** 2)
    eq = Eq(Differential(t)(u(t, x, y)), (Dx ** 2)(u(t, x, y)) + (Dy ** 2)(u(t, x, y)))
    bcs = [Eq(u(0, x, y), P), Eq(u(t, 0, y), P.subs(x, 0)), Eq(u(t, 1, y), P.subs(x, 1)),
           Eq(u(t, x, 0), P.subs(y, 0)), Eq(u(t, x, 1), P.subs(y, 1))]
    sys_ = PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, 0.0, 1.0), Interval(y, 0.0, 1.0)],
                     [t, x, y], [u(t, x, y)])
    prob = molgpu.discretize(sys_, MOLFiniteDifference({x: 300, y: 200}, t, grid_align=edge_align, dtype=dtype))
    xc, yc = prob.lowered.coords[0][1:-1], prob.lowered.coords[1][1:-1]
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    lap = sp.lambdify((x, y), sp.diff(P, x, 2) + sp.diff(P, y, 2))
    want = lap(X, Y).reshape(-1, order="F")
    print("  edge 2d %s" % dtype, flush=True)
    ug = torch.tensor(prob.u0, dtype=prob.torch_dtype(), device="cuda")
    du = torch.empty_like(ug)
    prob.f(du, ug, None, 0.0)
    got = du.cpu().numpy().astype(np.float64)
    # FP32: the Laplacian of O(1) values at h = 1/300 cancels to about eps32 * 8 / h^2
    tol = 1e-8 if dtype == "f64" else 8 * 1.2e-7 * 8 * 300 ** 2
    assert np.max(np.abs(got - want)) < tol
