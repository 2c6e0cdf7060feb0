"""Solution arrays from device state (mol_expand): full grids per variable with the observed
boundary values, as the reference's PDETimeSeriesSolution returns them
(src/interface/solution/timedep.jl:19-93; boundary equations generate_bc_eqs.jl:238-392,
corners :396-416).  Expected boundary values come from the closed forms of the boundary
equations at approximation order 2 (D1 boundary row [-3/2, 2, -1/2] / dx).
"""
import numpy as np
import pytest

import molgpu
import oracle as orc
from molgpu import systems

pytestmark = pytest.mark.gpu


def _ghost_low(a0, a1, g, u2, u3, dx):
    """Node 1 from a0 u1 + a1 (-3/2 u1 + 2 u2 - 1/2 u3) / dx = g."""
    return (g - a1 * (2 * u2 - 0.5 * u3) / dx) / (a0 - 1.5 * a1 / dx)


def _ghost_high(a0, a1, g, un1, un2, dx):
    """Node n from a0 un + a1 (3/2 un - 2 u(n-1) + 1/2 u(n-2)) / dx = g."""
    return (g - a1 * (-2 * un1 + 0.5 * un2) / dx) / (a0 + 1.5 * a1 / dx)


@pytest.mark.parametrize("bcl,bch", [((1, 0, 0.0), (1, 0, 0.0)), ((0, 1, 0.4), (1, 0, -0.3)),
                                     ((1, 2, 0.1), (2, -1, 0.4))])
def test_heat1d_solution_has_observed_boundary_values(bcl, bch):
    n, D = 41, 0.5
    sys_, disc = systems.heat1d(n=n, D=D, bcl=bcl, bch=bch)
    prob = molgpu.discretize(sys_, disc)
    print("  solve heat1d bcl=%s bch=%s" % (bcl, bch), flush=True)
    sol = molgpu.solve(prob, molgpu.RK4, dt=1e-4, saveat=0.01, tspan=(0.0, 0.03))
    u = sol["u"]
    assert u.shape == (len(sol.t), n)
    np.testing.assert_array_equal(u[:, 1:-1], sol.interior("u"))
    np.testing.assert_array_equal(sol[prob.lowered.axes[0]], prob.lowered.coords[0])
    dx = 1.0 / (n - 1)
    for k in range(len(sol.t)):
        lo = _ghost_low(*bcl, u[k, 1], u[k, 2], dx)
        hi = _ghost_high(*bch, u[k, -2], u[k, -3], dx)
        assert abs(u[k, 0] - lo) <= 1e-12 * (1 + abs(lo))
        assert abs(u[k, -1] - hi) <= 1e-12 * (1 + abs(hi))
    # the unknowns of the last save match the oracle's RK4
    ref = orc.diffusion1d_integrate(n, D, bcl, bch, prob.u0, 0.0, 1e-4, 300)
    assert np.max(np.abs(sol.u[-1] - ref)) <= 1e-12 * (1 + np.max(np.abs(ref)))


def test_diffreact2d_faces_and_corners():
    """Mixed Dirichlet / Neumann sides; nodes outside the box along both axes are 0."""
    n = 21
    bct = (0, 1, 1, 0)   # x-lo Dirichlet, x-hi Neumann, y-lo Neumann, y-hi Dirichlet
    sys_, disc = systems.diffreact2d(n, 0.05, 2.5, bct)
    prob = molgpu.discretize(sys_, disc)
    torch = pytest.importorskip("torch")
    u = torch.tensor(np.random.default_rng(3).uniform(0, 1, prob.lowered.n_unknowns), dtype=torch.float64,
                     device="cuda")
    print("  expand diffreact2d", flush=True)
    full = prob.expand(u, 0.0).cpu().numpy()[0]          # [i, j]
    box = prob.unflatten(u.cpu().numpy())[0]
    np.testing.assert_array_equal(full[1:-1, 1:-1], box)
    inner = slice(1, n - 1)
    assert np.all(full[0, inner] == 0.0)                                            # x-lo Dirichlet 0
    np.testing.assert_allclose(full[-1, inner], (4 * full[-2, inner] - full[-3, inner]) / 3, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(full[inner, 0], (4 * full[inner, 1] - full[inner, 2]) / 3, rtol=1e-13, atol=1e-14)
    assert np.all(full[inner, -1] == 0.0)                                           # y-hi Dirichlet 0
    for c in [(0, 0), (0, -1), (-1, 0), (-1, -1)]:
        assert full[c] == 0.0


def test_periodic_node_one_aliases_node_n():
    n = 33
    sys_, disc = systems.brusselator2d(n=n)
    prob = molgpu.discretize(sys_, disc)
    print("  solve brusselator n=%d" % n, flush=True)
    sol = molgpu.solve(prob, molgpu.RK4, dt=1e-5, saveat=1e-4, tspan=(0.0, 2e-4))
    for name in ("u", "v"):
        g = sol[name]
        assert g.shape == (len(sol.t), n, n)
        np.testing.assert_array_equal(g[:, 0, :], g[:, -1, :])
        np.testing.assert_array_equal(g[:, :, 0], g[:, :, -1])
        np.testing.assert_array_equal(g[:, 1:, 1:], sol.interior(name))
    # the initial condition on the full grid (generate_ic_defaults) at the saved t = 0
    x = prob.lowered.coords[0]
    X, Y = np.meshgrid(x, x, indexing="ij")
    np.testing.assert_allclose(sol["u"][0], 22 * (Y * (1 - Y)) ** 1.5, rtol=1e-14, atol=1e-14)
