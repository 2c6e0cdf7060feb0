"""GPU: the fixed-step loop's launch paths (direct, replayed step chunks, resident fused steps)
give the same trajectory and the same device clock, and the host-pointer right-hand side
(the literal ODEFunction drop-in) keeps no per-call allocations."""
import time

import numpy as np
import pytest

import oracle as orc
import molgpu
from molgpu import systems
from conftest import assert_parity

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _time_dependent_heat(n):
    """u_t = u_xx / 2 + sin(x) u, u(t, 0) = sin(5t), u_x(t, 1) = cos(3t): a wrong stage or step
    time moves both boundary values."""
    from molgpu.pdesys import Differential, Eq, Interval, PDESystem, parameters, variables
    from molgpu.discretization import MOLFiniteDifference
    import sympy as sp
    t, x = parameters("t x")
    u = variables("u")
    Dx = Differential(x)
    eq = Eq(Differential(t)(u(t, x)), 0.5 * (Dx ** 2)(u(t, x)) + sp.sin(x) * u(t, x))
    bcs = [Eq(u(0, x), x * (1 - x)), Eq(u(t, 0), sp.sin(5 * t)), Eq(Dx(u(t, 1)), sp.cos(3 * t))]
    return PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, 0.0, 1.0)], [t, x], [u(t, x)]), \
        MOLFiniteDifference({x: n}, t)


def _time_dependent_heat2d(n):
    """2D variant: u(t, 0, y) = sin(5t) y, u_y(t, x, 1) = cos(3t), Neumann 0 elsewhere."""
    from molgpu.pdesys import Differential, Eq, Interval, PDESystem, parameters, variables
    from molgpu.discretization import MOLFiniteDifference
    import sympy as sp
    t, x, y = parameters("t x y")
    u = variables("u")
    Dx, Dy = Differential(x), Differential(y)
    eq = Eq(Differential(t)(u(t, x, y)), 0.05 * ((Dx ** 2)(u(t, x, y)) + (Dy ** 2)(u(t, x, y))) + u(t, x, y))
    bcs = [Eq(u(0, x, y), x * (1 - x) * y),
           Eq(u(t, 0, y), sp.sin(5 * t) * y), Eq(Dx(u(t, 1, y)), 0),
           Eq(Dy(u(t, x, 0)), 0), Eq(Dy(u(t, x, 1)), sp.cos(3 * t))]
    doms = [Interval(t, 0.0, 1.0), Interval(x, 0.0, 1.0), Interval(y, 0.0, 1.0)]
    return PDESystem([eq], bcs, doms, [t, x, y], [u(t, x, y)]), MOLFiniteDifference({x: n, y: n}, t)


def _run(prob, u0, t0, dt, N, **kw):
    torch = _torch()
    u = torch.tensor(u0, dtype=torch.float64, device="cuda")
    prob.integrate(u, t0, dt, N, molgpu.RK4, **kw)
    torch.cuda.synchronize()
    return u.cpu().numpy()


@pytest.mark.parametrize("dim", [1, 2])
def test_graph_chunks_advance_the_clock(dim):
    """Replayed chunks must advance the device clock like direct launches: with time-dependent
    boundary values the graph run (chunks of 64, 16 and 4 steps and a direct remainder, called
    twice so the second call only replays) equals the direct launches, fused and not."""
    if dim == 1:
        prob = molgpu.discretize(*_time_dependent_heat(61))
        dt, N = 1e-4, 303          # 4 x 64 + 2 x 16 + 3 x 4 + 3
    else:
        prob = molgpu.discretize(*_time_dependent_heat2d(65))
        dt, N = 2e-5, 87           # 64 + 16 + 4 + 3
    assert prob.plan()["fused_tiles"] > 0, prob.plan()
    u0 = np.random.default_rng(7).uniform(-1, 1, prob.lowered.n_unknowns)
    t0 = 0.3
    print("  dim=%d N=%d direct" % (dim, N), flush=True)
    ref = _run(prob, u0, t0, dt, N, graph=False, fused=False)
    scale = 1 + np.max(np.abs(ref))
    for fused in (True, False):
        for graph in (False, True):
            for call in range(2 if graph else 1):
                print("  dim=%d fused=%s graph=%s call=%d" % (dim, fused, graph, call), flush=True)
                got = _run(prob, u0, t0, dt, N, graph=graph, fused=fused)
                assert np.max(np.abs(got - ref)) <= 1e-13 * scale, (fused, graph, call)
    # and the clock continues from where a previous call left it: two half runs = one run
    torch = _torch()
    u = torch.tensor(u0, dtype=torch.float64, device="cuda")
    h = N // 2
    prob.integrate(u, t0, dt, h, molgpu.RK4, graph=True)
    prob.integrate(u, t0 + h * dt, dt, N - h, molgpu.RK4, graph=True)
    torch.cuda.synchronize()
    assert np.max(np.abs(u.cpu().numpy() - ref)) <= 1e-12 * scale


def test_graph_chunks_follow_the_state_pointer():
    """A second state tensor gets chunks of its own; the first one's are still replayed."""
    torch = _torch()
    n = 129
    prob = molgpu.discretize(*systems.brusselator2d(n=n))
    u0 = np.asarray(prob.u0, dtype=np.float64)
    dt, N = 1e-6, 21
    ref = orc.bruss2d_integrate(n, u0, 0.0, dt, N, threads=8)
    a = torch.tensor(u0, device="cuda")
    b = torch.tensor(u0, device="cuda")
    for x in (a, b, a):
        x.copy_(torch.tensor(u0, device="cuda"))
        prob.integrate(x, 0.0, dt, N, molgpu.RK4, graph=True)
        torch.cuda.synchronize()
        assert_parity(x.cpu().numpy(), ref, rtol=1e-12, what="graph replay")


def test_host_pointer_entry_point_repeated_calls():
    """mol_rhs_h called 1000 times: every result equals the oracle's, and a call costs at most
    twice the host-to-device plus device-to-host copy of the state (no per-call device
    allocation or device-wide synchronization)."""
    torch = _torch()
    n = 257
    prob = molgpu.discretize(*systems.brusselator2d(n=n))
    rng = np.random.default_rng(3)
    u = rng.uniform(0.5, 1.5, prob.lowered.n_unknowns)
    du = np.empty_like(u)
    ref = orc.bruss2d_rhs(n, u, 1.3, threads=8)
    prob.f(du, u, prob.p, 1.3)          # first call allocates the handle's copies
    assert_parity(du, ref, 10.0 * 8 * (n - 1) ** 2 * np.max(np.abs(u)), what="rhs_h")
    first = du.copy()
    calls = 1000
    t0 = time.perf_counter()
    for _ in range(calls):
        prob.f(du, u, prob.p, 1.3)
    per_call = (time.perf_counter() - t0) / calls
    assert np.array_equal(du, first)
    # the copies alone: host to device and back through torch, same bytes
    ut = torch.from_numpy(u)
    for _ in range(10):
        ut.to("cuda").cpu()
    t0 = time.perf_counter()
    for _ in range(calls):
        d = ut.to("cuda")
        d.cpu()
    per_copy = (time.perf_counter() - t0) / calls
    print("  rhs_h %.1f us per call, copies %.1f us" % (per_call * 1e6, per_copy * 1e6))
    assert per_call <= 2 * per_copy, (per_call, per_copy)


@pytest.mark.parametrize("bcl,bch", [((1.0, 0.0, 0.3), (0.0, 1.0, -0.2)), ((1.0, 2.0, 0.1), (2.0, -1.0, 0.4))])
def test_fused_two_step_launches_1d_against_oracle(bcl, bch):
    """A 1D system too large for one resident tile runs fused launches of two RK4 steps (ghosts
    refilled from the Dirichlet / Neumann / Robin closures after every stage, and after the first
    step of a launch) plus one single step for the odd remainder: equal to the oracle's RK4."""
    torch = _torch()
    n = 4001
    prob = molgpu.discretize(*systems.heat1d(n=n, D=1.0, bcl=bcl, bch=bch))
    plan = prob.plan()
    assert plan["fused_tiles"] > 1 and plan["fused_steps_per_launch"] == 2, plan
    u0 = np.random.default_rng(5).uniform(-1, 1, prob.lowered.n_unknowns)
    dt, N = 2e-8, 37
    print("  1d fused two-step n=%d N=%d %s %s" % (n, N, bcl, bch), flush=True)
    got = _run(prob, u0, 0.0, dt, N)
    ref = orc.diffusion1d_integrate(n, 1.0, bcl, bch, u0, 0.0, dt, N)
    assert np.max(np.abs(got - ref)) <= 1e-13 * (1 + np.max(np.abs(ref)))
