"""GPU parity: the HIP right-hand side and stage loops against the CPU oracle (FP64: 1e-12)."""
import json
import os

import numpy as np
import pytest

import oracle as orc
import molgpu
from molgpu import systems
from conftest import assert_parity

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _torch():
    import torch
    return torch


SMALL_TILES = {1: (32, 1, 1), 2: (16, 8, 1), 3: (8, 4, 4)}


def gpu_rhs(prob, u, t=0.0, path=None):
    """RHS on the GPU.  With `path`, small tiles make sure that core tiles exist, and the plan
    is checked to use the requested core-tile kernel."""
    torch = _torch()
    print("  gpu rhs: %s unknowns=%d t=%g path=%s" % (prob.lowered.fields, prob.lowered.n_unknowns, t, path),
          flush=True)
    if path is not None:
        prob.set_tile(*SMALL_TILES[prob.lowered.ndim])
        plan = prob.plan()
        if plan["core_tiles"] > 0:
            assert plan["generated"] == (1 if path == "generated" else 0), plan
    dt = prob.torch_dtype()
    ug = torch.tensor(np.asarray(u), dtype=dt, device="cuda")
    du = torch.empty_like(ug)
    prob.f(du, ug, None, t)
    torch.cuda.synchronize()
    return du.cpu().numpy().astype(np.float64)


def lap_scale(u, coef, dxs):
    return np.max(np.abs(u)) * abs(coef) * sum(4.0 / dx ** 2 for dx in dxs) + 1.0


# ---------------------------------------------------------------- heat (config 1)
@pytest.mark.parametrize("bc", ["dirichlet", "neumann", "robin"])
@pytest.mark.parametrize("p", [2, 4])
def test_heat1d_rhs(bc, p, rhs_path):
    bcs = {"dirichlet": ((1, 0, 0), (1, 0, 0)), "neumann": ((0, 1, 0), (0, 1, 0)),
           "robin": ((2.0, 0.5, 0.3), (1.5, -0.25, -0.2))}[bc]
    n = 101
    sys_, disc = systems.heat1d(n=n, D=0.8, bcl=bcs[0], bch=bcs[1], approx_order=p)
    prob = molgpu.discretize(sys_, disc)
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 1, prob.lowered.n_unknowns)
    got = gpu_rhs(prob, u, path=rhs_path)
    if p == 2:
        ref = orc.diffusion1d_rhs(n, 0.8, bcs[0], bcs[1], u)
    else:
        ref = orc.diffusion1d_rhs(n, 0.8, bcs[0], bcs[1], u, p=4)
    assert_parity(got, ref, lap_scale(u, 0.8, [0.01]) * 4, what="heat %s p=%d" % (bc, p))


def test_heat1d_rk4_closed_form():
    """Config 1: RK4, dt = 1e-5, 10^4 steps to t = 0.1: R(z)^N sin(pi x) exactly."""
    torch = _torch()
    n, D, dt, N = 101, 1.0, 1e-5, 10_000
    sys_, disc = systems.heat1d(n=n, D=D)
    prob = molgpu.discretize(sys_, disc)
    u = torch.tensor(prob.u0, dtype=torch.float64, device="cuda")
    print("  heat rk4 10^4 steps", flush=True)
    prob.integrate(u, 0.0, dt, N, molgpu.RK4, graph=True)
    torch.cuda.synchronize()
    x = np.arange(1, n - 1) / (n - 1)
    dx = 1.0 / (n - 1)
    lam = -(4 * D / dx ** 2) * np.sin(np.pi * dx / 2) ** 2
    z = dt * lam
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    got = u.cpu().numpy()
    assert np.max(np.abs(got - R ** N * prob.u0)) < 1e-12
    assert np.max(np.abs(got - np.exp(-np.pi ** 2 * D * 0.1) * np.sin(np.pi * x))) < 5e-5


# ---------------------------------------------------------------- Brusselator (config 2)
def test_brusselator_matches_generated_code_on_gpu():
    fx = json.load(open(os.path.join(GOLD, "bruss_dx025.json")))
    sys_, disc = systems.brusselator2d(n=5)
    prob = molgpu.discretize(sys_, disc)
    for s in fx["samples"]:
        got = gpu_rhs(prob, np.array(s["u"]), 0.0)
        np.testing.assert_allclose(got, s["du"], rtol=1e-13, atol=1e-11)


@pytest.mark.parametrize("n,t", [(33, 0.0), (65, 1.2), (513, 1.2)])
def test_brusselator_rhs(n, t, rhs_path):
    sys_, disc = systems.brusselator2d(n=n)
    prob = molgpu.discretize(sys_, disc)
    rng = np.random.default_rng(n)
    u = rng.uniform(0.5, 1.5, prob.lowered.n_unknowns)
    got = gpu_rhs(prob, u, t, path=rhs_path)
    ref = orc.bruss2d_rhs(n, u, t, threads=8)
    dx = 1.0 / (n - 1)
    assert_parity(got, ref, lap_scale(u, 10.0, [dx, dx]), what="bruss n=%d" % n)


def test_brusselator_initial_condition_and_rk4():
    torch = _torch()
    n = 65
    sys_, disc = systems.brusselator2d(n=n)
    prob = molgpu.discretize(sys_, disc)
    g = np.arange(n) / (n - 1)
    X, Y = np.meshgrid(g[1:], g[1:], indexing="ij")
    u0 = np.concatenate([(22 * (Y * (1 - Y)) ** 1.5).reshape(-1, order="F"),
                         (27 * (X * (1 - X)) ** 1.5).reshape(-1, order="F")])
    np.testing.assert_allclose(prob.u0, u0, rtol=1e-14, atol=1e-15)
    dt, N = 2e-6, 50
    u = torch.tensor(prob.u0, dtype=torch.float64, device="cuda")
    print("  bruss rk4 %d steps" % N, flush=True)
    prob.integrate(u, 1.0, dt, N, molgpu.RK4)
    torch.cuda.synchronize()
    ref = orc.bruss2d_integrate(n, prob.u0, 1.0, dt, N, threads=8)
    assert_parity(u.cpu().numpy(), ref, rtol=1e-12, what="bruss rk4")


@pytest.mark.parametrize("graph", [False, True])
def test_brusselator_fused_rk4_step_equals_stage_launches(graph):
    """The fused RK4 step kernel (one launch per step) against the four stage launches of the
    same steps, across the forcing switch at t = 1.1, with an odd step count (the state ends
    in the workspace and is copied back)."""
    torch = _torch()
    n = 129
    prob = molgpu.discretize(*systems.brusselator2d(n=n))
    assert prob.plan()["fused_tiles"] > 0
    dt, N, t0 = 1e-7, 69, 1.1 - 3.3e-6     # stable: dt * 8 * alpha / dx^2 = 0.13
    res = {}
    for fused in (True, False):
        u = torch.tensor(prob.u0, dtype=torch.float64, device="cuda")
        print("  bruss fused=%s graph=%s" % (fused, graph), flush=True)
        prob.integrate(u, t0, dt, N, molgpu.RK4, graph=graph, fused=fused)
        torch.cuda.synchronize()
        res[fused] = u.cpu().numpy()
    diff = np.max(np.abs(res[True] - res[False]))
    print("  fused vs stage launches: max abs difference %.3e" % diff)
    assert_parity(res[True], res[False], rtol=1e-13, what="fused rk4")
    ref = orc.bruss2d_integrate(n, prob.u0, t0, dt, N, threads=8)
    assert_parity(res[True], ref, rtol=1e-12, what="fused rk4 vs oracle")


@pytest.mark.parametrize("case", ["brusselator_p4", "nonlinlap_periodic"])
def test_fused_rk4_reach_two_equals_stage_launches(case):
    torch = _torch()
    if case == "brusselator_p4":
        n = 257
        prob = molgpu.discretize(*systems.brusselator2d(n=n, approx_order=4))
        dt, N = 1e-7, 31
        ref = lambda u0: orc.bruss2d_integrate(n, u0, 0.0, dt, N, p=4, threads=8)
    else:
        n = 2001
        prob = molgpu.discretize(*systems.nonlinlap1d(n, True, 0.0, 0.0, 2))
        dt, N = 2e-8, 31
        ref = None
    assert prob.plan()["fused_tiles"] > 0, prob.plan()
    u0 = np.asarray(prob.u0, dtype=np.float64)
    res = {}
    for fused in (True, False):
        u = torch.tensor(u0, dtype=torch.float64, device="cuda")
        print("  %s fused=%s" % (case, fused), flush=True)
        prob.integrate(u, 0.0, dt, N, molgpu.RK4, graph=True, fused=fused)
        torch.cuda.synchronize()
        res[fused] = u.cpu().numpy()
    assert_parity(res[True], res[False], rtol=1e-13, what="fused vs stage launches")
    if ref is not None:
        assert_parity(res[True], ref(u0), rtol=1e-12, what="fused vs oracle")


def test_fp32_fused_rk4_step():
    """FP32 systems (without the compensated update) take the fused step too: equal to the
    stage launches to FP32 rounding, and within FP32 accuracy of the FP64 oracle."""
    torch = _torch()
    n = 257
    prob = molgpu.discretize(*systems.brusselator2d(n=n, dtype="f32"))
    assert prob.plan()["fused_tiles"] > 0
    dt, N = 2e-7, 40     # stable: dt * 8 * alpha / dx^2 = 1.05
    res = {}
    for fused in (True, False):
        u = torch.tensor(prob.u0, dtype=torch.float32, device="cuda")
        print("  bruss fp32 fused=%s" % fused, flush=True)
        prob.integrate(u, 0.0, dt, N, molgpu.RK4, graph=True, fused=fused)
        torch.cuda.synchronize()
        res[fused] = u.cpu().numpy().astype(np.float64)
    scale = np.max(np.abs(res[False]))
    assert np.max(np.abs(res[True] - res[False])) <= 1e-6 * scale
    ref = orc.bruss2d_integrate(n, np.asarray(prob.u0, dtype=np.float64), 0.0, dt, N, threads=8)
    assert np.max(np.abs(res[True] - ref)) <= 2e-5 * scale


@pytest.mark.parametrize("case", ["heat_dirichlet", "heat_neumann_robin", "diffreact_mixed", "diffreact_neumann"])
def test_fused_rk4_with_boundary_closures_equals_stage_launches(case):
    """Non-periodic systems take the fused step too (ghosts refilled from their closures after
    each stage): equal to the four stage launches, and to the oracle's RK4."""
    torch = _torch()
    if case.startswith("heat"):
        n = 1001
        bcl, bch = ((1, 0, 0.0), (1, 0, 0.0)) if case == "heat_dirichlet" else ((0, 1, 0.5), (2, -1, 0.4))
        prob = molgpu.discretize(*systems.heat1d(n=n, D=1.0, bcl=bcl, bch=bch))
        dt, N = 5e-8, 41
        ref = lambda u0: orc.diffusion1d_integrate(n, 1.0, bcl, bch, u0, 0.0, dt, N)
    else:
        n = 257
        bct = (0, 1, 1, 0) if case == "diffreact_mixed" else (1, 1, 1, 1)
        prob = molgpu.discretize(*systems.diffreact2d(n, 0.05, 2.5, bct))
        dt, N = 1e-4, 41

        def ref(u0):
            u = u0.copy()
            F = lambda w: orc.diffreact2d_rhs(n, 0.05, 2.5, bct, w)
            for _ in range(N):
                k1 = F(u); k2 = F(u + dt / 2 * k1); k3 = F(u + dt / 2 * k2); k4 = F(u + dt * k3)
                u = u + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
            return u
    assert prob.plan()["fused_tiles"] > 0, prob.plan()
    u0 = np.random.default_rng(5).uniform(0, 1, prob.lowered.n_unknowns)
    res = {}
    for fused in (True, False):
        u = torch.tensor(u0, dtype=torch.float64, device="cuda")
        print("  %s fused=%s" % (case, fused), flush=True)
        prob.integrate(u, 0.0, dt, N, molgpu.RK4, graph=True, fused=fused)
        torch.cuda.synchronize()
        res[fused] = u.cpu().numpy()
    assert_parity(res[True], res[False], rtol=1e-13, what="fused vs stage launches")
    assert_parity(res[True], ref(u0), rtol=1e-12, what="fused vs oracle")


# ---------------------------------------------------------------- Burgers WENO5 (config 3)
@pytest.mark.parametrize("n", [33, 129])
def test_burgers_weno_rhs(n, rhs_path):
    sys_, disc = systems.burgers2d(n=n)
    prob = molgpu.discretize(sys_, disc)
    rng = np.random.default_rng(n)
    for u in (prob.u0, rng.uniform(0.2, 1.2, prob.lowered.n_unknowns)):
        got = gpu_rhs(prob, u, path=rhs_path)
        ref = orc.burgers2d_rhs(n, u, threads=8)
        assert_parity(got, ref, np.max(np.abs(u)) ** 2 * 4 * (n - 1) + 1, what="burgers n=%d" % n)


def test_burgers_ssprk3_matches_oracle():
    torch = _torch()
    n = 65
    sys_, disc = systems.burgers2d(n=n)
    prob = molgpu.discretize(sys_, disc)
    umax = np.max(np.abs(prob.u0))
    dx = 1.0 / (n - 1)
    dt = 0.4 / (umax / dx + umax / dx)
    N = 20
    u = torch.tensor(prob.u0, dtype=torch.float64, device="cuda")
    print("  burgers ssprk3", flush=True)
    prob.integrate(u, 0.0, dt, N, molgpu.SSPRK33)
    torch.cuda.synchronize()
    ref = orc.burgers2d_integrate(n, prob.u0, 0.0, dt, N, threads=8)
    assert_parity(u.cpu().numpy(), ref, rtol=1e-12, what="burgers ssprk3")


def test_burgers_fp32_kahan_error_vs_fp64_oracle():
    torch = _torch()
    n = 65
    sys_, disc = systems.burgers2d(n=n, dtype="f32")
    prob = molgpu.discretize(sys_, disc)
    umax = np.max(np.abs(prob.u0))
    dx = 1.0 / (n - 1)
    dt = 0.4 / (2 * umax / dx)
    N = 40
    u = torch.tensor(prob.u0, dtype=torch.float32, device="cuda")
    prob.integrate(u, 0.0, dt, N, molgpu.SSPRK33, kahan=True)
    torch.cuda.synchronize()
    ref = orc.burgers2d_integrate(n, prob.u0, 0.0, dt, N, threads=8)
    err = np.max(np.abs(u.cpu().numpy() - ref))
    print("  fp32+kahan max error vs fp64 oracle: %.3e" % err)
    # FP32 WENO weights on barely resolved bumps (width 0.02 at dx = 1/64) amplify rounding;
    # the error is reported, the bound only guards against a broken FP32 path
    assert err < 2e-3


def test_weno_dirichlet_extrapolated_nodes(rhs_path):
    n, L, c = 41, 2.0, 1.0
    sys_, disc = systems.weno1d_dirichlet(n, L, c, 0.1, 0.05)
    prob = molgpu.discretize(sys_, disc)
    rng = np.random.default_rng(4)
    u = rng.uniform(0, 1, prob.lowered.n_unknowns)
    got = gpu_rhs(prob, u, path=rhs_path)
    ref = orc.weno1d_dirichlet_rhs(n, L, c, 1e-6, 0.1, 0.05, u)
    assert_parity(got, ref, 4 * (n - 1) / L + 1, what="weno dirichlet")


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("n,tile", [(41, (16, 1, 1)), (41, None), (3001, None)])
def test_weno_nonuniform_rhs(periodic, n, tile):
    """Node-centred WENO-5 on a stretched grid (generated kernel only: the interpreter has no
    nonuniform WENO operator)."""
    c = 1.3
    x = systems.stretched_grid(n, 2.0, seed=3)
    sys_, disc = systems.weno1d_nonuniform(x, c, 0.2, -0.1, periodic=periodic)
    prob = molgpu.discretize(sys_, disc)
    if tile is not None:
        prob.set_tile(*tile)
    assert prob.plan()["generated"] == 1
    u = np.random.default_rng(n).uniform(-1, 1, prob.lowered.n_unknowns)
    got = gpu_rhs(prob, u)
    ref = orc.weno_nu1d_rhs(x, c, 1e-6, periodic, 0.2, -0.1, u)
    assert_parity(got, ref, np.max(np.abs(ref)) + 1, what="weno nonuniform")


NU_BCS = [((1, 0, 0.3), (1, 0, -0.2)), ((0, 1, 0.5), (1, 0, 0.0)), ((1, 2, 0.1), (2, -1, 0.4))]


@pytest.mark.parametrize("bc", range(3))
@pytest.mark.parametrize("p,c", [(2, 0.8), (4, -0.8), (6, 0.0)])
@pytest.mark.parametrize("n,tile", [(41, (8, 1, 1)), (5001, None)])
def test_nonuniform_centered_upwind_rhs(bc, p, c, n, tile):
    """Per-node centered / upwind rows on a stretched grid, Dirichlet / Neumann / Robin ends."""
    bcl, bch = NU_BCS[bc]
    x = systems.stretched_grid(n, 1.5, seed=2)
    sys_, disc = systems.nonuniform1d(x, 0.7, c, p, 1, bcl, bch)
    prob = molgpu.discretize(sys_, disc)
    if tile is not None:
        prob.set_tile(*tile)
    u = np.random.default_rng(p).uniform(-1, 1, prob.lowered.n_unknowns)
    got = gpu_rhs(prob, u)
    ref = orc.nonuniform1d_rhs(x, 0.7, c, p, 1, bcl, bch, u)
    h = np.min(np.diff(x))
    assert_parity(got, ref, np.max(np.abs(u)) * (0.7 * 8 / h ** 2 + 2 / h) + 1, what="nonuniform fd")


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("kind", ["nonlinlap", "spherical"])
@pytest.mark.parametrize("n,tile", [(41, (8, 1, 1)), (3001, None)])
def test_nonuniform_nonlinear_and_spherical_laplacians(p, kind, n, tile):
    x = systems.stretched_grid(n, 1.0, seed=6)
    if kind == "nonlinlap":
        sys_, disc = systems.nonlinlap1d(x, False, 0.2, -0.3, p)
        ref_fn = lambda w: orc.nonlinlap1d_nu_rhs(x, p, 0.2, -0.3, w)
    else:
        sys_, disc = systems.spherical1d(x, 0.4, p)
        ref_fn = lambda w: orc.spherical1d_nu_rhs(x, p, 0.4, w)
    prob = molgpu.discretize(sys_, disc)
    if tile is not None:
        prob.set_tile(*tile)
    u = np.random.default_rng(p).uniform(-1, 1, prob.lowered.n_unknowns)
    got = gpu_rhs(prob, u)
    ref = ref_fn(u)
    h = np.min(np.diff(x))
    assert_parity(got, ref, 2 * 8 / h ** 2 + 1, what="nonuniform %s p=%d" % (kind, p))


@pytest.mark.parametrize("which", ["x", "xy", "xz3d"])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_nonuniform_multidimensional_diffusion(which, dtype):
    """Per-node rows on 2D / 3D grids (small tiles and the default), FP64 at 1e-12 and FP32
    against the FP64 oracle at a single-precision tolerance."""
    xg = systems.stretched_grid(129 if "3d" not in which else 40, 1.0, seed=4)
    yg = systems.stretched_grid(97, 1.0, seed=5)
    zg = systems.stretched_grid(33, 1.0, seed=7)
    grids = {"x": [xg, 97], "xy": [xg, yg], "xz3d": [xg, 36, zg]}[which]
    bct = (0, 1, 1, 0) + ((1, 0) if len(grids) == 3 else ())
    sys_, disc = systems.diffreact_grids(grids, 0.05, 2.5, bct, 2, dtype=dtype)
    for tile in ([(16, 8, 1), None] if len(grids) == 2 else [(8, 4, 4), None]):
        prob = molgpu.discretize(sys_, disc)
        if tile is not None:
            prob.set_tile(*tile)
        u = np.random.default_rng(2).uniform(0, 1, prob.lowered.n_unknowns)
        got = gpu_rhs(prob, u)
        ref = orc.diffreact_nd_rhs(grids, 2, 0.05, 2.5, bct, u)
        h = min(np.min(np.diff(g)) for g in grids if not np.isscalar(g))
        scale = 0.05 * len(grids) * 4 / h ** 2 + 3
        if dtype == "f64":
            assert_parity(got, ref, scale, what="nonuniform %s" % which)
        else:
            assert np.max(np.abs(got - ref)) <= 2e-6 * scale, np.max(np.abs(got - ref))


def test_nonuniform_fused_rk4_equals_stage_launches():
    torch = _torch()
    grids = [systems.stretched_grid(257, 1.0, seed=4), systems.stretched_grid(193, 1.0, seed=5)]
    bct = (0, 1, 1, 0)
    prob = molgpu.discretize(*systems.diffreact_grids(grids, 0.05, 2.5, bct, 2))
    assert prob.plan()["fused_tiles"] > 0, prob.plan()
    u0 = np.random.default_rng(5).uniform(0, 1, prob.lowered.n_unknowns)
    h = min(np.min(np.diff(g)) for g in grids)
    dt, N = 0.1 * h * h / 0.05 / 8, 21
    res = {}
    for fused in (True, False):
        u = torch.tensor(u0, dtype=torch.float64, device="cuda")
        print("  nonuniform diffreact fused=%s" % fused, flush=True)
        prob.integrate(u, 0.0, dt, N, molgpu.RK4, graph=True, fused=fused)
        torch.cuda.synchronize()
        res[fused] = u.cpu().numpy()
    assert_parity(res[True], res[False], rtol=1e-13, what="nonuniform fused vs stage launches")
    ref = u0.copy()
    F = lambda w: orc.diffreact_nd_rhs(grids, 2, 0.05, 2.5, bct, w)
    for _ in range(N):
        k1 = F(ref); k2 = F(ref + dt / 2 * k1); k3 = F(ref + dt / 2 * k2); k4 = F(ref + dt * k3)
        ref = ref + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    assert_parity(res[True], ref, rtol=1e-12, what="nonuniform fused vs oracle")


def test_nonuniform_operators_refuse_the_interpreter(monkeypatch):
    """With code generation switched off the handle fails loudly instead of running a kernel
    that has no per-node rows."""
    monkeypatch.setenv("MOL_JIT", "0")
    x = systems.stretched_grid(41, 1.0, seed=2)
    sys_, disc = systems.nonuniform1d(x, 0.7, 0.5, 2, 1)
    with pytest.raises(molgpu.native.MolError, match="generated kernels only"):
        prob = molgpu.discretize(sys_, disc)
        gpu_rhs(prob, np.zeros(prob.lowered.n_unknowns))


def test_nonuniform_rk4_matches_rk4_over_the_oracle():
    n, p, c = 61, 4, 0.6
    bcl, bch = (0, 1, 0.2), (1, 0, 0.1)
    x = systems.stretched_grid(n, 1.0, seed=7)
    sys_, disc = systems.nonuniform1d(x, 0.05, c, p, 1, bcl, bch)
    prob = molgpu.discretize(sys_, disc)
    u0 = np.asarray(prob.u0, dtype=np.float64)
    dt, nsteps = 2e-4, 40
    F = lambda w: orc.nonuniform1d_rhs(x, 0.05, c, p, 1, bcl, bch, w)
    ref = u0.copy()
    for _ in range(nsteps):
        k1 = F(ref); k2 = F(ref + dt / 2 * k1); k3 = F(ref + dt / 2 * k2); k4 = F(ref + dt * k3)
        ref = ref + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    torch = _torch()
    print("  nonuniform rk4 n=%d steps=%d" % (n, nsteps), flush=True)
    ug = torch.tensor(u0, dtype=torch.float64, device="cuda")
    prob.integrate(ug, 0.0, dt, nsteps, molgpu.RK4)
    torch.cuda.synchronize()
    got = ug.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-11 * (1 + np.max(np.abs(ref)))


# ---------------------------------------------------------------- diffusion-reaction (config 4 instance)
@pytest.mark.parametrize("bct", [(0, 0, 0, 0), (1, 1, 1, 1), (0, 1, 1, 0), (1, 0, 0, 1)])
def test_diffreact2d_rhs(bct, rhs_path):
    n = 48
    sys_, disc = systems.diffreact2d(n, 0.05, 2.5, bct)
    prob = molgpu.discretize(sys_, disc)
    rng = np.random.default_rng(sum(bct))
    u = rng.uniform(0, 1, prob.lowered.n_unknowns)
    got = gpu_rhs(prob, u, path=rhs_path)
    ref = orc.diffreact2d_rhs(n, 0.05, 2.5, bct, u)
    dx = 1.0 / (n - 1)
    assert_parity(got, ref, lap_scale(u, 0.05, [dx, dx]), what="diffreact %s" % (bct,))


# ---------------------------------------------------------------- nonlinear diffusion (config 5)
@pytest.mark.parametrize("n", [12, 34])
def test_nonlindiff3d_rhs(n, rhs_path):
    sys_, disc = systems.nonlindiff3d(n=n)
    prob = molgpu.discretize(sys_, disc)
    d0 = 0.5 + 0.5 * np.abs(systems.gaussian_field3d(n, 1))
    src = systems.gaussian_field3d(n, 2)
    prob.set_grid_field("d0", d0)
    prob.set_grid_field("s", src)
    rng = np.random.default_rng(n)
    u = rng.uniform(-1, 1, prob.lowered.n_unknowns)
    got = gpu_rhs(prob, u, path=rhs_path)
    ref = orc.nonlindiff3d_rhs(n, d0.reshape(-1, order="F"), src.reshape(-1, order="F"), u, threads=8)
    dx = 1.0 / (n - 1)
    assert_parity(got, ref, lap_scale(u, 2.0 * np.max(d0), [dx] * 3), what="nonlindiff3d n=%d" % n)


# ---------------------------------------------------------------- other operator families
@pytest.mark.parametrize("order", [1, 2])
def test_upwind_sign_selection(order, rhs_path):
    n = 61
    sys_, disc = systems.upwind1d(n, 0.1, 1.0, 0.3, -0.2, order)
    prob = molgpu.discretize(sys_, disc)
    rng = np.random.default_rng(order)
    u = rng.uniform(-1, 1, prob.lowered.n_unknowns)
    got = gpu_rhs(prob, u, path=rhs_path)
    ref = orc.upwind1d_rhs(n, order, 0.1, 1.0, 0.3, -0.2, u)
    assert_parity(got, ref, 4 * 1.1 * (n - 1) * 2, what="upwind order %d" % order)


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("which", ["x", "xy"])
def test_mixed_derivative_on_nonuniform_axes(p, which):
    """Per-node first-derivative rows of nonuniform axes in the mixed slot (generated kernel
    only), small tiles for core and boundary tiles, against the oracle."""
    torch = _torch()
    gx = systems.stretched_grid(41, 1.0, seed=8)
    gy = systems.stretched_grid(33, 1.0, seed=9)
    grids = (gx, 33) if which == "x" else (gx, gy)
    q = (0.3, -0.7, 1.1, 0.9, -1.3, 0.4)
    prob = molgpu.discretize(*systems.mixed2d(grids, 0.6, approx_order=p, q=q))
    u = np.random.default_rng(p).uniform(-1, 1, prob.lowered.n_unknowns)
    ref = orc.mixed2d_grid_rhs(list(grids), 0.6, u, p=p, q=q)
    for tile in [(16, 8, 1), None]:
        if tile:
            prob.set_tile(*tile)
        ug = torch.tensor(u, device="cuda")
        du = torch.empty_like(ug)
        print("  nonuniform mixed p=%d %s tile=%s" % (p, which, tile), flush=True)
        prob.f(du, ug, None, 0.0)
        torch.cuda.synchronize()
        assert_parity(du.cpu().numpy(), ref, rtol=1e-12, what="nonuniform mixed p=%d %s" % (p, which))


@pytest.mark.parametrize("p", [2, 4])
def test_mixed_derivative(p, rhs_path):
    n = 25
    sys_, disc = systems.mixed2d(n, 0.7, p)
    prob = molgpu.discretize(sys_, disc)
    rng = np.random.default_rng(p)
    u = rng.uniform(-1, 1, prob.lowered.n_unknowns)
    got = gpu_rhs(prob, u, path=rhs_path)
    ref = orc.mixed2d_rhs(n, 0.7, u, p=p)
    dx = 1.0 / (n - 1)
    assert_parity(got, ref, lap_scale(u, 4.0, [dx, dx]), what="mixed p=%d" % p)


@pytest.mark.parametrize("p", [2, 4])
def test_spherical_laplacian(p, rhs_path):
    n = 41
    sys_, disc = systems.spherical1d(n, 0.25, p)
    prob = molgpu.discretize(sys_, disc)
    rng = np.random.default_rng(p + 10)
    u = rng.uniform(0, 1, prob.lowered.n_unknowns)
    got = gpu_rhs(prob, u, path=rhs_path)
    ref = orc.spherical1d_rhs(n, 0.25, u, p=p)
    dx = 1.0 / (n - 1)
    assert_parity(got, ref, lap_scale(u, 8.0, [dx]) * 10, what="spherical p=%d" % p)


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("p", [2, 4])
def test_nonlinear_laplacian_1d(periodic, p, rhs_path):
    n = 41
    sys_, disc = systems.nonlinlap1d(n, periodic, 0.2, -0.4, p)
    prob = molgpu.discretize(sys_, disc)
    rng = np.random.default_rng(p)
    u = rng.uniform(-1, 1, prob.lowered.n_unknowns)
    got = gpu_rhs(prob, u, path=rhs_path)
    ref = orc.nonlinlap1d_rhs(n, u, periodic, 0.2, -0.4, p=p)
    dx = 1.0 / (n - 1)
    assert_parity(got, ref, lap_scale(u, 8.0, [dx]) * 4, what="nonlinlap periodic=%s p=%d" % (periodic, p))


def test_host_pointer_entry_point():
    n = 17
    sys_, disc = systems.brusselator2d(n=n)
    prob = molgpu.discretize(sys_, disc)
    rng = np.random.default_rng(21)
    u = rng.uniform(0.5, 1.5, prob.lowered.n_unknowns)
    du = np.empty_like(u)
    prob.f(du, u, prob.p, 1.3)
    assert_parity(du, orc.bruss2d_rhs(n, u, 1.3), lap_scale(u, 10.0, [1 / 16] * 2), what="rhs_h")


def test_kahan_recovers_lost_fp32_increments():
    """FP32 heat equation with increments far below the state's ulp: plain accumulation loses
    them, the compensated update keeps the FP64 trajectory to FP32 precision."""
    torch = _torch()
    n, D, dt, N = 101, 1.0, 2e-8, 20_000
    res = {}
    for kahan in (False, True):
        sys_, disc = systems.heat1d(n=n, D=D, dtype="f32")
        prob = molgpu.discretize(sys_, disc)
        u = torch.tensor(prob.u0, dtype=torch.float32, device="cuda")
        print("  fp32 heat kahan=%s" % kahan, flush=True)
        prob.integrate(u, 0.0, dt, N, molgpu.RK4, kahan=kahan, graph=True)
        torch.cuda.synchronize()
        res[kahan] = u.cpu().numpy().astype(np.float64)
    dx = 1.0 / (n - 1)
    lam = -(4 * D / dx ** 2) * np.sin(np.pi * dx / 2) ** 2
    z = dt * lam
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    exact = R ** N * np.sin(np.pi * np.arange(1, n - 1) / (n - 1))
    e_plain = np.max(np.abs(res[False] - exact))
    e_kahan = np.max(np.abs(res[True] - exact))
    print("  fp32 error plain %.3e, kahan %.3e" % (e_plain, e_kahan))
    assert e_kahan < 2e-6
    assert e_kahan < e_plain / 10
