"""The generated kernels, compiled for the host (tests/genhost), against the oracle.

The source the library generates for gfx950 runs here with its thread loops collapsed (one
"thread" per work-group covering the tile), so the forward and reverse-mode code paths —
interior-row and boundary-row variants, ghost closures, direct (unstaged) arrays, both halos —
are checked on the CPU at FP64 parity, without a GPU.  Small tiles put most nodes on boundary
tiles.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "genhost"))

import genhost  # noqa: E402
import oracle as orc  # noqa: E402
from conftest import assert_parity  # noqa: E402
from molgpu import systems  # noqa: E402
from molgpu.discretization import lower  # noqa: E402

TILES = {1: [(16, 1, 1), None], 2: [(8, 8, 1), None], 3: [(8, 4, 4), None]}


def _lap_scale(u, coef, dxs):
    return np.max(np.abs(u)) * abs(coef) * sum(4.0 / dx ** 2 for dx in dxs) + 1.0


def _cases():
    """(name, builder -> (lowered, u, t, aux, oracle rhs, scale))."""
    def heat(bc, p):
        def b():
            bcs = {"dirichlet": ((1, 0, 0), (1, 0, 0)), "neumann": ((0, 1, 0), (0, 1, 0)),
                   "robin": ((2.0, 0.5, 0.3), (1.5, -0.25, -0.2))}[bc]
            n = 61
            low = lower(*systems.heat1d(n=n, D=0.8, bcl=bcs[0], bch=bcs[1], approx_order=p))
            u = np.random.default_rng(1).uniform(-1, 1, low.n_unknowns)
            ref = orc.diffusion1d_rhs(n, 0.8, bcs[0], bcs[1], u, p=p)
            return low, u, 0.0, None, ref, _lap_scale(u, 0.8, [1 / (n - 1)]) * 4
        return b

    def bruss():
        n = 33
        low = lower(*systems.brusselator2d(n=n))
        u = np.random.default_rng(2).uniform(0.5, 1.5, low.n_unknowns)
        return low, u, 1.2, None, orc.bruss2d_rhs(n, u, 1.2), _lap_scale(u, 10.0, [1 / (n - 1)] * 2)

    def burgers():
        n = 33
        low = lower(*systems.burgers2d(n=n))
        u = np.random.default_rng(3).uniform(0.2, 1.2, low.n_unknowns)
        return low, u, 0.0, None, orc.burgers2d_rhs(n, u), np.max(np.abs(u)) ** 2 * 4 * (n - 1) + 1

    def weno_dir():
        n, L, c = 41, 2.0, 1.0
        low = lower(*systems.weno1d_dirichlet(n, L, c, 0.1, 0.05))
        u = np.random.default_rng(4).uniform(0, 1, low.n_unknowns)
        return low, u, 0.0, None, orc.weno1d_dirichlet_rhs(n, L, c, 1e-6, 0.1, 0.05, u), 4 * (n - 1) / L + 1

    def diffreact(bct):
        def b():
            n = 21
            low = lower(*systems.diffreact2d(n, 0.05, 2.5, bct))
            u = np.random.default_rng(5).uniform(0, 1, low.n_unknowns)
            return low, u, 0.0, None, orc.diffreact2d_rhs(n, 0.05, 2.5, bct, u), _lap_scale(u, 0.05, [1 / (n - 1)] * 2)
        return b

    def nonlin():
        n = 12
        low = lower(*systems.nonlindiff3d(n=n))
        d0 = 0.5 + 0.5 * np.abs(systems.gaussian_field3d(n, 1))
        src = systems.gaussian_field3d(n, 2)
        u = np.random.default_rng(6).uniform(-1, 1, low.n_unknowns)
        ref = orc.nonlindiff3d_rhs(n, d0.reshape(-1, order="F"), src.reshape(-1, order="F"), u)
        return low, u, 0.0, {"d0": d0, "s": src}, ref, _lap_scale(u, 2.0 * np.max(d0), [1 / (n - 1)] * 3)

    def upwind(order):
        def b():
            n = 61
            low = lower(*systems.upwind1d(n, 0.1, 1.0, 0.3, -0.2, order))
            u = np.random.default_rng(order).uniform(-1, 1, low.n_unknowns)
            return low, u, 0.0, None, orc.upwind1d_rhs(n, order, 0.1, 1.0, 0.3, -0.2, u), 4 * 1.1 * (n - 1) * 2
        return b

    def mixed(p):
        def b():
            n = 25
            low = lower(*systems.mixed2d(n, 0.7, p))
            u = np.random.default_rng(p).uniform(-1, 1, low.n_unknowns)
            return low, u, 0.0, None, orc.mixed2d_rhs(n, 0.7, u, p=p), _lap_scale(u, 4.0, [1 / (n - 1)] * 2)
        return b

    def spherical(p):
        def b():
            n = 41
            low = lower(*systems.spherical1d(n, 0.25, p))
            u = np.random.default_rng(p + 10).uniform(0, 1, low.n_unknowns)
            return low, u, 0.0, None, orc.spherical1d_rhs(n, 0.25, u, p=p), _lap_scale(u, 8.0, [1 / (n - 1)]) * 10
        return b

    def nonlinlap(periodic, p):
        def b():
            n = 41
            low = lower(*systems.nonlinlap1d(n, periodic, 0.2, -0.4, p))
            u = np.random.default_rng(p).uniform(-1, 1, low.n_unknowns)
            ref = orc.nonlinlap1d_rhs(n, u, periodic, 0.2, -0.4, p=p)
            return low, u, 0.0, None, ref, _lap_scale(u, 8.0, [1 / (n - 1)]) * 4
        return b

    return {
        "heat_dirichlet_p2": heat("dirichlet", 2), "heat_neumann_p4": heat("neumann", 4),
        "heat_robin_p4": heat("robin", 4), "brusselator": bruss, "burgers_weno": burgers,
        "weno_dirichlet": weno_dir, "diffreact_0000": diffreact((0, 0, 0, 0)),
        "diffreact_0110": diffreact((0, 1, 1, 0)), "nonlindiff3d": nonlin, "upwind_o1": upwind(1),
        "upwind_o2": upwind(2), "mixed_p2": mixed(2), "mixed_p4": mixed(4), "spherical_p2": spherical(2),
        "spherical_p4": spherical(4), "nonlinlap_periodic_p4": nonlinlap(True, 4),
        "nonlinlap_dirichlet_p2": nonlinlap(False, 2),
    }


CASES = _cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_generated_forward_kernel_matches_oracle(name):
    low, u, t, aux, ref, scale = CASES[name]()
    for tile in TILES[low.ndim]:
        got = genhost.rhs(low, u, t=t, aux=aux, tile=tile)
        assert_parity(got, ref, scale, what="%s tile %s" % (name, tile))


@pytest.mark.parametrize("name", ["heat_robin_p4", "brusselator", "burgers_weno", "weno_dirichlet",
                                  "diffreact_0110", "nonlindiff3d", "upwind_o2", "mixed_p4", "spherical_p2",
                                  "nonlinlap_periodic_p4"])
def test_generated_vjp_kernel_dot_product(name):
    """<(df/du)^T lam, v> of the generated reverse-mode kernel against the Richardson-extrapolated
    central difference of <lam, f(u + e v)> with the same system's generated forward kernel
    (itself pinned to the oracle above)."""
    low, u, t, aux, ref, scale = CASES[name]()
    rng = np.random.default_rng(17)
    lam = rng.standard_normal(low.n_unknowns)
    v = rng.standard_normal(low.n_unknowns)
    for tile in TILES[low.ndim]:
        out, g = genhost.vjp(low, u, lam, t=t, aux=aux, tile=tile)

        def val(e):
            return lam @ genhost.rhs(low, u + e * v, t=t, aux=aux, tile=tile)

        eps = 1e-4 * max(1.0, np.max(np.abs(u)))
        d1 = (val(eps) - val(-eps)) / (2 * eps)
        d2 = (val(eps / 2) - val(-eps / 2)) / eps
        fd = (4 * d2 - d1) / 3
        assert abs(out @ v - fd) <= 1e-7 * abs(fd) + 1e-9 * np.linalg.norm(out) * np.linalg.norm(v), (name, tile)


def test_generated_vjp_grid_field_gradients_match_oracle_differences():
    """(df/dd0)^T lam and (df/ds)^T lam of the 3D nonlinear diffusion against the oracle."""
    low, u, t, aux, ref, scale = CASES["nonlindiff3d"]()
    n = low.n[0]
    rng = np.random.default_rng(23)
    lam = rng.standard_normal(low.n_unknowns)
    w = rng.standard_normal(n ** 3)
    d0 = aux["d0"].reshape(-1, order="F")
    src = aux["s"].reshape(-1, order="F")
    for tile in TILES[3]:
        out, g = genhost.vjp(low, u, lam, t=t, aux=aux, tile=tile)

        def val(e):
            return lam @ orc.nonlindiff3d_rhs(n, d0 + e * w, src, u)

        d1 = (val(1e-4) - val(-1e-4)) / 2e-4
        d2 = (val(5e-5) - val(-5e-5)) / 1e-4
        fd = (4 * d2 - d1) / 3
        gd0, gs = g[low.aux.index("d0")], g[low.aux.index("s")]
        assert abs(gd0 @ w - fd) <= 1e-8 * abs(fd), (tile, gd0 @ w, fd)
        # s enters linearly at the node: its gradient is lam on the unknown nodes, 0 elsewhere
        box = np.zeros((n, n, n))
        (l0, l1, l2), (h0, h1, h2) = low.lo[0], low.hi[0]
        box[l0 - 1:h0, l1 - 1:h1, l2 - 1:h2] = lam.reshape((h0 - l0 + 1, h1 - l1 + 1, h2 - l2 + 1), order="F")
        np.testing.assert_allclose(gs, box.reshape(-1, order="F"), rtol=0, atol=1e-14)


@pytest.mark.parametrize("rb", [2, 4])
@pytest.mark.parametrize("name", ["brusselator", "burgers_weno", "diffreact_0110", "nonlindiff3d", "mixed_p4"])
def test_generated_forward_kernel_register_blocked(name, rb, monkeypatch):
    """Several nodes per thread along the slowest axis (MOL_RB forces it on small tiles),
    including tiles that overhang the unknown box, against the oracle."""
    monkeypatch.setenv("MOL_RB", str(rb))
    low, u, t, aux, ref, scale = CASES[name]()
    tiles = {2: [(8, 8, 1), (16, 4, 1)], 3: [(8, 4, 4), (4, 4, 8)]}[low.ndim]
    for tile in tiles:
        src = __import__("molgpu").codegen(low, tile=tile)
        assert "static constexpr int RB = %d;" % rb in src
        got = genhost.rhs(low, u, t=t, aux=aux, tile=tile)
        assert_parity(got, ref, scale, what="%s tile %s RB %d" % (name, tile, rb))


@pytest.mark.parametrize("grid", [8, 3])
@pytest.mark.parametrize("name", ["brusselator", "diffreact_0110", "nonlindiff3d", "heat_robin_p4"])
def test_generated_forward_kernel_persistent_launch(name, grid):
    """Fewer work-groups than tiles: each walks several tiles (core and boundary variants in
    turn) with the next tile's staging prefetched, against the oracle."""
    low, u, t, aux, ref, scale = CASES[name]()
    tile = {1: (16, 1, 1), 2: (8, 8, 1), 3: (4, 4, 4)}[low.ndim]
    if low.ndim == 3:
        # 3D tiles run one work-group per tile (the library launches the persistent grid only
        # for a source with the persistent loop)
        assert "mol_tiles_persistent<" not in __import__("molgpu").codegen(low, tile=tile)
        grid = None
    got = genhost.rhs(low, u, t=t, aux=aux, tile=tile, grid=grid)
    assert_parity(got, ref, scale, what="%s grid %s" % (name, grid))


@pytest.mark.parametrize("n,tile", [(33, None), (33, (16, 8, 1)), (37, (16, 16, 1)), (29, (8, 8, 1))])
def test_generated_fused_rk4_step_matches_stage_arithmetic(n, tile):
    """The fused RK4 step (four stages on a tile with a 4-node halo in LDS) of the periodic
    Brusselator, including ragged tilings, against the same step built from four evaluations
    of the oracle RHS with the stage kernels' update arithmetic, and the device clock advance."""
    low = lower(*systems.brusselator2d(n=n))
    u = np.random.default_rng(n).uniform(0.5, 1.5, low.n_unknowns)
    t, dt = 1.1 - 6e-6, 1e-5  # the forcing switches on at t = 1.1 in the last stage
    got, tnext = genhost.rk4_fused(low, u, t, dt, tile=tile)
    f = lambda y, s: orc.bruss2d_rhs(n, y, s)
    k1 = f(u, t)
    y1 = dt / 2 * k1 + u
    k2 = f(y1, t + dt / 2)
    y2 = dt / 2 * k2 + u
    k3 = f(y2, t + dt / 2)
    y3 = dt * k3 + u
    k4 = f(y3, t + dt)
    acc = (1.0 * k1 + 2.0 * k2) + 2.0 * k3
    ref = (dt / 6 * k4 + u) + dt / 6 * acc
    assert_parity(got, ref, np.max(np.abs(u)) * 2, what="fused n=%d tile %s" % (n, tile))
    assert tnext == t + dt


@pytest.mark.parametrize("name,tile", [("nonlinlap_periodic_p4", (13, 1, 1)), ("burgers_weno", (15, 8, 1)),
                                       ("brusselator", (15, 15, 1)), ("nonlindiff3d", (3, 4, 5))])
def test_generated_vjp_gather_on_ragged_tilings(name, tile):
    """Reverse-mode tile + gather kernels on tilings whose last tile is narrower than the halo
    (periodic images covered by two tiles; the gather's general cover path) and on tiles
    smaller than the halo: dot-product test against central differences of the forward kernel."""
    low, u, t, aux, ref, scale = CASES[name]()
    rng = np.random.default_rng(29)
    lam = rng.standard_normal(low.n_unknowns)
    v = rng.standard_normal(low.n_unknowns)
    out, g = genhost.vjp(low, u, lam, t=t, aux=aux, tile=tile)

    def val(e):
        return lam @ genhost.rhs(low, u + e * v, t=t, aux=aux, tile=None)

    eps = 1e-4 * max(1.0, np.max(np.abs(u)))
    d1 = (val(eps) - val(-eps)) / (2 * eps)
    d2 = (val(eps / 2) - val(-eps / 2)) / eps
    fd = (4 * d2 - d1) / 3
    assert abs(out @ v - fd) <= 1e-7 * abs(fd) + 1e-9 * np.linalg.norm(out) * np.linalg.norm(v), (name, tile)


def _rk4_numpy(F, u, t, dt):
    k1 = F(u, t)
    k2 = F(u + dt / 2 * k1, t + dt / 2)
    k3 = F(u + dt / 2 * k2, t + dt / 2)
    k4 = F(u + dt * k3, t + dt)
    return u + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)


@pytest.mark.parametrize("bcl,bch", [((1, 0, 0.0), (1, 0, 0.0)), ((0, 1, 0.5), (1, 0, 0.2)),
                                     ((1, 2, 0.1), (2, -1, 0.4))])
@pytest.mark.parametrize("tile", [(16, 1, 1), (64, 1, 1), None])
def test_fused_rk4_step_with_boundary_closures_1d(bcl, bch, tile):
    """Non-periodic fused step: ghost positions refilled from their closures after every stage
    (Dirichlet, Neumann, Robin ends) equal the oracle's RK4 step."""
    n = 101
    low = lower(*systems.heat1d(n=n, D=1.0, bcl=bcl, bch=bch))
    u = np.random.default_rng(1).uniform(-1, 1, low.n_unknowns)
    got, _ = genhost.rk4_fused(low, u, 0.0, 1e-5, tile=tile)
    ref = orc.diffusion1d_integrate(n, 1.0, bcl, bch, u, 0.0, 1e-5, 1)
    assert np.max(np.abs(got - ref)) <= 1e-14 * (1 + np.max(np.abs(ref)))


@pytest.mark.parametrize("bct", [(0, 0, 0, 0), (1, 1, 1, 1), (0, 1, 1, 0), (1, 0, 0, 1)])
@pytest.mark.parametrize("tile", [(8, 8, 1), (32, 16, 1)])
def test_fused_rk4_step_with_boundary_closures_2d(bct, tile):
    n = 41
    low = lower(*systems.diffreact2d(n, 0.05, 2.5, bct))
    u = np.random.default_rng(sum(bct)).uniform(0, 1, low.n_unknowns)
    dt = 2e-4
    got, _ = genhost.rk4_fused(low, u, 0.0, dt, tile=tile)
    ref = _rk4_numpy(lambda w, t: orc.diffreact2d_rhs(n, 0.05, 2.5, bct, w, t), u, 0.0, dt)
    assert np.max(np.abs(got - ref)) <= 1e-13 * (1 + np.max(np.abs(ref)))


def test_fused_rk4_step_with_time_dependent_boundary_values():
    """u(t, 0) = sin(5t), u_x(t, 1) = cos(3t): each stage's ghosts use the stage's time; checked
    against RK4 over the generated right-hand side evaluated at the stage times."""
    from molgpu.pdesys import Differential, Eq, Interval, PDESystem, parameters, variables
    from molgpu.discretization import MOLFiniteDifference
    import sympy as sp
    t, x = parameters("t x")
    u = variables("u")
    Dx = Differential(x)
    eq = Eq(Differential(t)(u(t, x)), 0.5 * (Dx ** 2)(u(t, x)) + sp.sin(x) * u(t, x))
    bcs = [Eq(u(0, x), x * (1 - x)), Eq(u(t, 0), sp.sin(5 * t)), Eq(Dx(u(t, 1)), sp.cos(3 * t))]
    low = lower(PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, 0.0, 1.0)], [t, x], [u(t, x)]),
                MOLFiniteDifference({x: 61}, t))
    u0 = np.random.default_rng(4).uniform(-1, 1, low.n_unknowns)
    t0, dt = 0.3, 1e-3
    got, tn = genhost.rk4_fused(low, u0, t0, dt, tile=(16, 1, 1))
    ref = _rk4_numpy(lambda w, tt: genhost.rhs(low, w, t=tt), u0, t0, dt)
    assert tn == pytest.approx(t0 + dt, abs=1e-15)
    assert np.max(np.abs(got - ref)) <= 1e-13 * (1 + np.max(np.abs(ref)))


def _td_heat1d(n):
    from molgpu.pdesys import Differential, Eq, Interval, PDESystem, parameters, variables
    from molgpu.discretization import MOLFiniteDifference
    import sympy as sp
    t, x = parameters("t x")
    u = variables("u")
    Dx = Differential(x)
    eq = Eq(Differential(t)(u(t, x)), 0.5 * (Dx ** 2)(u(t, x)) + sp.sin(x) * u(t, x))
    bcs = [Eq(u(0, x), x * (1 - x)), Eq(u(t, 0), sp.sin(5 * t)), Eq(Dx(u(t, 1)), sp.cos(3 * t))]
    return lower(PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, 0.0, 1.0)], [t, x], [u(t, x)]),
                 MOLFiniteDifference({x: n}, t))


@pytest.mark.parametrize("case", ["bruss33", "bruss37_16x16", "bruss29_8x8", "diffreact_0000", "diffreact_1111",
                                  "diffreact_0110", "heat_td"])
def test_fused_two_steps_per_launch(case, monkeypatch):
    """Two RK4 steps per fused launch (8-reach halo; the first step's result, with its running
    sums for the frame positions in LDS and its ghosts refilled at t + dt, is the second step's
    tile + halo): equal to two steps of RK4 over the right-hand side, clock advanced twice."""
    monkeypatch.setenv("MOL_FUSED_STEPS", "2")
    tile = None
    if case.startswith("bruss"):
        n = int(case[5:7])
        tile = {"bruss33": None, "bruss37_16x16": (16, 16, 1), "bruss29_8x8": (8, 8, 1)}[case]
        low = lower(*systems.brusselator2d(n=n))
        t0, dt = 1.1 - 1.5e-5, 1e-5   # the forcing switches on inside the second step
        F = lambda w, tt: orc.bruss2d_rhs(n, w, tt)
        u = np.random.default_rng(n).uniform(0.5, 1.5, low.n_unknowns)
    elif case.startswith("diffreact"):
        bct = tuple(int(c) for c in case[-4:])
        n = 41
        low = lower(*systems.diffreact2d(n, 0.05, 2.5, bct))
        t0, dt, tile = 0.0, 2e-4, (16, 8, 1)
        F = lambda w, tt: orc.diffreact2d_rhs(n, 0.05, 2.5, bct, w, tt)
        u = np.random.default_rng(sum(bct)).uniform(0, 1, low.n_unknowns)
    else:
        low = _td_heat1d(61)
        t0, dt, tile = 0.3, 1e-3, (16, 1, 1)
        F = lambda w, tt: genhost.rhs(low, w, t=tt)
        u = np.random.default_rng(4).uniform(-1, 1, low.n_unknowns)
    src = __import__("molgpu").codegen(low, tile=tile, fused=True)
    assert "// resident steps per launch: 2" in src
    got, tn = genhost.rk4_fused(low, u, t0, dt, tile=tile, nsub=2)
    ref = _rk4_numpy(F, _rk4_numpy(F, u, t0, dt), t0 + dt, dt)
    assert tn == (t0 + dt) + dt
    assert np.max(np.abs(got - ref)) <= 1e-13 * (1 + np.max(np.abs(ref))), case
    # the same kernel takes a single step when asked for one
    one, tn1 = genhost.rk4_fused(low, u, t0, dt, tile=tile, nsub=1)
    assert tn1 == t0 + dt
    ref1 = _rk4_numpy(F, u, t0, dt)
    assert np.max(np.abs(one - ref1)) <= 1e-13 * (1 + np.max(np.abs(ref1))), case


@pytest.mark.parametrize("name", ["brusselator_p4", "nonlinlap_periodic"])
def test_fused_rk4_step_reach_two_periodic(name):
    """Periodic systems of reach 2 (4th-order Laplacian, nonlinear Laplacian) take the fused step
    with an 8-node halo; equal to RK4 over the generated right-hand side."""
    if name == "brusselator_p4":
        low = lower(*systems.brusselator2d(n=65, approx_order=4))
    else:
        low = lower(*systems.nonlinlap1d(201, True, 0, 0, 2))
    u = np.asarray(low.u0, dtype=np.float64) + 0.01 * np.random.default_rng(1).standard_normal(low.n_unknowns)
    got, _ = genhost.rk4_fused(low, u, 1.2, 1e-6)
    ref = _rk4_numpy(lambda w, tt: genhost.rhs(low, w, t=tt), u, 1.2, 1e-6)
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))


@pytest.mark.parametrize("bcl,bch", [((1, 0, 0.0), (1, 0, 0.0)), ((0, 1, 0.5), (2, -1, 0.4))])
def test_fused_resident_steps_1d(bcl, bch):
    """A 1D system held by one tile runs several RK4 steps per launch with the state resident
    (ghosts refilled between steps): equal to the oracle's steps, and the clock equals the
    sequence of single-step additions."""
    n = 101
    low = lower(*systems.heat1d(n=n, D=1.0, bcl=bcl, bch=bch))
    import molgpu
    assert "// resident steps per launch: any" in molgpu.codegen(low, fused=True)
    u = np.random.default_rng(1).uniform(-1, 1, low.n_unknowns)
    got, tn = genhost.rk4_fused(low, u, 0.3, 1e-5, nsub=7)
    ref = orc.diffusion1d_integrate(n, 1.0, bcl, bch, u, 0.3, 1e-5, 7)
    tref = 0.3
    for _ in range(7):
        tref = tref + 1e-5
    assert tn == tref
    assert np.max(np.abs(got - ref)) <= 1e-14 * (1 + np.max(np.abs(ref)))


def test_fused_resident_steps_periodic():
    low = lower(*systems.nonlinlap1d(101, True, 0, 0, 2))
    u = np.asarray(low.u0, dtype=np.float64)
    ref, t = u.copy(), 0.0
    for _ in range(5):
        ref = _rk4_numpy(lambda w, tt: genhost.rhs(low, w, t=tt), ref, t, 1e-5)
        t += 1e-5
    got, _ = genhost.rk4_fused(low, u, 0.0, 1e-5, nsub=5)
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_fused_resident_steps_with_time_dependent_boundary_values():
    from molgpu.pdesys import Differential, Eq, Interval, PDESystem, parameters, variables
    from molgpu.discretization import MOLFiniteDifference
    import sympy as sp
    t, x = parameters("t x")
    u = variables("u")
    Dx = Differential(x)
    eq = Eq(Differential(t)(u(t, x)), 0.5 * (Dx ** 2)(u(t, x)) + sp.sin(x) * u(t, x))
    bcs = [Eq(u(0, x), x * (1 - x)), Eq(u(t, 0), sp.sin(5 * t)), Eq(Dx(u(t, 1)), sp.cos(3 * t))]
    low = lower(PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, 0.0, 1.0)], [t, x], [u(t, x)]),
                MOLFiniteDifference({x: 61}, t))
    u0 = np.random.default_rng(4).uniform(-1, 1, low.n_unknowns)
    t0, dt = 0.3, 1e-3
    ref, tt = u0.copy(), t0
    for _ in range(3):
        ref = _rk4_numpy(lambda w, s: genhost.rhs(low, w, t=s), ref, tt, dt)
        tt = tt + dt
    got, tn = genhost.rk4_fused(low, u0, t0, dt, nsub=3)
    assert tn == tt
    assert np.max(np.abs(got - ref)) <= 1e-13 * (1 + np.max(np.abs(ref)))


def test_operations_without_consumers_do_not_widen_the_halo():
    """The outer boundary row of the 3D nonlinear Laplacian has a zero third weight
    (half_offset_centred_difference.jl boundary rows); the interpolation it would weight is not
    generated, so the boundary tiles stage the interior reach (2), not 3."""
    import re
    import molgpu
    low = lower(*systems.nonlindiff3d(n=24))
    assert low.halo == [3, 3, 3]                 # the lowering's conservative span of the rows
    for vjp in (False, True):
        src = molgpu.codegen(low, tile=(8, 4, 4), vjp=vjp)
        assert set(re.findall(r"static constexpr int H[XYZ] = (\d+);", src)) == {"2"}, vjp


def test_reverse_gather_stores_with_every_variable_staged():
    """With every variable staged the gather stores each unknown (ghost adjoints through the
    transposed closures in the gather, no atomics on the output)."""
    import molgpu
    src = molgpu.codegen(lower(*systems.nonlindiff3d(n=12)), tile=(8, 4, 4), vjp=True)
    assert "// adjoint gather stores every unknown" in src
    assert "va.out[i] = (T)s;" in src and "va.out[i] +=" not in src
    body = src[src.index("mol_vjp_gather(VArgs va)"):]
    assert "T(0x1.5555555555556p+0) * mol_gather_pos" in body   # Neumann ghost u1 = 4/3 u2 - 1/3 u3


def test_reverse_mode_with_a_variable_read_only_at_its_node():
    """v of the diffusion + ODE system is not staged: its adjoint goes straight to the output,
    so the gather accumulates (output cleared first) instead of storing."""
    import molgpu
    low = lower(*systems.diffusion_ode1d(41))
    assert "// adjoint gather stores every unknown" not in molgpu.codegen(low, vjp=True)
    rng = np.random.default_rng(31)
    u = rng.uniform(0.2, 1.2, low.n_unknowns)
    lam = rng.standard_normal(low.n_unknowns)
    v = rng.standard_normal(low.n_unknowns)
    for tile in [(16, 1, 1), None]:
        out, _ = genhost.vjp(low, u, lam, tile=tile)

        def val(e):
            return lam @ genhost.rhs(low, u + e * v, tile=tile)

        d1 = (val(1e-4) - val(-1e-4)) / 2e-4
        d2 = (val(5e-5) - val(-5e-5)) / 1e-4
        fd = (4 * d2 - d1) / 3
        assert abs(out @ v - fd) <= 1e-7 * abs(fd) + 1e-9 * np.linalg.norm(out) * np.linalg.norm(v), tile
