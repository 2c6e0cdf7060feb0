"""PDE systems written with the public API: the benchmark workloads and the test systems.

They play the role of the reference's benchmark systems (benchmark/weno/suite.jl) and of the
systems in its test suite (test/Brusselator/brusselator_eq.jl, test/Diffusion/..., ...).
"""
from __future__ import annotations

import numpy as np
import sympy as sp

from .discretization import MOLFiniteDifference, UpwindScheme, WENOScheme
from .pdesys import Differential, Eq, Interval, PDESystem, grid_field, ifelse, parameters, variables


def heat1d(n: int = 101, D: float = 1.0, bcl=(1.0, 0.0, 0.0), bch=(1.0, 0.0, 0.0), approx_order: int = 2,
           dtype: str = "f64", grid_align: str = "center"):
    """u_t = D u_xx on [0, 1]; boundary a0 u + a1 u_x = g at each end; u(0, x) = sin(pi x).
    n: nodes of a centre-aligned grid, cells of an edge-aligned one (grid_align="edge")."""
    t, x = parameters("t x")
    u = variables("u")
    Dt, Dx = Differential(t), Differential(x)
    eq = Eq(Dt(u(t, x)), D * (Dx ** 2)(u(t, x)))

    def bc(side, coefs):
        a0, a1, g = coefs
        xv = 0 if side == 0 else 1
        return Eq(a0 * u(t, xv) + a1 * Dx(u(t, xv)), g)

    bcs = [Eq(u(0, x), sp.sin(sp.pi * x)), bc(0, bcl), bc(1, bch)]
    sys_ = PDESystem([eq], bcs, [Interval(t, 0.0, 0.1), Interval(x, 0.0, 1.0)], [t, x], [u(t, x)])
    return sys_, MOLFiniteDifference({x: n}, t, approx_order=approx_order, dtype=dtype, grid_align=grid_align)


def brusselator2d(n: int = 513, alpha: float = 10.0, approx_order: int = 2, dtype: str = "f64"):
    """2D Brusselator with periodic BCs (test/Brusselator/brusselator_eq.jl:8-52), n nodes per axis."""
    x, y, t = parameters("x y t")
    u, v = variables("u v")
    Dt, Dx, Dy = Differential(t), Differential(x), Differential(y)
    lap = lambda w: (Dx ** 2)(w) + (Dy ** 2)(w)
    f = ifelse((x - 0.3) ** 2 + (y - 0.6) ** 2 <= 0.1 ** 2, 1.0, 0.0) * ifelse(t >= 1.1, 1.0, 0.0) * 5.0
    U, V = u(x, y, t), v(x, y, t)
    eqs = [Eq(Dt(U), 1.0 + V * U ** 2 - 4.4 * U + alpha * lap(U) + f),
           Eq(Dt(V), 3.4 * U - V * U ** 2 + alpha * lap(V))]
    bcs = [Eq(u(x, y, 0), 22 * (y * (1 - y)) ** 1.5), Eq(u(0, y, t), u(1, y, t)), Eq(u(x, 0, t), u(x, 1, t)),
           Eq(v(x, y, 0), 27 * (x * (1 - x)) ** 1.5), Eq(v(0, y, t), v(1, y, t)), Eq(v(x, 0, t), v(x, 1, t))]
    sys_ = PDESystem(eqs, bcs, [Interval(x, 0.0, 1.0), Interval(y, 0.0, 1.0), Interval(t, 0.0, 11.5)],
                     [x, y, t], [U, V])
    return sys_, MOLFiniteDifference({x: n, y: n}, t, approx_order=approx_order, dtype=dtype)


def gaussian_bumps(seed: int, count: int, cmin=0.0, cmax=1.0, wmin=0.02, wmax=0.08, amin=0.1, amax=0.5):
    rng = np.random.default_rng(seed)
    c = rng.uniform(cmin, cmax, (count, 2))
    w = rng.uniform(wmin, wmax, count)
    a = rng.uniform(amin, amax, count)
    return c, w, a


def burgers2d(n: int = 1025, eps: float = 1e-6, seed: int = 0, dtype: str = "f64"):
    """u_t + u u_x + u u_y = 0 with WENO5 on periodic [0,1]^2; u0 = 0.5 + 8 seeded Gaussian bumps
    (nearest periodic image)."""
    t, x, y = parameters("t x y")
    u = variables("u")
    Dt, Dx, Dy = Differential(t), Differential(x), Differential(y)
    U = u(t, x, y)
    eq = Eq(Dt(U), -U * Dx(U) - U * Dy(U))
    c, w, a = gaussian_bumps(seed, 8)
    ic = 0.5
    for k in range(8):
        dxp = sp.Min(sp.Abs(x - c[k, 0]), 1 - sp.Abs(x - c[k, 0]))
        dyp = sp.Min(sp.Abs(y - c[k, 1]), 1 - sp.Abs(y - c[k, 1]))
        ic = ic + a[k] * sp.exp(-(dxp ** 2 + dyp ** 2) / (2 * w[k] ** 2))
    bcs = [Eq(u(0, x, y), ic), Eq(u(t, 0, y), u(t, 1, y)), Eq(u(t, x, 0), u(t, x, 1))]
    sys_ = PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, 0.0, 1.0), Interval(y, 0.0, 1.0)],
                     [t, x, y], [U])
    return sys_, MOLFiniteDifference({x: n, y: n}, t, advection_scheme=WENOScheme(eps), dtype=dtype)


def burgers2d_nonuniform(n: int = 1025, seed: int = 0, spread: float = 0.3, eps: float = 1e-6,
                         dtype: str = "f64"):
    """The WENO5 Burgers system of `burgers2d` on seeded stretched node arrays along x and y
    (n nodes each, periodic)."""
    sys_, disc = burgers2d(n=n, eps=eps, seed=seed, dtype=dtype)
    x, y = sys_.ivs[1], sys_.ivs[2]
    nodes = {x: stretched_grid(n, 1.0, seed + 1, spread), y: stretched_grid(n, 1.0, seed + 2, spread)}
    return sys_, MOLFiniteDifference(nodes, disc.time, advection_scheme=WENOScheme(eps), dtype=dtype)


def diffreact2d(n: int, D: float, k: float, bctype=(0, 0, 0, 0), dtype: str = "f64"):
    """u_t = D Lap u + k u (1 - u) on [0,1]^2; bctype per side (x-lo, x-hi, y-lo, y-hi):
    0 Dirichlet u = 0, 1 Neumann du/dn = 0."""
    t, x, y = parameters("t x y")
    u = variables("u")
    Dt, Dx, Dy = Differential(t), Differential(x), Differential(y)
    U = u(t, x, y)
    Dp, kp = sp.Symbol("D", real=True), sp.Symbol("k", real=True)
    eq = Eq(Dt(U), Dp * ((Dx ** 2)(U) + (Dy ** 2)(U)) + kp * U * (1 - U))
    bcs = [Eq(u(0, x, y), sp.sin(sp.pi * x) * sp.sin(sp.pi * y))]
    for (axis, side), kind in zip([(0, 0), (0, 1), (1, 0), (1, 1)], bctype):
        val = 0 if side == 0 else 1
        args = (t, val, y) if axis == 0 else (t, x, val)
        var = x if axis == 0 else y
        if kind == 0:
            bcs.append(Eq(u(*args), 0))
        else:
            bcs.append(Eq(Differential(var)(u(*args)), 0))
    sys_ = PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, 0.0, 1.0), Interval(y, 0.0, 1.0)],
                     [t, x, y], [U], {Dp: D, kp: k})
    return sys_, MOLFiniteDifference({x: n, y: n}, t, dtype=dtype)


def diffreact_grids(grids, D: float, k: float, bctype, approx_order: int = 2, dtype: str = "f64"):
    """u_t = D Lap u + k u (1 - u) on [0,1]^d, d = len(grids); grids[a]: a node count or the nodes
    (spanning [0, 1]); bctype per side (x-lo, x-hi, y-lo, ...): 0 Dirichlet u = 0, 1 Neumann 0."""
    names = "x y z"[:2 * len(grids) - 1]
    xs = list(parameters("t " + names))
    t, xs = xs[0], xs[1:]
    u = variables("u")
    Dt = Differential(t)
    U = u(t, *xs)
    eq = Eq(Dt(U), D * sum((Differential(x) ** 2)(U) for x in xs) + k * U * (1 - U))
    ic = 1
    for x in xs:
        ic = ic * sp.sin(sp.pi * x)
    bcs = [Eq(u(0, *xs), ic)]
    for a, x in enumerate(xs):
        for side in (0, 1):
            args = [t] + list(xs)
            args[a + 1] = side
            if bctype[2 * a + side] == 0:
                bcs.append(Eq(u(*args), 0))
            else:
                bcs.append(Eq(Differential(x)(u(*args)), 0))
    sys_ = PDESystem([eq], bcs, [Interval(t, 0.0, 1.0)] + [Interval(x, 0.0, 1.0) for x in xs], [t] + xs, [U])
    return sys_, MOLFiniteDifference({x: g for x, g in zip(xs, grids)}, t, approx_order=approx_order, dtype=dtype)


def diffusion_ode1d(n: int = 41, D: float = 0.3):
    """u_t = D u_xx - u v, v_t = u^2 - v/2 on [0, 1]: v has no spatial derivative, so every
    node of v is an unknown and v is read only at its node (u(t,0) = 0, u_x(t,1) = 0)."""
    t, x = parameters("t x")
    u, v = variables("u v")
    Dx = Differential(x)
    U, V = u(t, x), v(t, x)
    eqs = [Eq(Differential(t)(U), D * (Dx ** 2)(U) - U * V), Eq(Differential(t)(V), U * U - 0.5 * V)]
    bcs = [Eq(u(0, x), x * (1 - x)), Eq(v(0, x), 1 + x), Eq(u(t, 0), 0), Eq(Dx(u(t, 1)), 0)]
    return (PDESystem(eqs, bcs, [Interval(t, 0.0, 1.0), Interval(x, 0.0, 1.0)], [t, x], [U, V]),
            MOLFiniteDifference({x: n}, t))


def nonlindiff3d(n: int = 128, approx_order: int = 2, dtype: str = "f64"):
    """u_t = div(d0 (1 + u^2) grad u) + s on [0,1]^3 with Neumann BCs; d0 and s are grid fields."""
    t, x, y, z = parameters("t x y z")
    u = variables("u")
    d0 = grid_field("d0")
    s = grid_field("s")
    Dt, Dx, Dy, Dz = Differential(t), Differential(x), Differential(y), Differential(z)
    U = u(t, x, y, z)
    a = d0(x, y, z) * (1 + U ** 2)
    eq = Eq(Dt(U), Dx(a * Dx(U)) + Dy(a * Dy(U)) + Dz(a * Dz(U)) + s(x, y, z))
    bcs = [Eq(u(0, x, y, z), 0)]
    for var in (x, y, z):
        for val in (0, 1):
            args = [t, x, y, z]
            args[[x, y, z].index(var) + 1] = val
            bcs.append(Eq(Differential(var)(u(*args)), 0))
    sys_ = PDESystem([eq], bcs, [Interval(t, 0.0, 1.0)] + [Interval(v, 0.0, 1.0) for v in (x, y, z)],
                     [t, x, y, z], [U])
    return sys_, MOLFiniteDifference({x: n, y: n, z: n}, t, approx_order=approx_order, dtype=dtype)


def gaussian_field3d(n: int, seed: int, count: int = 16):
    """Seeded sum of `count` Gaussians on the n^3 grid (full grid, array indexed [x, y, z])."""
    rng = np.random.default_rng(seed)
    g = np.arange(n) / (n - 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    out = np.zeros((n, n, n))
    for _ in range(count):
        c = rng.uniform(0, 1, 3)
        w = rng.uniform(0.05, 0.2)
        a = rng.uniform(-1, 1)
        out += a * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * w * w))
    return out


def upwind1d(n: int, c0: float, c1: float, gl: float, gh: float, order: int = 1):
    t, x = parameters("t x")
    u = variables("u")
    Dt, Dx = Differential(t), Differential(x)
    eq = Eq(Dt(u(t, x)), -(c0 + c1 * sp.sin(2 * sp.pi * x)) * Dx(u(t, x)))
    bcs = [Eq(u(0, x), sp.cos(x)), Eq(u(t, 0), gl), Eq(u(t, 1), gh)]
    sys_ = PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, 0.0, 1.0)], [t, x], [u(t, x)])
    return sys_, MOLFiniteDifference({x: n}, t, advection_scheme=UpwindScheme(order))


def weno1d_dirichlet(n: int, L: float, c: float, gl: float, gh: float, eps: float = 1e-6):
    t, x = parameters("t x")
    u = variables("u")
    Dt, Dx = Differential(t), Differential(x)
    eq = Eq(Dt(u(t, x)), -c * Dx(u(t, x)))
    bcs = [Eq(u(0, x), sp.exp(-(x - L / 2) ** 2)), Eq(u(t, 0), gl), Eq(u(t, L), gh)]
    sys_ = PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, 0.0, L)], [t, x], [u(t, x)])
    return sys_, MOLFiniteDifference({x: n}, t, advection_scheme=WENOScheme(eps))


def stretched_grid(n: int, L: float, seed: int = 0, spread: float = 0.5):
    """n strictly increasing nodes on [0, L]: uniform spacing perturbed by seeded factors in
    [1 - spread, 1 + spread] and rescaled to end at L."""
    rng = np.random.default_rng(seed)
    h = 1.0 + spread * rng.uniform(-1.0, 1.0, n - 1)
    x = np.concatenate([[0.0], np.cumsum(h)])
    x = x * (L / x[-1])
    x[-1] = L
    return x


def weno1d_nonuniform(x_nodes, c: float, gl: float = 0.0, gh: float = 0.0, periodic: bool = False,
                      eps: float = 1e-6):
    """u_t = -c u_x with WENO5 on the given (nonuniform) nodes: Dirichlet u = gl, gh at the ends,
    or periodic; u0 = a Gaussian bump."""
    t, x = parameters("t x")
    u = variables("u")
    Dt, Dx = Differential(t), Differential(x)
    x_nodes = np.asarray(x_nodes, dtype=np.float64)
    L0, L1 = float(x_nodes[0]), float(x_nodes[-1])
    eq = Eq(Dt(u(t, x)), -c * Dx(u(t, x)))
    ic = sp.exp(-((x - (L0 + L1) / 2) / (0.2 * (L1 - L0))) ** 2)
    if periodic:
        bcs = [Eq(u(0, x), ic), Eq(u(t, L0), u(t, L1))]
    else:
        bcs = [Eq(u(0, x), ic), Eq(u(t, L0), gl), Eq(u(t, L1), gh)]
    sys_ = PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, L0, L1)], [t, x], [u(t, x)])
    return sys_, MOLFiniteDifference({x: x_nodes}, t, advection_scheme=WENOScheme(eps))


def nonuniform1d(x_nodes, D: float, c: float, approx_order: int = 2, upwind_order: int = 1,
                 bcl=(1.0, 0.0, 0.0), bch=(1.0, 0.0, 0.0)):
    """u_t = D u_xx - c u_x on the given (nonuniform) nodes; a0 u + a1 u_x = g at each end;
    u0 = a Gaussian bump."""
    t, x = parameters("t x")
    u = variables("u")
    Dt, Dx = Differential(t), Differential(x)
    x_nodes = np.asarray(x_nodes, dtype=np.float64)
    L0, L1 = float(x_nodes[0]), float(x_nodes[-1])
    rhs = D * (Dx ** 2)(u(t, x)) if D != 0 else 0
    if c != 0:
        rhs = rhs - c * Dx(u(t, x))
    eq = Eq(Dt(u(t, x)), rhs)

    def bc(xv, coefs):
        a0, a1, g = coefs
        return Eq(a0 * u(t, xv) + a1 * Dx(u(t, xv)), g)

    ic = sp.exp(-((x - (L0 + L1) / 2) / (0.2 * (L1 - L0))) ** 2)
    bcs = [Eq(u(0, x), ic), bc(L0, bcl), bc(L1, bch)]
    sys_ = PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, L0, L1)], [t, x], [u(t, x)])
    return sys_, MOLFiniteDifference({x: x_nodes}, t, approx_order=approx_order,
                                     advection_scheme=UpwindScheme(upwind_order))


def mixed2d(n, c: float, approx_order: int = 2, q=(0.0, 1.0, 2.0, 0.0, 0.0, 0.0)):
    """u_t = u_xx + u_yy + c u_xy on [0,1]^2, Dirichlet u = q0 + q1 x + q2 y + q3 x^2 + q4 x y +
    q5 y^2 on all sides.  n: nodes per axis, or a pair (per axis: a node count or the nodes, both
    ends 0 and 1 included)."""
    t, x, y = parameters("t x y")
    u = variables("u")
    Dt, Dx, Dy = Differential(t), Differential(x), Differential(y)
    U = u(t, x, y)
    eq = Eq(Dt(U), (Dx ** 2)(U) + (Dy ** 2)(U) + c * Dx(Dy(U)))
    q = [sp.nsimplify(v) if float(v) == int(v) else float(v) for v in q]
    g = lambda xx, yy: q[0] + q[1] * xx + q[2] * yy + q[3] * xx ** 2 + q[4] * xx * yy + q[5] * yy ** 2
    bcs = [Eq(u(0, x, y), 0), Eq(u(t, 0, y), g(0, y)), Eq(u(t, 1, y), g(1, y)),
           Eq(u(t, x, 0), g(x, 0)), Eq(u(t, x, 1), g(x, 1))]
    sys_ = PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, 0.0, 1.0), Interval(y, 0.0, 1.0)],
                     [t, x, y], [U])
    nx, ny = (n, n) if np.isscalar(n) else n
    return sys_, MOLFiniteDifference({x: nx, y: ny}, t, approx_order=approx_order)


def spherical1d(n, g1: float, approx_order: int = 2):
    """u_t = r^-2 d/dr(r^2 u_r) on [0, 1], u_r(0) = 0, u(1) = g1; n nodes, or the nodes themselves."""
    t, r = parameters("t r")
    u = variables("u")
    Dt, Dr = Differential(t), Differential(r)
    eq = Eq(Dt(u(t, r)), 1 / r ** 2 * Dr(r ** 2 * Dr(u(t, r))))
    bcs = [Eq(u(0, r), sp.cos(r)), Eq(Dr(u(t, 0)), 0), Eq(u(t, 1), g1)]
    sys_ = PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(r, 0.0, 1.0)], [t, r], [u(t, r)])
    return sys_, MOLFiniteDifference({r: n}, t, approx_order=approx_order)


def nonlinlap1d(n, periodic: bool, gl: float = 0.0, gh: float = 0.0, approx_order: int = 2):
    """u_t = d/dx((1 + u^2) u_x) on [0, 1]; n nodes, or the nodes themselves (spanning [0, 1])."""
    t, x = parameters("t x")
    u = variables("u")
    Dt, Dx = Differential(t), Differential(x)
    eq = Eq(Dt(u(t, x)), Dx((1 + u(t, x) ** 2) * Dx(u(t, x))))
    if periodic:
        bcs = [Eq(u(0, x), sp.sin(2 * sp.pi * x)), Eq(u(t, 0), u(t, 1))]
    else:
        bcs = [Eq(u(0, x), sp.sin(2 * sp.pi * x)), Eq(u(t, 0), gl), Eq(u(t, 1), gh)]
    sys_ = PDESystem([eq], bcs, [Interval(t, 0.0, 1.0), Interval(x, 0.0, 1.0)], [t, x], [u(t, x)])
    return sys_, MOLFiniteDifference({x: n}, t, approx_order=approx_order)
