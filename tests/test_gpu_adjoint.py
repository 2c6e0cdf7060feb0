"""GPU reverse mode: the generated vector-Jacobian product and the checkpointed RK4 adjoint.

Pins: (1) for a linear right-hand side the VJP is the exact transpose of the oracle's matrix;
(2) for nonlinear ones, <J^T lam, v> equals the central difference of <lam, f(u + e v)> built
from the CPU oracle (or, for slot kinds without an oracle system here, from the GPU forward
kernel); (3) the adjoint gradient of L = 1/2 |u(T) - target|^2 with respect to the per-cell
field d0 matches the oracle's forward-mode (dual-number) derivative through the oracle's own
RK4 to 1e-8, and Richardson-extrapolated central finite differences of the GPU solve on random
cells.
"""
import numpy as np
import pytest

import oracle as orc
import molgpu
from molgpu import systems

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _dev(x, dtype=None):
    torch = _torch()
    return torch.tensor(np.asarray(x), dtype=dtype or torch.float64, device="cuda")


def _nonlin(n, seed=1):
    sys_, disc = systems.nonlindiff3d(n=n)
    prob = molgpu.discretize(sys_, disc)
    d0 = 0.5 + 0.5 * np.abs(systems.gaussian_field3d(n, seed))
    src = systems.gaussian_field3d(n, seed + 1)
    prob.set_grid_field("d0", d0)
    prob.set_grid_field("s", src)
    return prob, d0.reshape(-1, order="F"), src.reshape(-1, order="F")


def _grad_of(prob, g, name):
    """Gradient with respect to grid field `name` out of the stacked FP64 gradient array."""
    ng = int(np.prod(prob.lowered.n))
    a = prob.lowered.aux.index(name)
    return g[a * ng:(a + 1) * ng].cpu().numpy()


def _on_box(prob, field):
    """A full-grid field restricted to the unknown box (the state layout)."""
    (l0, l1, l2), (h0, h1, h2) = prob.lowered.lo[0], prob.lowered.hi[0]
    return field[l0 - 1:h0, l1 - 1:h1, l2 - 1:h2].reshape(-1, order="F")


def test_vjp_is_exact_transpose_for_linear_heat_with_robin_boundaries():
    """f(u) = A u + b (Robin, 4th order): VJP(lam) = A^T lam, A from oracle RHS columns."""
    torch = _torch()
    n = 41
    bcl, bch = (2.0, 0.5, 0.3), (1.5, -0.25, -0.2)
    sys_, disc = systems.heat1d(n=n, D=0.8, bcl=bcl, bch=bch, approx_order=4)
    prob = molgpu.discretize(sys_, disc)
    prob.set_tile(16)                    # boundary and interior tiles both present
    N = prob.lowered.n_unknowns
    b = orc.diffusion1d_rhs(n, 0.8, bcl, bch, np.zeros(N), p=4)
    A = np.stack([orc.diffusion1d_rhs(n, 0.8, bcl, bch, np.eye(N)[j], p=4) - b for j in range(N)], axis=1)
    rng = np.random.default_rng(3)
    lam = rng.standard_normal(N)
    u = rng.standard_normal(N)
    print("  heat vjp n=%d" % n, flush=True)
    got = prob.vjp(_dev(u), _dev(lam)).cpu().numpy()
    torch.cuda.synchronize()
    ref = A.T @ lam
    scale = np.abs(A.T) @ np.abs(lam)
    assert np.max(np.abs(got - ref) / (scale + 1e-300)) < 1e-13


@pytest.mark.parametrize("n", [12, 20])
def test_vjp_dot_product_nonlinear_diffusion_vs_oracle(n):
    """<J^T lam, v> and <(df/dd0)^T lam, w> against central differences of the oracle RHS."""
    torch = _torch()
    prob, d0, src = _nonlin(n)
    N = prob.lowered.n_unknowns
    rng = np.random.default_rng(n)
    u = rng.uniform(-1, 1, N)
    lam = rng.standard_normal(N)
    v = rng.standard_normal(N)
    w = rng.standard_normal(n ** 3)
    g = torch.zeros(2 * n ** 3, dtype=torch.float64, device="cuda")
    print("  nonlin vjp n=%d" % n, flush=True)
    got = prob.vjp(_dev(u), _dev(lam), t=0.3, g_aux=g).cpu().numpy()
    gd0 = _grad_of(prob, g, "d0")
    gs = _grad_of(prob, g, "s")

    def F(uu, dd):
        return orc.nonlindiff3d_rhs(n, dd, src, uu, t=0.3, threads=8)

    def cdiff(fun, eps):
        return (fun(eps) - fun(-eps)) / (2 * eps)

    for eps_u in (1e-4,):
        du = cdiff(lambda e: lam @ F(u + e * v, d0), eps_u)
        du2 = cdiff(lambda e: lam @ F(u + e * v, d0), eps_u / 2)
        fd = (4 * du2 - du) / 3
        assert abs(got @ v - fd) <= 1e-8 * abs(fd) + 1e-10 * np.linalg.norm(got) * np.linalg.norm(v)
    dd = cdiff(lambda e: lam @ F(u, d0 + e * w), 1e-4)
    dd2 = cdiff(lambda e: lam @ F(u, d0 + e * w), 5e-5)
    fd = (4 * dd2 - dd) / 3
    assert abs(gd0 @ w - fd) <= 1e-8 * abs(fd) + 1e-10 * np.linalg.norm(gd0) * np.linalg.norm(w)
    # the source term enters linearly: df/ds = identity on the unknown nodes
    assert abs(gs.sum() - lam.sum()) <= 1e-10 * np.abs(lam).sum()


def _gpu_dot_fd(prob, u, lam, v, t=0.0, eps=1e-5):
    """Richardson-extrapolated central difference of <lam, f(u + e v)> with the GPU forward kernel."""
    torch = _torch()

    def val(e):
        du = torch.empty(len(u), dtype=torch.float64, device="cuda")
        prob.f(du, _dev(u + e * v), None, t)
        return float(lam @ du.cpu().numpy())

    d1 = (val(eps) - val(-eps)) / (2 * eps)
    d2 = (val(eps / 2) - val(-eps / 2)) / eps
    return (4 * d2 - d1) / 3


@pytest.mark.parametrize("name", ["brusselator", "burgers_weno", "mixed_p4", "spherical", "upwind_o2",
                                  "nonlinlap_periodic_p4", "weno_dirichlet", "diffreact_mixed_bcs",
                                  "weno_nonuniform", "weno_nonuniform_periodic", "nonuniform_centered_upwind",
                                  "nonuniform_nonlinlap_p4", "nonuniform_spherical", "node_only_variable"])
def test_vjp_dot_product_every_slot_kind(name):
    """Reverse mode of every operator-slot kind against the forward kernel's directional
    derivative (small tiles: boundary-row variants and ghost closures included)."""
    torch = _torch()
    mk = {
        "brusselator": (lambda: systems.brusselator2d(33), (0.5, 1.5), 1.2),
        "burgers_weno": (lambda: systems.burgers2d(33), (0.2, 1.2), 0.0),
        "mixed_p4": (lambda: systems.mixed2d(21, 0.7, 4), (-1, 1), 0.0),
        "spherical": (lambda: systems.spherical1d(41, 0.25, 2), (-1, 1), 0.0),
        "upwind_o2": (lambda: systems.upwind1d(61, 0.1, 1.0, 0.3, -0.2, 2), (-1, 1), 0.0),
        "nonlinlap_periodic_p4": (lambda: systems.nonlinlap1d(61, True, 0.2, -0.4, 4), (-1, 1), 0.0),
        "weno_dirichlet": (lambda: systems.weno1d_dirichlet(61, 1.0, 1.0, 0.1, 0.05), (-1, 1), 0.0),
        "diffreact_mixed_bcs": (lambda: systems.diffreact2d(25, 0.05, 2.5, (0, 1, 2, 0)), (0, 1), 0.0),
        "weno_nonuniform": (lambda: systems.weno1d_nonuniform(systems.stretched_grid(61, 1.0, 4), 1.0, 0.1, 0.05),
                            (-1, 1), 0.0),
        "weno_nonuniform_periodic": (lambda: systems.weno1d_nonuniform(systems.stretched_grid(61, 1.0, 4), 1.0,
                                                                       periodic=True), (-1, 1), 0.0),
        "nonuniform_centered_upwind": (lambda: systems.nonuniform1d(systems.stretched_grid(61, 1.0, 4), 0.3, 0.7, 4,
                                                                    1, (0, 1, 0.5), (1, 2, 0.1)), (-1, 1), 0.0),
        "nonuniform_nonlinlap_p4": (lambda: systems.nonlinlap1d(systems.stretched_grid(61, 1.0, 5), False, 0.2, -0.4,
                                                                4), (-1, 1), 0.0),
        "nonuniform_spherical": (lambda: systems.spherical1d(systems.stretched_grid(61, 1.0, 5), 0.25, 2), (-1, 1),
                                 0.0),
        # v is read only at its node: the output is cleared and the gather accumulates
        "node_only_variable": (lambda: systems.diffusion_ode1d(61), (0.2, 1.2), 0.0),
    }
    build, (lo, hi), t = mk[name]
    sys_, disc = build()
    prob = molgpu.discretize(sys_, disc)
    prob.set_tile(*{1: (16, 1, 1), 2: (8, 8, 1), 3: (8, 4, 4)}[prob.lowered.ndim])
    N = prob.lowered.n_unknowns
    rng = np.random.default_rng(7)
    u = rng.uniform(lo, hi, N)
    lam = rng.standard_normal(N)
    v = rng.standard_normal(N)
    print("  vjp dot %s" % name, flush=True)
    got = prob.vjp(_dev(u), _dev(lam), t=t).cpu().numpy()
    fd = _gpu_dot_fd(prob, u, lam, v, t)
    assert abs(got @ v - fd) <= 1e-7 * abs(fd) + 1e-9 * np.linalg.norm(got) * np.linalg.norm(v), (got @ v, fd)


def _loss_and_state(prob, u0, dt, nsteps, target):
    torch = _torch()
    u = _dev(u0)
    prob.integrate(u, 0.0, dt, nsteps, molgpu.RK4)
    r = u.cpu().numpy() - target
    return 0.5 * float(r @ r), u


def test_adjoint_gradient_matches_forward_mode_oracle_and_finite_differences():
    """dL/dd0 from the checkpointed discrete adjoint vs the oracle's dual-number forward mode
    (ForwardDiff counterpart) along a random direction, and vs central differences on cells."""
    torch = _torch()
    n, nsteps, every = 16, 24, 5
    prob, d0, src = _nonlin(n, seed=3)
    N = prob.lowered.n_unknowns
    rng = np.random.default_rng(11)
    u0 = 0.3 * _on_box(prob, systems.gaussian_field3d(n, 5))
    target = 0.2 * _on_box(prob, systems.gaussian_field3d(n, 6))
    dx = 1.0 / (n - 1)
    dt = 0.1 * dx * dx / (2.0 * np.max(d0))
    # forward with checkpoints, then the adjoint
    u = _dev(u0)
    print("  adjoint forward", flush=True)
    prob.forward_checkpointed(u, 0.0, dt, nsteps, every)
    uT = u.cpu().numpy()
    lam = _dev(uT - target)
    g = torch.zeros(2 * n ** 3, dtype=torch.float64, device="cuda")
    print("  adjoint backward", flush=True)
    prob.adjoint(lam, g)
    gd0 = _grad_of(prob, g, "d0")
    # the checkpointed forward is the plain forward
    L_gpu, u_plain = _loss_and_state(prob, u0, dt, nsteps, target)
    assert np.array_equal(u_plain.cpu().numpy(), uT)
    # oracle forward mode along a random direction
    vdir = rng.standard_normal(n ** 3)
    L_orc, dL = orc.nonlindiff3d_jvp(n, d0, vdir, src, u0, target, dt, nsteps)
    assert abs(L_gpu - L_orc) <= 1e-12 * abs(L_orc)
    assert abs(gd0 @ vdir - dL) <= 1e-8 * abs(dL), (gd0 @ vdir, dL)
    # central differences of the GPU solve on random cells (Richardson extrapolated)
    cells = rng.choice(n ** 3, size=6, replace=False)
    d0_dev = _dev(d0)
    prob.set_grid_field("d0", d0_dev)
    for c in cells:
        def L_at(e):
            d0_dev.copy_(_dev(d0))
            d0_dev[c] += e
            return _loss_and_state(prob, u0, dt, nsteps, target)[0]
        h = 1e-3 * d0[c]
        D1 = (L_at(h) - L_at(-h)) / (2 * h)
        D2 = (L_at(h / 2) - L_at(-h / 2)) / h
        fd = (4 * D2 - D1) / 3
        assert abs(gd0[c] - fd) <= 1e-6 * abs(fd) + 1e-9 * np.max(np.abs(gd0)), (c, gd0[c], fd)
    d0_dev.copy_(_dev(d0))


def test_autograd_function_returns_the_adjoint_gradients():
    """molgpu.autograd.rk4_solve: gradients w.r.t. u0 and d0 equal the direct adjoint calls."""
    torch = _torch()
    from molgpu.autograd import rk4_solve
    n, nsteps, every = 12, 10, 4
    prob, d0, src = _nonlin(n, seed=4)
    dt = 0.1 * (1.0 / (n - 1)) ** 2 / (2.0 * np.max(d0))
    u0 = _dev(0.3 * _on_box(prob, systems.gaussian_field3d(n, 7))).requires_grad_(True)
    d0_t = _dev(d0).requires_grad_(True)
    target = _dev(0.1 * _on_box(prob, systems.gaussian_field3d(n, 8)))
    print("  autograd solve", flush=True)
    uT = rk4_solve(prob, u0, 0.0, dt, nsteps, every, {"d0": d0_t})
    loss = 0.5 * ((uT - target) ** 2).sum()
    loss.backward()
    # direct calls
    u = u0.detach().clone()
    prob.forward_checkpointed(u, 0.0, dt, nsteps, every)
    lam = (u - target).contiguous()
    g = torch.zeros(2 * n ** 3, dtype=torch.float64, device="cuda")
    prob.adjoint(lam, g)
    # the reverse sweep sums tap adjoints with atomics: equal up to the order of summation
    assert torch.allclose(u0.grad, lam, rtol=1e-12, atol=1e-14 * float(lam.abs().max()))
    a = prob.lowered.aux.index("d0")
    gd = g[a * n ** 3:(a + 1) * n ** 3]
    assert torch.allclose(d0_t.grad, gd, rtol=1e-12, atol=1e-14 * float(g.abs().max()))


def test_adjoint_with_stored_stages_equals_checkpointed_adjoint():
    """ckpt_every=0 keeps every stage state in the forward run: the same forward state, and the
    adjoint (no recomputation) equal to the checkpointed adjoint up to the order of the atomic
    sums; the automatic choice picks stored stages for a small run."""
    torch = _torch()
    n, nsteps = 12, 9
    prob, d0, src = _nonlin(n, seed=5)
    dt = 0.1 * (1.0 / (n - 1)) ** 2 / (2.0 * np.max(d0))
    u0 = 0.3 * _on_box(prob, systems.gaussian_field3d(n, 9))
    target = _dev(0.1 * _on_box(prob, systems.gaussian_field3d(n, 10)))
    res = {}
    for every in (4, 0):
        print("  stored-stage adjoint, ckpt_every=%d" % every, flush=True)
        u = _dev(u0)
        prob.forward_checkpointed(u, 0.0, dt, nsteps, every)
        lam = (u - target).contiguous()
        g = torch.zeros(2 * n ** 3, dtype=torch.float64, device="cuda")
        prob.adjoint(lam, g)
        res[every] = (u.clone(), lam.clone(), g.clone())
    assert torch.equal(res[0][0], res[4][0])
    for k in (1, 2):
        a, b = res[0][k], res[4][k]
        assert torch.allclose(a, b, rtol=1e-12, atol=1e-14 * float(b.abs().max())), k
    assert prob.checkpoint_interval(nsteps) == 0


def test_adjoint_with_a_variable_read_only_at_its_node():
    """The discrete RK4 adjoint of the diffusion + ODE system (accumulating gather: stage
    adjoints cleared per step) against Richardson-extrapolated central differences of the GPU
    solve along a random direction of the initial state."""
    sys_, disc = systems.diffusion_ode1d(41)
    prob = molgpu.discretize(sys_, disc)
    N = prob.lowered.n_unknowns
    rng = np.random.default_rng(11)
    u0 = rng.uniform(0.2, 1.2, N)
    target = rng.uniform(0.0, 1.0, N)
    w = rng.standard_normal(N)
    dt, nsteps = 1e-4, 40
    print("  adjoint diffusion_ode1d", flush=True)
    u = _dev(u0)
    prob.forward_checkpointed(u, 0.0, dt, nsteps, 0)
    lam = (u - _dev(target)).contiguous()
    prob.adjoint(lam)
    grad = lam.cpu().numpy()

    def loss(e):
        return _loss_and_state(prob, u0 + e * w, dt, nsteps, target)[0]

    d1 = (loss(1e-4) - loss(-1e-4)) / 2e-4
    d2 = (loss(5e-5) - loss(-5e-5)) / 1e-4
    fd = (4 * d2 - d1) / 3
    assert abs(grad @ w - fd) <= 1e-7 * abs(fd) + 1e-10 * np.linalg.norm(grad) * np.linalg.norm(w), (grad @ w, fd)


# ---------------------------------------------------------------- config 5 in FP32
def _nonlin32(n, seed=1):
    """The 3D nonlinear diffusion in FP32 with FP32-representable grid fields (so that the FP64
    oracle sees the same inputs and the difference is FP32 arithmetic only)."""
    sys_, disc = systems.nonlindiff3d(n=n, dtype="f32")
    prob = molgpu.discretize(sys_, disc)
    d0 = (0.5 + 0.5 * np.abs(systems.gaussian_field3d(n, seed))).astype(np.float32).astype(np.float64)
    src = systems.gaussian_field3d(n, seed + 1).astype(np.float32).astype(np.float64)
    prob.set_grid_field("d0", d0)
    prob.set_grid_field("s", src)
    return prob, d0.reshape(-1, order="F"), src.reshape(-1, order="F")


def test_fp32_nonlinear_diffusion_rhs_and_rk4_vs_fp64_oracle():
    """FP32 state and arithmetic at 34^3: one RHS and 20 RK4 steps against the FP64 oracle, with
    the FP32 error stated against the magnitude of the summed terms."""
    torch = _torch()
    n = 34
    prob, d0, src = _nonlin32(n)
    N = prob.lowered.n_unknowns
    u = np.random.default_rng(2).uniform(-1, 1, N).astype(np.float32).astype(np.float64)
    ug = _dev(u, torch.float32)
    du = torch.empty_like(ug)
    print("  fp32 nonlin rhs n=%d" % n, flush=True)
    prob.f(du, ug, None, 0.0)
    torch.cuda.synchronize()
    got = du.cpu().numpy().astype(np.float64)
    ref = orc.nonlindiff3d_rhs(n, d0, src, u, threads=8)
    dx = 1.0 / (n - 1)
    # each of the 3 axes sums 2 half-node fluxes of size a * |du| / dx^2, a <= d0 (1 + u^2)
    scale = 3 * 2 * 2 * np.max(d0) * 2.0 * np.max(np.abs(u)) / dx ** 2 + np.max(np.abs(src))
    err = np.max(np.abs(got - ref))
    print("  fp32 rhs max error %.3e (scale %.3e)" % (err, scale))
    assert err <= 1e-6 * scale
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-5
    # 20 RK4 steps against RK4 over the FP64 oracle
    dt = 0.1 * dx * dx / (2.0 * np.max(d0))
    steps = 20
    print("  fp32 nonlin rk4 %d steps" % steps, flush=True)
    prob.integrate(ug, 0.0, dt, steps, molgpu.RK4, graph=True)
    w = u.copy()
    F = lambda x, t: orc.nonlindiff3d_rhs(n, d0, src, x, t=t, threads=8)
    for k in range(steps):
        t = k * dt
        k1 = F(w, t); k2 = F(w + dt / 2 * k1, t + dt / 2); k3 = F(w + dt / 2 * k2, t + dt / 2)
        k4 = F(w + dt * k3, t + dt)
        w = w + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    torch.cuda.synchronize()
    gotT = ug.cpu().numpy().astype(np.float64)
    errT = np.max(np.abs(gotT - w))
    print("  fp32 rk4 max error %.3e (max |u| %.3e)" % (errT, np.max(np.abs(w))))
    assert errT <= 2e-6 * (1 + np.max(np.abs(w)))


def test_fp32_adjoint_d0_gradient_vs_fp64():
    """FP32 forward and reverse sweep with FP64 gradient reduction: the d0 gradient of
    L = 1/2 |u(T) - target|^2 against the FP64 adjoint and the oracle's forward mode."""
    torch = _torch()
    n, nsteps, every = 16, 24, 5
    prob32, d0, src = _nonlin32(n, seed=3)
    prob64 = molgpu.discretize(*systems.nonlindiff3d(n=n))
    prob64.set_grid_field("d0", d0.reshape((n, n, n), order="F"))
    prob64.set_grid_field("s", src.reshape((n, n, n), order="F"))
    u0 = (0.3 * _on_box(prob64, systems.gaussian_field3d(n, 5))).astype(np.float32).astype(np.float64)
    target = (0.2 * _on_box(prob64, systems.gaussian_field3d(n, 6))).astype(np.float32).astype(np.float64)
    dx = 1.0 / (n - 1)
    dt = 0.1 * dx * dx / (2.0 * np.max(d0))
    grads = {}
    for name, prob, dtype in (("f64", prob64, torch.float64), ("f32", prob32, torch.float32)):
        print("  %s adjoint n=%d" % (name, n), flush=True)
        u = _dev(u0, dtype)
        prob.forward_checkpointed(u, 0.0, dt, nsteps, every)
        lam = (u - _dev(target, dtype)).contiguous()
        g = torch.zeros(2 * n ** 3, dtype=torch.float64, device="cuda")
        prob.adjoint(lam, g)
        grads[name] = _grad_of(prob, g, "d0")
    rel = np.linalg.norm(grads["f32"] - grads["f64"]) / np.linalg.norm(grads["f64"])
    vdir = np.random.default_rng(11).standard_normal(n ** 3)
    _, dL = orc.nonlindiff3d_jvp(n, d0, vdir, src, u0, target, dt, nsteps)
    dir_err = abs(grads["f32"] @ vdir - dL) / abs(dL)
    print("  fp32 d0 gradient: norm-wise error vs fp64 %.3e, directional vs oracle %.3e" % (rel, dir_err))
    assert abs(grads["f64"] @ vdir - dL) <= 1e-8 * abs(dL)
    assert rel <= 1e-4
    assert dir_err <= 1e-4
