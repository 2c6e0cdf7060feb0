"""GPU batches: many uneven diffusion-reaction instances in one persistent launch per stage,
against the oracle per instance and against each instance run on its own."""
import numpy as np
import pytest

import oracle as orc
import molgpu
from molgpu import systems
from molgpu.batch import BatchProblem
from conftest import assert_parity

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _instances(count, seed, dtype="f64", nmin=10, nmax=41):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax))
        D = float(rng.uniform(0.01, 0.1))
        k = float(rng.uniform(0.5, 5.0))
        bct = tuple(int(b) for b in rng.integers(0, 2, 4))
        sys_, disc = systems.diffreact2d(n, D, k, bct, dtype=dtype)
        out.append(((n, D, k, bct), molgpu.discretize(sys_, disc)))
    return out


@pytest.mark.parametrize("tile", [(16, 8, 1), (7, 5, 1), (64, 4, 1), (256, 8, 1)])
def test_batched_rhs_matches_oracle_per_instance(tile):
    """Tiles narrower than the instances (ragged last tiles), and tiles wider than every instance
    (each tile is a band of whole rows, trimmed to the instance's box)."""
    torch = _torch()
    inst = _instances(14, 1)
    batch = BatchProblem([p for _, p in inst], tile=tile)
    plan = batch.plan()
    assert plan["classes"] == len({meta[3] for meta, _ in inst})
    rng = np.random.default_rng(2)
    u = rng.uniform(0, 1, batch.n_unknowns)
    ug = torch.tensor(u, device="cuda")
    du = torch.empty_like(ug)
    print("  batch rhs: %d instances, %d unknowns, plan %s" % (len(inst), batch.n_unknowns, plan), flush=True)
    batch.f(du, ug, 0.0)
    got = batch.split(du.cpu().numpy())
    for ((n, D, k, bct), _), ui, gi in zip(inst, batch.split(u), got):
        ref = orc.diffreact2d_rhs(n, D, k, bct, ui)
        dx = 1.0 / (n - 1)
        assert_parity(gi, ref, np.max(np.abs(ui)) * D * 8 / dx ** 2 + k + 1, what="batch n=%d %s" % (n, bct))


@pytest.mark.parametrize("graph", [False, True])
def test_batched_rk4_equals_instances_run_alone(graph):
    torch = _torch()
    inst = _instances(9, 3)
    batch = BatchProblem([p for _, p in inst])
    dt, N = 1e-4, 20
    u = torch.tensor(batch.u0, device="cuda")
    print("  batch rk4 graph=%s" % graph, flush=True)
    batch.integrate(u, 0.0, dt, N, molgpu.RK4, graph=graph)
    got = batch.split(u.cpu().numpy())
    for (_, p), gi in zip(inst, got):
        ui = torch.tensor(p.u0, device="cuda")
        p.integrate(ui, 0.0, dt, N, molgpu.RK4)
        assert_parity(gi, ui.cpu().numpy(), rtol=1e-13, what="batch vs single")


def test_batched_fp32_kahan_error_vs_fp64():
    """FP32 state with the compensated update, against the FP64 batch (error reported)."""
    torch = _torch()
    inst64 = _instances(6, 4)
    inst32 = _instances(6, 4, dtype="f32")
    b64 = BatchProblem([p for _, p in inst64])
    b32 = BatchProblem([p for _, p in inst32])
    dt, N = 2e-4, 200
    u64 = torch.tensor(b64.u0, device="cuda")
    u32 = torch.tensor(b32.u0, dtype=torch.float32, device="cuda")
    print("  batch fp32 kahan", flush=True)
    b64.integrate(u64, 0.0, dt, N, molgpu.RK4)
    b32.integrate(u32, 0.0, dt, N, molgpu.RK4, kahan=True)
    err = float((u32.double() - u64).abs().max())
    print("  fp32+kahan max error vs fp64: %.3e" % err)
    assert err < 1e-5


def test_autotuner_times_every_candidate_and_keeps_the_fastest(tmp_path, monkeypatch):
    """tune_batch times each candidate tile on the device, returns the fastest, caches it under
    the batch's key, and a batch built with the chosen tile matches the oracle."""
    from molgpu import autotune
    torch = _torch()
    monkeypatch.setenv("MOL_TUNE_CACHE", str(tmp_path / "tune.json"))
    inst = _instances(8, 5)
    lows = [p.lowered for _, p in inst]
    cands = [(16, 4, 1), (8, 8, 1), (32, 2, 1)]
    print("  autotune %d instances over %s" % (len(lows), cands), flush=True)
    tile, table = autotune.tune_batch(lows, candidates=cands)
    assert sorted(table) == sorted("%d,%d,%d" % c for c in cands)
    assert all(v > 0 for v in table.values())
    assert tile in cands and table["%d,%d,%d" % tile] == min(table.values())
    # the cache answers the second call without timing
    again, table2 = autotune.tune_batch(lows, candidates=[(4, 4, 1)])
    assert again == tile and table2 == table
    batch = BatchProblem([p for _, p in inst], tile=tile)
    rng = np.random.default_rng(6)
    u = rng.uniform(0, 1, batch.n_unknowns)
    ug = torch.tensor(u, device="cuda")
    du = torch.empty_like(ug)
    batch.f(du, ug, 0.0)
    for ((n, D, k, bct), _), ui, gi in zip(inst, batch.split(u), batch.split(du.cpu().numpy())):
        ref = orc.diffreact2d_rhs(n, D, k, bct, ui)
        dx = 1.0 / (n - 1)
        assert_parity(gi, ref, np.max(np.abs(ui)) * D * 8 / dx ** 2 + k + 1, what="tuned batch n=%d" % n)


def _heat3d(n, D, bc):
    """u_t = D (u_xx + u_yy + u_zz) + x y, Dirichlet (bc 0) or Neumann (bc 1) on every face."""
    from molgpu.pdesys import Differential, Eq, Interval, PDESystem, parameters, variables
    from molgpu.discretization import MOLFiniteDifference
    t, x, y, z = parameters("t x y z")
    u = variables("u")
    Dx, Dy, Dz = Differential(x), Differential(y), Differential(z)
    eq = Eq(Differential(t)(u(t, x, y, z)),
            D * ((Dx ** 2)(u(t, x, y, z)) + (Dy ** 2)(u(t, x, y, z)) + (Dz ** 2)(u(t, x, y, z))) + x * y)
    bcs = [Eq(u(0, x, y, z), x * y * z)]
    for var, D_ in ((x, Dx), (y, Dy), (z, Dz)):
        for end in (0, 1):
            args = {x: x, y: y, z: z}
            args[var] = end
            w = u(t, args[x], args[y], args[z])
            bcs.append(Eq(w, 0) if bc == 0 else Eq(D_(w), 0))
    doms = [Interval(t, 0.0, 1.0)] + [Interval(v, 0.0, 1.0) for v in (x, y, z)]
    return PDESystem([eq], bcs, doms, [t, x, y, z], [u(t, x, y, z)]), MOLFiniteDifference({x: n, y: n, z: n}, t)


def _family(name, seed):
    rng = np.random.default_rng(seed)
    out = []
    if name == "heat1d":
        for _ in range(9):
            n = int(rng.integers(12, 300))
            bcl = [(1.0, 0.0, 0.3), (0.0, 1.0, -0.2), (1.0, 0.5, 0.1)][int(rng.integers(0, 3))]
            bch = [(1.0, 0.0, -0.4), (0.0, 1.0, 0.25), (2.0, -1.0, 0.0)][int(rng.integers(0, 3))]
            out.append(molgpu.discretize(*systems.heat1d(n=n, D=float(rng.uniform(0.1, 1.0)), bcl=bcl, bch=bch)))
    elif name == "bruss_periodic":
        for _ in range(5):
            out.append(molgpu.discretize(*systems.brusselator2d(n=int(rng.integers(9, 70)))))
    else:
        for _ in range(4):
            n = int(rng.integers(7, 24))
            out.append(molgpu.discretize(*_heat3d(n, float(rng.uniform(0.05, 0.5)), int(rng.integers(0, 2)))))
    return out


@pytest.mark.parametrize("name,tile", [("heat1d", (64, 1, 1)), ("heat1d", (512, 1, 1)), ("bruss_periodic", (256, 16, 1)),
                                       ("bruss_periodic", (16, 8, 1)), ("heat3d", (8, 4, 4)), ("heat3d", (32, 8, 2))])
def test_batched_rhs_equals_each_instance_alone(name, tile):
    """1D instances with Dirichlet / Neumann / Robin ends, periodic 2D instances (the frame of a
    whole-row tile wraps around the ring) and 3D instances (planes of the frame, interior boxes of
    three axes): the batched right-hand side of every instance equals that instance's own
    generated kernel (itself checked against the oracle elsewhere)."""
    torch = _torch()
    probs = _family(name, 11)
    batch = BatchProblem(probs, tile=tile)
    rng = np.random.default_rng(12)
    u = rng.uniform(0.5, 1.5, batch.n_unknowns)
    ug = torch.tensor(u, device="cuda")
    du = torch.empty_like(ug)
    print("  batch %s tile %s: %d instances, %d unknowns" % (name, tile, len(probs), batch.n_unknowns), flush=True)
    batch.f(du, ug, 0.7)
    got = batch.split(du.cpu().numpy())
    for p, ui, gi in zip(probs, batch.split(u), got):
        ut = torch.tensor(ui, device="cuda")
        ref = torch.empty_like(ut)
        p.f(ref, ut, p.p, 0.7)
        r = ref.cpu().numpy()
        # terms of the operator rows: up to 4 alpha / dx^2 |u| (alpha <= 10) plus the reaction
        scale = 40 * np.max(np.abs(ui)) * max(p.lowered.n) ** 2 + 10
        assert_parity(gi, r, scale, what="batch %s vs alone" % name)
