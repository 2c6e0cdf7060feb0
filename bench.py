"""Benchmark of the MI355X path of the MethodOfLines.jl right-hand side and stage loop.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--graph 0|1] [--no-cpu] [--quick]

Headline (`value`): stencil RHS Gpoint/s of the 2D Brusselator (512^2 unknowns per variable,
FP64, periodic, 2nd-order Laplacian) driven by the device-resident RK4 loop at dt = 1e-5.
One step = one classic RK4 step = 4 right-hand sides over all 512^2 points (a point is one
unknown grid node with both variables); value = points * 4 * K / elapsed / 1e9.  The steps run
as fused launches of two steps each (the eight stages on a tile with an 8-node halo in LDS)
where the system allows it; `roofline.stage_kernels` reports the four-launches-per-step path
too.

The same JSON line carries the other parts of the metric:
  roofline     the dominant kernel of the timed region (the fused step kernel) against HBM peak,
               priced at the algorithmic bytes of the steps a launch takes;
  weno5        one FP64 WENO5 Burgers RHS at 1024^2 (and 4096^2, beyond the 256 MB MALL), and the
               1024^2 SSPRK3 solve (CFL 0.4, 500 steps) in FP64, FP32 + Kahan and plain FP32 with
               the FP32 errors against FP64 and the FP64 error against the oracle;
  adjoint      3D nonlinear diffusion 128^3 (2.0 M unknowns), 200 RK4 steps: plain forward time,
               then per checkpoint mode (every 20 steps with recomputation, and the automatic
               choice, which keeps every stage state when it fits in device memory) the
               recording forward and the adjoint sweep; ratio = adjoint sweep / plain forward
               of the automatic mode; gradient errors against the oracle's forward mode and
               against central differences;
  batch        512 diffusion-reaction instances with uneven grids, FP32 + Kahan, 200 RK4 steps;
  heat         1D heat, 101 nodes, RK4 at dt = 1e-5 to t = 0.1: time per step and the error
               against the analytic solution and against the oracle's RK4;
  cpu_baseline the C++ CPU restatement (the oracle) on the host, 1 thread, bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "methodoflines.jl_amd"))

HBM_PEAK_GBS = 8000.0

# HBM bytes per RK4 stage launch of the headline workload, from two rocprofv3 --pmc runs of
# `bench.py --quick --no-cpu --steps 50` (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch, averaged
# over the stage launches), each scaled by what the counter reads for a device copy of 64 MiB
# (`bench.py --section copy`: FETCH_SIZE x 2.0, WRITE_SIZE x 1.0); summary in
# profiles/brusselator512_rk4_stage_kernel.md.  Below the 17.8 MB of algorithmic bytes: halo
# re-reads and the stage vectors written by one launch and read by the next hit in L2.
STAGE_TRAFFIC_BYTES = (3960.27 * 2.0 + 5869.53 * 1.0) * 1024
# The same measurement for the fused RK4 kernel (two steps per launch, 32x32 tiles, 8-node
# halo): FETCH_SIZE 2689.9 KiB, WRITE_SIZE 4096.0 KiB per launch (16.8 MB algorithmic for the
# two steps; a launch reads u once with its halo and writes it once).
FUSED_TRAFFIC_BYTES = (2689.9 * 2.0 + 4096.0 * 1.0) * 1024
# Issue floor of one fused launch from its SQ counters (profiles/brusselator512_rk4_fused_step.md):
# 676.0 VALU instructions per wave over 4096 waves; a wave64 VALU instruction occupies a SIMD for
# 4 cycles, and the chip has 1024 SIMDs at 2.4 GHz.
FUSED_VALU_PER_WAVE, FUSED_WAVES = 676.0, 4096
VALU_FLOOR_US = FUSED_VALU_PER_WAVE * FUSED_WAVES * 4 / (1024 * 2.4e9) * 1e6


def progress(msg):
    """One line on stderr per section of a long run (the JSON line alone goes to stdout)."""
    print("[bench] %s" % msg, file=sys.stderr, flush=True)


def host_cpu():
    """Model name and logical CPU count of the host (from /proc/cpuinfo)."""
    model, n = None, 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    n += 1
                    model = model or line.split(":", 1)[1].strip()
    except OSError:
        pass
    return {"model": model, "logical_cpus": n or os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def rhs_time_us(prob, u, reps):
    """Average duration of back-to-back right-hand-side launches on the current stream."""
    import ctypes as C
    import torch
    h = prob.handle
    du = torch.empty_like(u)
    pp = prob.device_params()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = (h.h, C.c_void_p(du.data_ptr()), C.c_void_p(u.data_ptr()),
            C.c_void_p(pp.data_ptr()) if pp is not None else None, C.c_double(0.0), s)
    for _ in range(5):
        h.lib.mol_rhs(*args)
    torch.cuda.synchronize()
    e0, e1 = events()
    e0.record()
    for _ in range(reps):
        h.lib.mol_rhs(*args)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e3 / reps


def bench_weno(n, reps):
    import torch
    import molgpu
    from molgpu import systems
    sys_, disc = systems.burgers2d(n=n + 1)
    prob = molgpu.discretize(sys_, disc)
    u = torch.tensor(prob.u0, dtype=torch.float64, device="cuda")
    us = rhs_time_us(prob, u, reps)
    pts = n * n
    ach = pts * 16 / (us * 1e-6) / 1e9
    return {"grid": "%dx%d" % (n, n), "rhs_us": us, "gpoint_s": pts / (us * 1e-6) / 1e9,
            "bytes_per_launch": pts * 16, "achieved_gbs": ach, "hbm_frac": ach / HBM_PEAK_GBS}


def bench_weno_nonuniform(n, reps):
    """The WENO5 Burgers RHS on seeded stretched node arrays along both axes (n^2 unknowns, FP64):
    the nonuniform operator of weno_f_nonuniform, its coordinate-only coefficients read from the
    per-axis tables the library builds at handle creation."""
    import torch
    import molgpu
    from molgpu import systems
    prob = molgpu.discretize(*systems.burgers2d_nonuniform(n=n + 1))
    u = torch.tensor(prob.u0, dtype=torch.float64, device="cuda")
    us = rhs_time_us(prob, u, reps)
    pts = n * n
    ach = pts * 16 / (us * 1e-6) / 1e9
    return {"grid": "%dx%d stretched" % (n, n), "rhs_us": us, "gpoint_s": pts / (us * 1e-6) / 1e9,
            "bytes_per_launch": pts * 16, "achieved_gbs": ach, "hbm_frac": ach / HBM_PEAK_GBS}


def bench_weno_solve(n, nsteps, oracle_steps):
    """The WENO5 Burgers workload as an integration: SSPRK3 at CFL 0.4 for `nsteps` steps in FP64,
    in FP32 with the compensated update and in plain FP32; errors of the FP32 runs against the
    FP64 run, and of the FP64 run against the oracle's SSPRK3 over `oracle_steps` steps."""
    import numpy as np
    import torch
    import molgpu
    from molgpu import systems
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    p64 = molgpu.discretize(*systems.burgers2d(n=n + 1))
    p32 = molgpu.discretize(*systems.burgers2d(n=n + 1, dtype="f32"))
    u0 = np.asarray(p64.u0, dtype=np.float64)
    dx = dy = 1.0 / n
    umax = float(np.max(np.abs(u0)))
    dt = 0.4 / (umax / dx + umax / dy)          # CFL 0.4 over both axes
    res = {"grid": "%dx%d" % (n, n), "scheme": "SSPRK33", "cfl": 0.4, "dt": dt, "steps": nsteps}
    finals = {}
    for name, prob, kahan in (("fp64", p64, False), ("fp32_kahan", p32, True), ("fp32", p32, False)):
        dtype = torch.float64 if name == "fp64" else torch.float32
        u = torch.tensor(u0, dtype=dtype, device="cuda")
        prob.integrate(u, 0.0, dt, 5, molgpu.SSPRK33, kahan=kahan, graph=True)
        u = torch.tensor(u0, dtype=dtype, device="cuda")
        torch.cuda.synchronize()
        e0, e1 = events()
        e0.record()
        prob.integrate(u, 0.0, dt, nsteps, molgpu.SSPRK33, kahan=kahan, graph=True)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        res[name] = {"ms": ms, "ms_per_step": ms / nsteps,
                     "gpoint_rhs_s": n * n * 3 * nsteps / (ms / 1e3) / 1e9}
    # errors in chunks of 50 steps, at the last chunk where the FP64 solution is finite (the
    # scheme applies WENO5 to u_x without upwinding, as the reference does, and the shocks of
    # this initial state break it down near t = 0.15 at 1024^2)
    us = {name: torch.tensor(u0, dtype=torch.float64 if name == "fp64" else torch.float32, device="cuda")
          for name in ("fp64", "fp32_kahan", "fp32")}
    done = 0
    while done < nsteps:
        k = min(50, nsteps - done)
        for name, prob, kahan in (("fp64", p64, False), ("fp32_kahan", p32, True), ("fp32", p32, False)):
            prob.integrate(us[name], done * dt, dt, k, molgpu.SSPRK33, kahan=kahan, graph=True)
        if not bool(torch.isfinite(us["fp64"]).all()):
            break
        done += k
        for name in ("fp64", "fp32_kahan", "fp32"):
            finals[name] = us[name].double().clone()
    res["finite_steps"] = done
    if finals:
        res["max_abs_u"] = float(finals["fp64"].abs().max())
        for name in ("fp32_kahan", "fp32"):
            res[name]["max_abs_err_vs_fp64"] = float((finals[name] - finals["fp64"]).abs().max())
    # FP64 run against the oracle over the first steps
    u = torch.tensor(u0, device="cuda")
    p64.integrate(u, 0.0, dt, oracle_steps, molgpu.SSPRK33)
    ref = orc.burgers2d_integrate(n + 1, u0, 0.0, dt, oracle_steps, threads=min(16, os.cpu_count() or 1))
    res["fp64_vs_oracle"] = {"steps": oracle_steps,
                             "max_abs_err": float(np.max(np.abs(u.cpu().numpy() - ref))),
                             "max_abs_u": float(np.max(np.abs(ref)))}
    return res


def bench_rhs3d(n, reps):
    """One FP64 RHS of the 3D nonlinear diffusion (config 5 system) at n^3 nodes."""
    import numpy as np
    import torch
    import molgpu
    from molgpu import systems
    s, d = systems.nonlindiff3d(n)
    p = molgpu.discretize(s, d)
    p.set_grid_field("d0", 0.5 + 0.5 * np.abs(systems.gaussian_field3d(n, 1)))
    p.set_grid_field("s", systems.gaussian_field3d(n, 2))
    u = torch.tensor(np.random.default_rng(0).uniform(-1, 1, p.lowered.n_unknowns), device="cuda")
    us = rhs_time_us(p, u, reps)
    pts = p.lowered.n_unknowns
    ach = pts * 32 / (us * 1e-6) / 1e9
    lam = torch.randn_like(u)
    for _ in range(3):
        p.vjp(u, lam)
    torch.cuda.synchronize()
    e0, e1 = events()
    e0.record()
    for _ in range(20):
        p.vjp(u, lam)
    e1.record()
    torch.cuda.synchronize()
    return {"grid": "%d^3" % n, "unknowns": pts, "plan": p.plan(), "rhs_us": us, "achieved_gbs": ach,
            "hbm_frac": ach / HBM_PEAK_GBS, "vjp_us": e0.elapsed_time(e1) * 1e3 / 20}


def bench_adjoint(n, nsteps, every, fd_cells):
    """Config 5: forward (checkpointed) and adjoint times; gradient error against the oracle's
    forward mode on a 16^3 copy and against central differences of the GPU solve on cells."""
    import numpy as np
    import torch
    import molgpu
    from molgpu import systems
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    def make(m):
        s, d = systems.nonlindiff3d(m)
        p = molgpu.discretize(s, d)
        d0 = 0.5 + 0.5 * np.abs(systems.gaussian_field3d(m, 1))
        src = systems.gaussian_field3d(m, 2)
        p.set_grid_field("d0", d0)
        p.set_grid_field("s", src)
        (l0, l1, l2), (h0, h1, h2) = p.lowered.lo[0], p.lowered.hi[0]
        box = lambda f: f[l0 - 1:h0, l1 - 1:h1, l2 - 1:h2].reshape(-1, order="F")
        u0 = 0.3 * box(systems.gaussian_field3d(m, 5))
        tgt = 0.2 * box(systems.gaussian_field3d(m, 6))
        # explicit-diffusion step: dt = 0.5 * 2.78 * dx^2 / (12 * max D) with D <= 2 max d0
        dx = 1.0 / (m - 1)
        dt = 0.5 * 2.78 * dx * dx / (12.0 * 2.0 * d0.max())
        return p, d0.reshape(-1, order="F"), src.reshape(-1, order="F"), u0, tgt, dt

    p, d0, src, u0, tgt, dt = make(n)
    ng = n ** 3
    u0d = torch.tensor(u0, device="cuda")
    tgd = torch.tensor(tgt, device="cuda")
    g = torch.zeros(2 * ng, dtype=torch.float64, device="cuda")

    def run(ck, grid_grads=True):
        u = u0d.clone()
        e0, e1 = events()
        e2 = torch.cuda.Event(enable_timing=True)
        e0.record()
        p.forward_checkpointed(u, 0.0, dt, nsteps, ck)
        e1.record()
        lam = (u - tgd).contiguous()
        if grid_grads:
            g.zero_()
        p.adjoint(lam, g if grid_grads else None)
        e2.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1), e1.elapsed_time(e2), u

    def plain():
        u = u0d.clone()
        e0, e1 = events()
        e0.record()
        p.integrate(u, 0.0, dt, nsteps, molgpu.RK4, graph=True)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    plain()
    fwd_ms = plain()
    modes = {}
    auto = p.checkpoint_interval(nsteps)
    for ck in sorted({every, auto}, reverse=True):
        run(ck)
        fc_ms, a_ms, uT = run(ck)
        modes["ckpt_every=%d" % ck] = {"forward_ckpt_ms": fc_ms, "adjoint_ms": a_ms, "ratio": a_ms / fwd_ms,
                                       "forward_plus_adjoint_ratio": (fc_ms + a_ms) / fwd_ms}
    # the initial-state gradient alone (no grid-field gradients), every stage state kept
    run(0, False)
    _, a0_ms, _ = run(0, False)
    modes["ckpt_every=0, initial state only"] = {"adjoint_ms": a0_ms, "ratio": a0_ms / fwd_ms}
    # g holds the gradient of the last mode run with grid-field gradients (the automatic choice)
    ia = p.lowered.aux.index("d0")
    gd0 = g[ia * ng:(ia + 1) * ng].cpu().numpy()
    # finite differences of the GPU solve on seeded cells (Richardson-extrapolated central)
    d0t = torch.tensor(d0, device="cuda")
    p.set_grid_field("d0", d0t)

    def final(e, c):
        d0t.copy_(torch.tensor(d0, device="cuda"))
        d0t[c] += e
        u = u0d.clone()
        p.integrate(u, 0.0, dt, nsteps, molgpu.RK4, graph=True)
        return u

    def dloss(e, c):
        # L(+e) - L(-e) = 1/2 sum (u+ - u-)(u+ + u- - 2 target): no cancellation of two losses
        up, um = final(e, c), final(-e, c)
        return 0.5 * float(((up - um) * (up + um - 2 * tgd)).sum())

    rng = np.random.default_rng(7)
    cells = rng.choice(ng, size=fd_cells, replace=False)
    fds = []
    for c in cells:
        h = 1e-2 * d0[c]
        D1 = dloss(h, c) / (2 * h)
        D2 = dloss(h / 2, c) / h
        fds.append((4 * D2 - D1) / 3)        # Richardson: O(h^4)
    fds = np.array(fds)
    gc = gd0[cells]
    # norm-wise relative error of the gradient on the sampled cells; the per-cell maximum is
    # limited by the finite differences' rounding on cells whose gradient is small
    fd_err = float(np.linalg.norm(gc - fds) / np.linalg.norm(fds))
    fd_cell = float(np.max(np.abs(gc - fds) / np.maximum(np.abs(fds), 1e-300)))
    fd_abs, fd_max = float(np.max(np.abs(gc - fds))), float(np.max(np.abs(fds)))
    d0t.copy_(torch.tensor(d0, device="cuda"))
    # oracle forward mode on a 16^3 copy (same construction), direction seeded
    p16, d016, src16, u016, tgt16, dt16 = make(16)
    u = torch.tensor(u016, device="cuda")
    p16.forward_checkpointed(u, 0.0, dt16, nsteps, auto)
    lam = (u - torch.tensor(tgt16, device="cuda")).contiguous()
    g16 = torch.zeros(2 * 16 ** 3, dtype=torch.float64, device="cuda")
    p16.adjoint(lam, g16)
    v = np.random.default_rng(3).standard_normal(16 ** 3)
    _, dL = orc.nonlindiff3d_jvp(16, d016, v, src16, u016, tgt16, dt16, nsteps)
    i16 = p16.lowered.aux.index("d0")
    fm_err = abs(float(g16[i16 * 16 ** 3:(i16 + 1) * 16 ** 3].cpu().numpy() @ v) - dL) / abs(dL)
    rhs_us = rhs_time_us(p, u0d, 20)
    best = modes["ckpt_every=%d" % auto]
    return {"grid": "%d^3" % n, "unknowns": p.lowered.n_unknowns, "steps": nsteps, "dt": dt,
            "forward_ms": fwd_ms, "ckpt_every": auto, "adjoint_ms": best["adjoint_ms"], "ratio": best["ratio"],
            "modes": modes, "rhs_us": rhs_us, "grad_rel_err_forward_mode_16": fm_err,
            "grad_rel_err_fd_cells": fd_err, "grad_err_fd_cells_vs_max": fd_abs / max(fd_max, 1e-300),
            "grad_max_cell_rel_err_fd": fd_cell,
            "fd_cells": int(fd_cells)}


def bench_adjoint_fp32(n, nsteps, every):
    """Config 5 in FP32: state, stencil arithmetic and stage adjoints in FP32, the grid-field
    gradient reduced in FP64.  Plain forward, and per checkpoint mode the recording forward and
    the adjoint sweep; the final state and the d0 gradient against the FP64 run of the same
    (FP32-representable) inputs."""
    import numpy as np
    import torch
    import molgpu
    from molgpu import systems
    r32 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)
    d0 = r32(0.5 + 0.5 * np.abs(systems.gaussian_field3d(n, 1)))
    src = r32(systems.gaussian_field3d(n, 2))
    probs = {}
    for name in ("f64", "f32"):
        s, d = systems.nonlindiff3d(n, dtype=name)
        probs[name] = molgpu.discretize(s, d)
        probs[name].set_grid_field("d0", d0)
        probs[name].set_grid_field("s", src)
    p64, p32 = probs["f64"], probs["f32"]
    (l0, l1, l2), (h0, h1, h2) = p64.lowered.lo[0], p64.lowered.hi[0]
    box = lambda f: f[l0 - 1:h0, l1 - 1:h1, l2 - 1:h2].reshape(-1, order="F")
    u0 = r32(0.3 * box(systems.gaussian_field3d(n, 5)))
    tgt = r32(0.2 * box(systems.gaussian_field3d(n, 6)))
    dx = 1.0 / (n - 1)
    dt = 0.5 * 2.78 * dx * dx / (12.0 * 2.0 * d0.max())
    ng = n ** 3
    ia = p64.lowered.aux.index("d0")

    def adjoint(p, dtype, ck):
        u = torch.tensor(u0, dtype=dtype, device="cuda")
        g = torch.zeros(2 * ng, dtype=torch.float64, device="cuda")
        e0, e1 = events()
        e2 = torch.cuda.Event(enable_timing=True)
        e0.record()
        p.forward_checkpointed(u, 0.0, dt, nsteps, ck)
        e1.record()
        lam = (u - torch.tensor(tgt, dtype=dtype, device="cuda")).contiguous()
        p.adjoint(lam, g)
        e2.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1), e1.elapsed_time(e2), u.double(), g[ia * ng:(ia + 1) * ng]

    _, _, u64, g64 = adjoint(p64, torch.float64, every)
    u = torch.tensor(u0, dtype=torch.float32, device="cuda")
    p32.integrate(u, 0.0, dt, nsteps, molgpu.RK4, graph=True)
    u = torch.tensor(u0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    e0, e1 = events()
    e0.record()
    p32.integrate(u, 0.0, dt, nsteps, molgpu.RK4, graph=True)
    e1.record()
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1)
    modes = {}
    g32 = None
    for ck in sorted({every, p32.checkpoint_interval(nsteps)}, reverse=True):
        adjoint(p32, torch.float32, ck)
        fc_ms, a_ms, u32, g32 = adjoint(p32, torch.float32, ck)
        modes["ckpt_every=%d" % ck] = {"forward_ckpt_ms": fc_ms, "adjoint_ms": a_ms, "ratio": a_ms / fwd_ms,
                                       "forward_plus_adjoint_ratio": (fc_ms + a_ms) / fwd_ms}
    return {"grid": "%d^3" % n, "steps": nsteps, "dt": dt, "forward_ms": fwd_ms, "modes": modes,
            "state_max_abs_err_vs_fp64": float((u32 - u64).abs().max()), "state_max_abs": float(u64.abs().max()),
            "grad_d0_rel_err_vs_fp64": float((g32 - g64).norm() / g64.norm()),
            "grad_d0_max_err_vs_max": float((g32 - g64).abs().max() / g64.abs().max())}


def bench_heat(n=101, D=1.0, dt=1e-5, t_end=0.1):
    """Config 0: 1D heat u_t = D u_xx, Dirichlet 0, u0 = sin(pi x), RK4 at dt to t_end: time per
    step, error against the analytic solution exp(-pi^2 D t) sin(pi x) and against the oracle's RK4
    over the same steps."""
    import numpy as np
    import torch
    import molgpu
    from molgpu import systems
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    prob = molgpu.discretize(*systems.heat1d(n=n, D=D))
    nsteps = int(round(t_end / dt))
    u = torch.tensor(prob.u0, dtype=torch.float64, device="cuda")
    prob.integrate(u, 0.0, dt, 100, molgpu.RK4, graph=True)
    u = torch.tensor(prob.u0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    e0, e1 = events()
    e0.record()
    prob.integrate(u, 0.0, dt, nsteps, molgpu.RK4, graph=True)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    got = u.cpu().numpy()
    x = prob.lowered.coords[0][1:-1]
    exact = np.exp(-np.pi ** 2 * D * t_end) * np.sin(np.pi * x)
    ref = orc.diffusion1d_integrate(n, D, (1.0, 0.0, 0.0), (1.0, 0.0, 0.0), prob.u0, 0.0, dt, nsteps)
    return {"n": n, "steps": nsteps, "dt": dt, "ms": ms, "us_per_step": ms * 1e3 / nsteps,
            "max_err_vs_analytic": float(np.max(np.abs(got - exact))),
            "max_diff_vs_oracle_rk4": float(np.max(np.abs(got - ref)))}


def bench_batch(count, nsteps, quick):
    """Config 4: `count` diffusion-reaction instances, seeded sizes 32..256 per side, FP32 +
    Kahan, RK4 in one persistent launch per stage (graph-captured)."""
    import numpy as np
    import torch
    import molgpu
    from molgpu import systems
    from molgpu.batch import BatchProblem
    from molgpu.discretization import lower
    from molgpu.autotune import tune_batch

    rng = np.random.default_rng(2024)
    lows, lows64 = [], []
    t0 = time.perf_counter()
    dts = []
    for _ in range(count):
        n = int(rng.integers(32, 257))
        D = float(rng.uniform(0.01, 0.1))
        k = float(rng.uniform(0.5, 5.0))
        bct = tuple(int(b) for b in rng.integers(0, 2, 4))
        u0 = rng.uniform(0, 1, (n, n))
        for _ in range(3):     # Jacobi smoothing: 5-point average, boundary rows kept
            u0[1:-1, 1:-1] = 0.2 * (u0[1:-1, 1:-1] + u0[:-2, 1:-1] + u0[2:, 1:-1] + u0[1:-1, :-2] + u0[1:-1, 2:])
        for dtype, out in (("f32", lows), ("f64", lows64)):
            sys_, disc = systems.diffreact2d(n, D, k, bct, dtype=dtype)
            low = lower(sys_, disc)
            (l0, l1), (h0, h1) = low.lo[0], low.hi[0]
            box = u0[l0 - 1:h0, l1 - 1:h1].reshape(-1, order="F")
            low.u0 = box.astype(np.float32) if dtype == "f32" else box.astype(np.float32).astype(np.float64)
            out.append(low)
        dx = 1.0 / (n - 1)
        dts.append(0.5 * 2.78 * dx * dx / (8 * D))
    lower_s = time.perf_counter() - t0
    dt = min(dts)                      # one global step, stable for every instance
    tile, table = tune_batch(lows, quick=quick)
    b = BatchProblem(lows, tile=tile)
    u = torch.tensor(b.u0, dtype=torch.float32, device="cuda")
    b.integrate(u, 0.0, dt, 4, molgpu.RK4, kahan=True, graph=True)
    torch.cuda.synchronize()
    u = torch.tensor(b.u0, dtype=torch.float32, device="cuda")
    e0, e1 = events()
    e0.record()
    b.integrate(u, 0.0, dt, nsteps, molgpu.RK4, kahan=True, graph=True)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    pts = b.n_unknowns
    # error of the FP32 runs (with and without the compensated update) against the same batch in
    # FP64 from the same (FP32-representable) initial state
    b64 = BatchProblem(lows64, tile=tile)
    u64 = torch.tensor(b64.u0, dtype=torch.float64, device="cuda")
    b64.integrate(u64, 0.0, dt, nsteps, molgpu.RK4, graph=True)
    u32 = torch.tensor(b.u0, dtype=torch.float32, device="cuda")
    b.integrate(u32, 0.0, dt, nsteps, molgpu.RK4, kahan=False, graph=True)
    torch.cuda.synchronize()
    scale = float(u64.abs().max())
    err_k = float((u.double() - u64).abs().max())
    err_p = float((u32.double() - u64).abs().max())
    # algorithmic bytes of an FP32 + Kahan RK4 step per unknown: stage 1 reads u and writes y1 and
    # acc (12 B); stages 2 and 3 read y, u, acc and write y', acc (20 B each); stage 4 reads y3, u,
    # acc and the compensation, writes the compensation and u (24 B): 76 B
    step_bytes = pts * 76
    ach = step_bytes / (ms / nsteps * 1e-3) / 1e9
    return {"instances": count, "unknowns": pts, "classes": b.plan()["classes"], "tile": list(tile),
            "bytes_per_step": step_bytes, "achieved_gbs": ach, "hbm_frac": ach / HBM_PEAK_GBS,
            "tuning": table, "steps": nsteps, "dt": dt, "ms": ms, "ms_per_step": ms / nsteps,
            "gpoint_rhs_s": pts * 4 * nsteps / (ms / 1e3) / 1e9, "lowering_s": lower_s,
            "finite": bool(torch.isfinite(u).all()),
            "err_vs_fp64": {"fp32_kahan_max_abs": err_k, "fp32_plain_max_abs": err_p, "max_abs_u": scale}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    # direct launches by default: at the driver's 20 steps they beat two graph replays
    # (6.7 against 7.4 us per step); at 1000 steps the two are within 1 %
    ap.add_argument("--graph", type=int, default=0)
    ap.add_argument("--unfused", action="store_true", help="time the four stage launches per step only")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline only (no weno/adjoint/batch sections)")
    ap.add_argument("--section", default="all",
                    choices=["all", "weno", "weno4096", "wenonu", "wenosolve", "adjoint", "adjoint32", "batch", "rhs3d",
                             "copy", "heat"],
                    help="run one section alone (for profiling); prints its JSON")
    ap.add_argument("--adj-steps", type=int, default=200)
    ap.add_argument("--reps", type=int, default=200, help="back-to-back RHS launches of the weno/rhs3d sections")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import molgpu
    from molgpu import systems

    if args.section != "all":
        if args.section == "weno":
            res = {"1024": bench_weno(1024, args.reps)}
        elif args.section == "weno4096":
            res = {"4096": bench_weno(4096, args.reps)}
        elif args.section == "wenonu":
            res = {"1024": bench_weno_nonuniform(1024, args.reps)}
        elif args.section == "wenosolve":
            res = bench_weno_solve(1024, 500, 3)
        elif args.section == "adjoint":
            res = bench_adjoint(128, args.adj_steps, 20, 2)
        elif args.section == "adjoint32":
            res = bench_adjoint_fp32(128, args.adj_steps, 20)
        elif args.section == "batch":
            res = bench_batch(512, 50, quick=True)
        elif args.section == "heat":
            res = bench_heat()
        elif args.section == "copy":
            # calibration of the HBM byte counters: device copies of a known size
            x = torch.ones(8 << 20, dtype=torch.float64, device="cuda")
            y = torch.empty_like(x)
            for _ in range(10):
                y.copy_(x)
            torch.cuda.synchronize()
            res = {"bytes_read_per_copy": x.numel() * 8, "bytes_written_per_copy": x.numel() * 8, "copies": 10}
        else:
            res = bench_rhs3d(128, args.reps)
        print(json.dumps({"section": args.section, "result": res}))
        return

    n = 513
    sys_, disc = systems.brusselator2d(n=n)
    prob = molgpu.discretize(sys_, disc)
    points = (n - 1) * (n - 1)
    dt = 1e-5
    stream = torch.cuda.current_stream()
    u = torch.tensor(prob.u0, dtype=torch.float64, device="cuda")
    u_start = u.clone()
    fused = prob.plan()["fused_tiles"] > 0 and not args.unfused

    def stepper(use_fused, nsteps):
        """mol_integrate through the C ABI with its arguments prepared (as a Julia ccall binding
        would hold them), so the timed region holds the library's own work only."""
        import ctypes as C
        from molgpu import native
        h = prob.handle
        pp = prob.device_params()
        flags = (native.MOL_FLAG_GRAPH if args.graph else 0) | (0 if use_fused else native.MOL_FLAG_UNFUSED)
        a = (h.h, C.c_void_p(u.data_ptr()), C.c_void_p(pp.data_ptr()) if pp is not None else None, C.c_double(0.0),
             C.c_double(dt), C.c_int64(nsteps), C.c_int32(molgpu.RK4.scheme), C.c_int32(flags),
             C.c_void_p(stream.cuda_stream))

        def run():
            native.check(h.lib, h.lib.mol_integrate(*a), h.h)
        return run

    def timed(use_fused):
        warm, step = stepper(use_fused, max(args.warmup, 1)), stepper(use_fused, args.steps)
        u.copy_(u_start)
        warm()
        torch.cuda.synchronize()
        u.copy_(u_start)
        torch.cuda.synchronize()
        ev0, ev1 = events()
        t0 = time.perf_counter()
        ev0.record(stream)
        step()
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1), time.perf_counter() - t0

    # the stage-kernel path (four launches per step) first, for the comparison line
    ms_st, _ = timed(False)
    ms, wall = timed(fused)
    ms_per_step = ms / args.steps
    value = points * 4 * args.steps / (ms / 1e3) / 1e9

    # Stage-kernel path: the generated RHS kernel in RK4 stage mode, four launches per step.
    # Algorithmic bytes per step and point (2 variables, FP64): stage 1 reads u, writes y1 and acc
    # (48 B); stages 2-3 read y, u, acc, write y', acc (80 B each); stage 4 reads y3, u, acc and
    # writes u (64 B): 272 B per step, 68 B per launch on average.
    st_us = ms_st * 1e3 / (4 * args.steps)
    st_bytes = points * 272 // 4
    st_ach = st_bytes / (st_us * 1e-6) / 1e9
    stage = {"kernel": "mol_rhs_tiles (generated; RK4 stage epilogue)", "avg_us": st_us, "bytes_per_launch": st_bytes,
             "achieved": st_ach, "frac": st_ach / HBM_PEAK_GBS, "traffic": STAGE_TRAFFIC_BYTES,
             "value": points * 4 * args.steps / (ms_st / 1e3) / 1e9}
    spl = max(1, prob.plan()["fused_steps_per_launch"]) if fused else 1
    if fused:
        # Fused path: a launch takes spl whole RK4 steps (spl = 2: an 8-node halo).  Algorithmic
        # bytes of one step: read u, write u, 32 B per point (2 variables, FP64); a launch of
        # spl steps is priced at spl such steps (it moves the bytes of one, which is the point of
        # fusing them).  Launch time = the time of spl steps.
        launch_us = ms * 1e3 * spl / args.steps
        bytes_launch = points * 32 * spl
        kname = "mol_rk4_fused (generated; %d whole RK4 step%s per launch)" % (spl, "s" if spl > 1 else "")
        traffic = FUSED_TRAFFIC_BYTES if spl == 2 else None
    else:
        launch_us, bytes_launch, kname, traffic = st_us, st_bytes, stage["kernel"], STAGE_TRAFFIC_BYTES
    achieved = bytes_launch / (launch_us * 1e-6) / 1e9
    rhs_us = rhs_time_us(prob, u_start, 200)
    rhs_ach = points * 32 / (rhs_us * 1e-6) / 1e9

    out = {
        "metric": "stencil RHS Gpoint/s (2D Brusselator 512^2, FP64, RK4 dt=1e-5)",
        "value": value,
        "unit": "Gpoint-RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: the configured initial condition u=22(y(1-y))^1.5, v=27(x(1-x))^1.5",
        "config": {"workload": "brusselator2d periodic 512x512 unknowns/variable, 2 variables, approx_order 2, "
                               "RK4 fixed dt=1e-5", "points": points, "rhs_per_step": 4,
                   "graph": bool(args.graph), "fused_step": fused, "steps_per_launch": spl},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kname, "avg_us": launch_us, "bytes_per_launch": bytes_launch,
                     # floors of the launch: HBM (algorithmic bytes at peak) and VALU issue (the
                     # binding one of the two; the rest is the serial load / compute / store phases)
                     "hbm_floor_us": bytes_launch / (HBM_PEAK_GBS * 1e3),
                     "valu_floor_us": VALU_FLOOR_US if fused and spl == 2 else None,
                     "stage_kernels": stage,
                     "rhs_only": {"avg_us": rhs_us, "bytes_per_launch": points * 32, "achieved": rhs_ach,
                                  "frac": rhs_ach / HBM_PEAK_GBS}},
        "wall_s": wall,
    }
    if not args.quick:
        progress("weno5 sections")
        out["weno5"] = {"1024": bench_weno(1024, 200), "4096": bench_weno(4096, 20),
                        "nonuniform_1024": bench_weno_nonuniform(1024, 200),
                        "solve_1024": bench_weno_solve(1024, 500, 3)}
        progress("adjoint section")
        out["adjoint"] = bench_adjoint(128, 200, 20, 32)
        progress("adjoint FP32 section")
        out["adjoint"]["fp32"] = bench_adjoint_fp32(128, 200, 20)
        progress("heat section")
        out["heat"] = bench_heat()
        progress("batch section")
        out["batch"] = bench_batch(512, 200, quick=False)

    cpu = None
    if not args.no_cpu:
        progress("CPU baseline (bounded sample)")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        ucpu = prob.u0.copy()

        def sample(threads, secs):
            reps = 0
            c0 = time.perf_counter()
            while time.perf_counter() - c0 < secs:      # bounded sample of CPU work
                orc.bruss2d_rhs(n, ucpu, 1.2, threads=threads)
                reps += 1
            return reps, points * reps / (time.perf_counter() - c0) / 1e9

        # the host share of one GPU: at most 16 threads, whatever the machine reports
        allc = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()))
        reps1, v1 = sample(1, 8.0)
        repsn, vn = sample(allc, 8.0)
        cpu = {"value": vn, "unit": "Gpoint-RHS/s", "cores": allc, "kind": "port",
               "sample": "%d right-hand sides of the same 512^2 Brusselator in the C++ oracle (CPU restatement, "
                         "not MethodOfLines.jl), FP64, OpenMP over %d threads; %d more on 1 thread"
                         % (repsn, allc, reps1),
               "one_thread": v1, "host": host_cpu()}
    out["cpu_baseline"] = cpu
    print(json.dumps(out))


if __name__ == "__main__":
    main()
