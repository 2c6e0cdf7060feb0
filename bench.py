"""Benchmark: stencil RHS Gpoint/s of the MI355X path on the 2D Brusselator (512^2 unknowns per
variable, FP64, periodic, 2nd-order Laplacian), driven by the device-resident RK4 stage loop
at dt = 1e-5.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--graph 0|1] [--no-cpu]

One step = one classic RK4 step = 4 right-hand-side evaluations over all 512^2 points (a point
is one unknown grid node with both variables).  value = points * 4 * K / elapsed / 1e9.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "methodoflines.jl_amd"))

HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--rhs-reps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import molgpu
    from molgpu import systems

    n = 513
    sys_, disc = systems.brusselator2d(n=n)
    prob = molgpu.discretize(sys_, disc)
    points = (n - 1) * (n - 1)
    dt = 1e-5
    stream = torch.cuda.current_stream()
    u = torch.tensor(prob.u0, dtype=torch.float64, device="cuda")
    u_start = u.clone()
    # warmup
    prob.integrate(u, 0.0, dt, max(args.warmup, 1), molgpu.RK4, graph=bool(args.graph))
    torch.cuda.synchronize()
    u.copy_(u_start)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    prob.integrate(u, 0.0, dt, args.steps, molgpu.RK4, graph=bool(args.graph))
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1)
    ms_per_step = ms / args.steps
    value = points * 4 * args.steps / (ms / 1e3) / 1e9

    # dominant kernel: the right-hand-side kernel; average launch duration over back-to-back launches
    du = torch.empty_like(u)
    uu = u_start.clone()
    for _ in range(5):
        prob.f(du, uu, None, 0.0)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.rhs_reps):
        prob.f(du, uu, None, 1.2)
    e1.record(stream)
    torch.cuda.synchronize()
    rhs_us = e0.elapsed_time(e1) * 1e3 / args.rhs_reps
    rhs_bytes = points * 2 * 16          # read u, v; write du, dv (FP64)
    achieved = rhs_bytes / (rhs_us * 1e-6) / 1e9
    rhs_gpts = points / (rhs_us * 1e-6) / 1e9

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        threads = 1
        ucpu = prob.u0.copy()
        reps = 0
        c0 = time.perf_counter()
        while time.perf_counter() - c0 < 10.0:      # bounded sample: ~10 s of CPU work
            orc.bruss2d_rhs(n, ucpu, 1.2, threads=threads)
            reps += 1
        csec = time.perf_counter() - c0
        cpu = {"value": points * reps / csec / 1e9, "unit": "Gpoint-RHS/s", "cores": threads, "kind": "port",
               "sample": "%d right-hand sides of the same 512^2 Brusselator in the C++ oracle (CPU restatement, "
                         "not MethodOfLines.jl), FP64, 1 thread" % reps}

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
                   "graph": bool(args.graph)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "rhs_generic_kernel<double> (mol_rhs)", "avg_us": rhs_us,
                     "bytes_per_launch": rhs_bytes, "rhs_gpoint_s": rhs_gpts},
        "cpu_baseline": cpu,
        "wall_s": wall,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
