"""GPU: mol_integrate_noalloc, the integration entry point that allocates nothing (stage vectors
and device clock in the library's per-device scratch array), called through the C ABI, against
mol_integrate on the same inputs."""
import ctypes as C

import numpy as np
import pytest

import molgpu
from molgpu import native, systems

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _integrate(prob, u, t0, dt, nsteps, scheme, flags, noalloc):
    torch = _torch()
    h = prob.handle
    p = prob.device_params()
    fn = h.lib.mol_integrate_noalloc if noalloc else h.lib.mol_integrate
    stream = torch.cuda.current_stream().cuda_stream
    return fn(h.h, u.data_ptr(), p.data_ptr() if p is not None else None, float(t0), float(dt), int(nsteps),
              scheme, flags, C.c_void_p(stream))


def _both(prob, u0, t0, dt, nsteps, scheme, flags, dtype):
    torch = _torch()
    out = {}
    for noalloc in (False, True):
        u = torch.tensor(u0, dtype=dtype, device="cuda")
        print("  %s noalloc=%s scheme=%d flags=%d unknowns=%d" % (prob.lowered.fields, noalloc, scheme, flags,
                                                                  len(u0)), flush=True)
        rc = _integrate(prob, u, t0, dt, nsteps, scheme, flags, noalloc)
        native.check(prob.handle.lib, rc, prob.handle.h)
        torch.cuda.synchronize()
        out[noalloc] = u.cpu().numpy()
    return out


def test_noalloc_equals_integrate_on_consecutive_calls():
    """Three calls in a row, each on a smaller input than the one before: RK4 stage launches of
    the FP64 Brusselator at 129^2 (y1, y2, acc in the scratch), then FP32 diffusion-reaction at
    41^2 with the compensated update (the compensation must start from zero although the
    previous call left stage values there), then the fused RK4 step at 33^2.  Each equals
    mol_integrate with the handle's own workspace, value for value."""
    torch = _torch()
    prob = molgpu.discretize(*systems.brusselator2d(n=129))
    u0 = np.asarray(prob.u0, dtype=np.float64)
    r = _both(prob, u0, 1.05, 2e-6, 37, native.MOL_RK4, native.MOL_FLAG_UNFUSED, torch.float64)
    assert np.all(np.isfinite(r[True]))
    assert np.array_equal(r[True], r[False])

    prob = molgpu.discretize(*systems.diffreact2d(41, 0.05, 2.5, (0, 1, 1, 0), dtype="f32"))
    u0 = np.random.default_rng(2).uniform(0, 1, prob.lowered.n_unknowns).astype(np.float32)
    r = _both(prob, u0, 0.0, 1e-4, 25, native.MOL_RK4, native.MOL_FLAG_KAHAN, torch.float32)
    assert np.all(np.isfinite(r[True]))
    assert np.array_equal(r[True], r[False])

    prob = molgpu.discretize(*systems.brusselator2d(n=33))
    assert prob.plan()["fused_tiles"] > 0, prob.plan()
    u0 = np.asarray(prob.u0, dtype=np.float64)
    r = _both(prob, u0, 0.0, 1e-5, 9, native.MOL_RK4, 0, torch.float64)
    assert np.array_equal(r[True], r[False])
    r = _both(prob, u0, 0.0, 1e-5, 9, native.MOL_SSPRK33, 0, torch.float64)
    assert np.array_equal(r[True], r[False])


def test_noalloc_rejects_graphs_and_oversized_integrations():
    """MOL_FLAG_GRAPH and integrations whose stage vectors exceed the scratch are refused before
    anything is enqueued; the state is left untouched."""
    torch = _torch()
    prob = molgpu.discretize(*systems.brusselator2d(n=33))
    u = torch.tensor(prob.u0, dtype=torch.float64, device="cuda")
    before = u.clone()
    rc = _integrate(prob, u, 0.0, 1e-5, 4, native.MOL_RK4, native.MOL_FLAG_GRAPH, True)
    assert rc == -1
    # 2 x 2048^2 FP64 unknowns: three stage vectors of 64 MiB each
    big = molgpu.discretize(*systems.brusselator2d(n=2049))
    ub = torch.zeros(big.lowered.n_unknowns, dtype=torch.float64, device="cuda")
    rc = _integrate(big, ub, 0.0, 1e-5, 4, native.MOL_RK4, native.MOL_FLAG_UNFUSED, True)
    assert rc == -2
    assert "bytes of scratch" in big.handle.lib.mol_last_error(big.handle.h).decode()
    torch.cuda.synchronize()
    assert torch.equal(u, before)
    assert float(ub.abs().max()) == 0.0
