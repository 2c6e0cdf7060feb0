"""GPU: the device error norms (max |u - ref|, ||u - ref||_2, ||ref||_2) against numpy on the same
state, in both state precisions, with and without a reference, from a single unknown up to the
Brusselator 512^2 state; and the final-time error check of the heat equation made on the device."""
import numpy as np
import pytest

import molgpu
from molgpu import systems

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _host_norms(u, ref):
    d = u.astype(np.float64) - (ref if ref is not None else 0.0)
    r = ref if ref is not None else np.zeros_like(d)
    return np.max(np.abs(d)), np.sqrt(np.sum(d * d)), np.sqrt(np.sum(r * r))


def _check(prob, u_host, ref_host):
    torch = _torch()
    u = torch.tensor(u_host, dtype=prob.torch_dtype(), device="cuda")
    ref = torch.tensor(ref_host, dtype=torch.float64, device="cuda") if ref_host is not None else None
    got = prob.error_norms(u, ref).cpu().numpy()
    want = _host_norms(u.cpu().numpy(), ref_host)
    # the maximum is exact; the sums differ from numpy's pairwise order by rounding only
    assert got[0] == want[0]
    assert abs(got[1] - want[1]) <= 1e-13 * want[1] + 1e-300
    assert abs(got[2] - want[2]) <= 1e-13 * want[2] + 1e-300


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("with_ref", [True, False])
def test_norms_heat(dtype, with_ref):
    sys_, disc = systems.heat1d(n=101, dtype=dtype)
    prob = molgpu.discretize(sys_, disc)
    rng = np.random.default_rng(3)
    n = prob.lowered.n_unknowns
    print("  norms heat %s ref=%s n=%d" % (dtype, with_ref, n), flush=True)
    _check(prob, rng.uniform(-1, 1, n), rng.uniform(-1, 1, n) if with_ref else None)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_norms_brusselator512(dtype):
    """1024 partial blocks, each striding over many values: every value is visited once."""
    sys_, disc = systems.brusselator2d(n=513, dtype=dtype)
    prob = molgpu.discretize(sys_, disc)
    rng = np.random.default_rng(4)
    n = prob.lowered.n_unknowns
    print("  norms brusselator %s n=%d" % (dtype, n), flush=True)
    u = rng.standard_normal(n)
    ref = u + 1e-3 * rng.standard_normal(n)
    ref[n // 3] += 7.0  # a single large difference, far from the first block
    _check(prob, u, ref)


def test_norms_single_unknown_and_zero_state():
    sys_, disc = systems.heat1d(n=3)
    prob = molgpu.discretize(sys_, disc)
    assert prob.lowered.n_unknowns == 1
    print("  norms single unknown", flush=True)
    _check(prob, np.array([-2.5]), np.array([0.5]))
    _check(prob, np.zeros(1), None)


def test_norms_reject_wrong_reference():
    torch = _torch()
    sys_, disc = systems.heat1d(n=11)
    prob = molgpu.discretize(sys_, disc)
    u = torch.zeros(prob.lowered.n_unknowns, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        prob.error_norms(u, torch.zeros(3, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        prob.error_norms(u, torch.zeros(prob.lowered.n_unknowns, dtype=torch.float32, device="cuda"))


def test_heat_final_time_error_on_device():
    """u_t = u_xx, Dirichlet 0, u0 = sin(pi x): at t = 0.1 the device norms of u - exp(-pi^2 t) sin(pi x)
    equal the host ones and lie within the reference test's tolerance (atol 0.01,
    test/Diffusion/MOL_1D_Linear_Diffusion.jl:82)."""
    torch = _torch()
    n, dt, N = 101, 1e-5, 10_000
    sys_, disc = systems.heat1d(n=n)
    prob = molgpu.discretize(sys_, disc)
    u = torch.tensor(prob.u0, dtype=torch.float64, device="cuda")
    print("  heat final-time error", flush=True)
    prob.integrate(u, 0.0, dt, N, molgpu.RK4)
    x = np.arange(1, n - 1) / (n - 1)
    exact = np.exp(-np.pi ** 2 * 0.1) * np.sin(np.pi * x)
    got = prob.error_norms(u, torch.tensor(exact, device="cuda")).cpu().numpy()
    want = _host_norms(u.cpu().numpy(), exact)
    assert got[0] == want[0]
    assert abs(got[1] - want[1]) <= 1e-13 * want[1]
    assert got[0] < 0.01 and got[0] < 5e-5
