"""GPU: two variables joined by an interface condition (interface_boundary.jl:79-111,
generate_bc_eqs.jl:35-58) against the oracle, and the reference's interface diffusion test."""
import numpy as np
import pytest

import oracle as orc
import molgpu
from molgpu import native
from conftest import assert_parity
from test_interface import interface_system

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


@pytest.mark.parametrize("p", [2, 4])
def test_interface_rhs_matches_oracle(p):
    torch = _torch()
    n1, n2, dx = 65, 97, 0.5 / 64
    prob = molgpu.discretize(*interface_system(n1, n2, D1=0.4, D2=1.3, gl=0.25, gr=-0.5, p=p))
    u = np.random.default_rng(p).uniform(-1, 1, prob.lowered.n_unknowns)
    ref = orc.interface1d_rhs(n1, n2, dx, 0.4, 1.3, u, p=p, gl=0.25, gr=-0.5)
    for tile in [(16, 1, 1), None]:
        if tile:
            prob.set_tile(*tile)
        ug = torch.tensor(u, device="cuda")
        du = torch.empty_like(ug)
        print("  interface p=%d tile=%s" % (p, tile), flush=True)
        prob.f(du, ug, None, 0.0)
        torch.cuda.synchronize()
        assert_parity(du.cpu().numpy(), ref, rtol=1e-12, what="interface p=%d" % p)


def test_interface_diffusion_decays_like_the_reference_test():
    """Test 14 of test/Diffusion/MOL_1D_Linear_Diffusion.jl: c1 on [0, 0.5], c2 on [0.5, 1], 10
    nodes each, c1(t, 0.5) ~ c2(t, 0.5), zero Dirichlet ends; at t = 1 both are 0 to 1e-3.  RK4
    at a stable fixed step, equal to RK4 over the oracle's right-hand side."""
    torch = _torch()
    n = 10
    dx = 0.5 / (n - 1)
    prob = molgpu.discretize(*interface_system(n, n))
    assert prob.lowered.n_unknowns == (n - 1) + (n - 2)
    dt, N = 5e-4, 2000
    u = torch.tensor(prob.u0, device="cuda")
    print("  interface diffusion %d steps" % N, flush=True)
    prob.integrate(u, 0.0, dt, N, molgpu.RK4, graph=True)
    torch.cuda.synchronize()
    got = u.cpu().numpy()
    assert np.max(np.abs(got)) < 1e-3
    w = np.asarray(prob.u0, dtype=np.float64)
    F = lambda x: orc.interface1d_rhs(n, n, dx, 1.0, 1.0, x)
    for _ in range(N):
        k1 = F(w); k2 = F(w + dt / 2 * k1); k3 = F(w + dt / 2 * k2); k4 = F(w + dt * k3)
        w = w + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    assert np.max(np.abs(got - w)) <= 1e-12 * max(1.0, np.max(np.abs(prob.u0)))


def test_interface_paths_not_built_are_refused():
    """The reverse mode and the full-grid expansion of interface systems are refused with
    MOL_ERR_UNSUPPORTED before any launch; the integration takes the stage kernels."""
    torch = _torch()
    prob = molgpu.discretize(*interface_system(17, 17))
    assert prob.plan()["fused_tiles"] == 0
    u = torch.tensor(prob.u0, device="cuda")
    with pytest.raises(native.MolError) as e:
        prob.vjp(u, torch.ones_like(u))
    assert e.value.code == -4
    with pytest.raises(native.MolError) as e:
        prob.expand(u, 0.0)
    assert e.value.code == -4
