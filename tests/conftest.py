import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "methodoflines.jl_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no HIP device in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(params=["generated", "interpreter"])
def rhs_path(request, monkeypatch):
    """Run a right-hand-side test on both device paths of the core tiles: the kernel generated
    from the stencil IR (compiled at run time) and the bytecode interpreter kernel."""
    monkeypatch.setenv("MOL_JIT", "1" if request.param == "generated" else "0")
    return request.param


def assert_parity(got, ref, scale=None, rtol=1e-12, what=""):
    """FP64 parity: norm-wise relative error <= rtol, and max error <= rtol * scale, where
    scale bounds the magnitude of the summed terms (cancellation-aware)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    assert np.all(np.isfinite(got)), "non-finite values " + what
    err = np.abs(got - ref)
    nref = np.linalg.norm(ref)
    if nref > 0:
        rel = np.linalg.norm(got - ref) / nref
        assert rel <= rtol, "%s: norm-wise relative error %.3e > %.1e" % (what, rel, rtol)
    if scale is None:
        scale = max(np.max(np.abs(ref)), 1e-300)
    assert err.max() <= rtol * scale, "%s: max error %.3e > %.1e * %.3e" % (what, err.max(), rtol, scale)
