"""Host-side parts of the drop-in backend (no GPU needed)."""
import types

import numpy as np
import pytest
import torch

from oracle import fd as ofd


def test_radius_inferred_from_laplacian():
    from parsec_python_amd.backend import _radius_of
    for r in (1, 2, 3, 4, 6):
        A = ofd.laplacian(16, 16, 16, r)
        assert _radius_of(A) == r


def test_to_numpy_accepts_torch():
    from parsec_python_amd.backend import SolverBackend
    be = SolverBackend(*(["gpu"] + [None] * 8 + [torch]))
    t = torch.arange(4, dtype=torch.float64)
    assert isinstance(be.to_numpy_array(t), np.ndarray)
    assert be.to_host_scalar(torch.tensor(2.5)) == 2.5
    assert be.to_host_scalar(np.float64(1.5)) == 1.5
    a = np.ones(3)
    assert be.to_numpy_array(a) is a


def test_select_backend_cpu_request_rejected():
    from parsec_python_amd.backend import select_solver_backend
    with pytest.raises(ValueError):
        select_solver_backend(types.SimpleNamespace(use_gpu=0))


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the behaviour without a GPU")
def test_select_backend_without_gpu_fails_loudly():
    from parsec_python_amd.backend import select_solver_backend
    with pytest.raises((SystemExit, RuntimeError)):
        select_solver_backend(types.SimpleNamespace(use_gpu=1))


def test_backend_options_validated():
    """Unknown precision or Hartree preconditioner names are rejected before any GPU work."""
    from parsec_python_amd.backend import GpuBackendImpl
    with pytest.raises(ValueError):
        GpuBackendImpl("fp16")
    with pytest.raises(ValueError):
        GpuBackendImpl("mixed", pcg_prec="ilu")
    impl = GpuBackendImpl("mixed", pcg_prec="mg")
    assert impl.pcg_prec == "mg" and impl.ctx is None
