"""CPU checks of the Python front end's argument validation: vectors handed to
the library as raw device pointers must be contiguous float64 tensors on the
problem's device with a whole number of vectors, or a ValueError is raised
before any pointer reaches the library (no GPU needed: every case here is
refused on the host)."""
import types

import numpy as np
import pytest
import torch

import dpdhip
from dpdhip import _check_vectors


def _fake_problem(dim=10):
    # a Problem shell: the checks run before the library is called
    p = dpdhip.Problem.__new__(dpdhip.Problem)
    p.dim = dim
    p.ctx = types.SimpleNamespace(device=0, h=None)
    p.h = None
    p.version = 0
    return p


def test_vectors_on_the_host_are_refused():
    p = _fake_problem()
    x = torch.zeros(20, dtype=torch.float64)
    for call in (lambda: p.sigma(x), lambda: p.sigma_t(x), lambda: p.davidson(2, R=x),
                 lambda: p.grad(x[:10], x[:10], out={})):
        with pytest.raises(ValueError, match="float64 tensor on cuda:0"):
            call()
    with pytest.raises(ValueError):
        dpdhip.sigma_group([p], [x])
    with pytest.raises(ValueError):
        dpdhip.davidson_batch([p], 2, Rs=[x])


def test_check_vectors_rules():
    x = torch.zeros(20, dtype=torch.float32)
    with pytest.raises(ValueError, match="float64"):
        _check_vectors(x, "R", 10, 0)
    with pytest.raises(ValueError, match="float64"):
        _check_vectors(np.zeros(20), "R", 10, 0)
    # strided views and wrong lengths are refused even before the device test
    # would matter; emulate a device tensor with a stand-in object
    class Dev:
        def __init__(self, n, contiguous=True, dtype=torch.float64, index=0):
            self._n, self._c = n, contiguous
            self.dtype, self.is_cuda = dtype, True
            self.device = types.SimpleNamespace(index=index)

        def numel(self):
            return self._n

        def is_contiguous(self):
            return self._c

    import dpdhip as D
    real = D._torch

    class FakeTorch:
        Tensor = Dev
        float64 = torch.float64

    D._torch = lambda: FakeTorch
    try:
        assert _check_vectors(Dev(30), "R", 10, 0) == 3
        assert _check_vectors(Dev(30), "R", 10, 0, nvec=3) == 3
        for bad, msg in ((Dev(25), "whole number"), (Dev(0), "whole number"), (Dev(30, contiguous=False), "contig"),
                         (Dev(30, index=1), "cuda:0"), (Dev(30, dtype=torch.float32), "float64")):
            with pytest.raises(ValueError, match=msg):
                _check_vectors(bad, "R", 10, 0)
        with pytest.raises(ValueError, match="expected 2 vector"):
            _check_vectors(Dev(30), "R", 10, 0, nvec=2)
    finally:
        D._torch = real
