"""GPU: the allocation-free gradient entry point (pdfc_mlp_gradient_noalloc) against the
allocating one (pdfc_mlp_gradient), through the C ABI.

Both run the same kernels in the same summation order, so the gradient must be bit-identical
and the mean loss equal. Two calls in a row, the second smaller than the first, show that the
shared fixed array left over from a larger call does not leak into a smaller one; a size
above the maximum is rejected by the host before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest

from climatenet_amd import abi

from .test_oracle_mlp import random_norm, random_params

pytestmark = pytest.mark.gpu


def _entry(lib):
    f = lib.pdfc_mlp_gradient_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return f


def test_noalloc_gradient_equals_allocating_gradient(gpu):
    import torch

    L = gpu.lib()
    noalloc = _entry(L)
    max_rows = 2048  # PDFC_MLP_NOALLOC_MAX_ROWS
    mlp = gpu.ContactMLP(seed=3)
    mlp.set_params(random_params(31), random_norm(32))
    rng = np.random.default_rng(33)
    for rows in (max_rows, 200):  # the largest call, then a smaller one with a ragged tile
        x = rng.normal(size=(rows, abi.MLP_FEATURES)).astype(np.float32)
        t = rng.normal(size=(rows, abi.MLP_OUTPUTS)).astype(np.float32)
        g_ref, loss_ref = mlp.gradient(x, t)
        dx = torch.from_numpy(x).to("cuda:0")
        dt = torch.from_numpy(t).to("cuda:0")
        dg = torch.full((abi.MLP_PARAMS,), float("nan"), dtype=torch.float32, device="cuda:0")
        dl = torch.full((1,), float("nan"), dtype=torch.float32, device="cuda:0")
        torch.cuda.synchronize()
        print("pdfc_mlp_gradient_noalloc rows", rows, flush=True)
        rc = noalloc(mlp.h, dx.data_ptr(), dt.data_ptr(), rows, dg.data_ptr(), dl.data_ptr())
        assert rc == abi.PDFC_OK, L.pdfc_last_error().decode()
        torch.cuda.synchronize()
        g = dg.cpu().numpy()
        loss = float(dl.cpu().numpy()[0])
        assert np.array_equal(g, g_ref), np.abs(g - g_ref).max()
        assert loss == loss_ref, (loss, loss_ref)
    # above the maximum: rejected on the host, outputs untouched
    dg.fill_(-1.0)
    torch.cuda.synchronize()
    rc = noalloc(mlp.h, dx.data_ptr(), dt.data_ptr(), max_rows + 1, dg.data_ptr(), dl.data_ptr())
    assert rc == abi.PDFC_ERR_INVALID
    assert "maximum" in L.pdfc_last_error().decode()
    assert bool((dg == -1.0).all())
    mlp.close()
