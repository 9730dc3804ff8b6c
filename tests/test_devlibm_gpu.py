"""The device math functions (ww3gpu_devlibm_eval) against the host glibc (through the
oracle library) on a large sample of float inputs, on the GPU."""
import numpy as np
import pytest

import oracle as O
from ww3gpu import capi

pytestmark = pytest.mark.gpu


def device_eval(torch, fn, x):
    X = torch.from_numpy(x).cuda()
    Y = torch.empty_like(X)
    rc = capi.lib().ww3gpu_devlibm_eval(fn, X.data_ptr(), None, Y.data_ptr(), X.numel(),
                                        torch.cuda.current_stream().cuda_stream)
    assert rc == capi.Status.OK
    torch.cuda.synchronize()
    return Y.cpu().numpy()


def test_expf_sample(gpu_runtime):
    torch = gpu_runtime
    rng = np.random.default_rng(123)
    u = rng.integers(0, 2 ** 32, 1 << 24, dtype=np.uint64).astype(np.uint32)
    dense = np.linspace(-110.0, 95.0, 1 << 22, dtype=np.float32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 88.72283, 88.72284, -103.97208,
                        -103.27893, 1e-45, -1e-45, float.fromhex("0x1.62e42ep6"),
                        -float.fromhex("0x1.9fe368p6")], np.float32)
    x = np.concatenate([u.view(np.float32), dense, special])
    y = device_eval(torch, 0, x)
    ref = O.expf(x)
    same = (y.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(y) & np.isnan(ref))
    assert same.all(), f"{(~same).sum()} differ, first x={x[~same][0]!r}"
