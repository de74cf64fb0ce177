"""The matrix-core lane layout the batched element kernel assumes, checked on
the device: C = A B for 16x16 matrices through 16x16x4 MFMA tiles, in FP32
and FP64, against numpy."""

import numpy as np
import pytest
import torch

from hybrid import _native as nat

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,code,tol", [(torch.float64, 0, 1e-14), (torch.float32, 1, 1e-5)])
def test_mfma_tile_layout(dtype, code, tol):
    rng = np.random.default_rng(0)
    A = rng.standard_normal((16, 16))
    B = rng.standard_normal((16, 16))
    a = torch.as_tensor(A, dtype=dtype, device="cuda")
    b = torch.as_tensor(B, dtype=dtype, device="cuda")
    c = torch.zeros(16, 16, dtype=dtype, device="cuda")
    nat.check(nat.lib().fsg_mfma_selftest(code, nat.ptr(a), nat.ptr(b), nat.ptr(c),
                                          nat.stream_handle(c.device)))
    got = c.double().cpu().numpy()
    np.testing.assert_allclose(got, A @ B, rtol=0, atol=tol * np.abs(A).sum(1).max() * 4)
