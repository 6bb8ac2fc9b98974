"""The allocation-free slice wrap (lqmc_wrap_green_static, a static device workspace)
against lqmc_wrap_green with a caller workspace, through the C ABI: bitwise equal on two
calls in a row, the second smaller than the first (the shared workspace keeps the first
call's contents), both within 1e-11 of B^-1 G B in numpy; sizes past the 16 MiB workspace
are refused before anything runs."""
import ctypes

import numpy as np
import pytest

from latticeqmc_amd import _native as nat
from oracle import model

pytestmark = pytest.mark.gpu


def _inputs(n, count, seed):
    side = int(round(n ** 0.5))
    k = model.square_k(side, side)
    _, lam, ek, eki = model.params(k, 8.0, 80, 4.0)
    rs = np.random.RandomState(seed)
    v = np.where(rs.rand(count, n) < 0.5, np.exp(-lam), np.exp(lam))
    g = rs.rand(count, n, n)
    return ek, eki, v, g


@pytest.fixture(scope="module")
def lib():
    return nat.load()


def _run(lib, n, count, seed):
    import torch
    ek, eki, v, g = _inputs(n, count, seed)
    want = np.array([np.linalg.solve(v[m][:, None] * ek, g[m] @ (v[m][:, None] * ek)) for m in range(count)])
    a = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (ek, eki, v, g)]
    b = [t.clone() for t in a]
    ws = torch.empty(lib.lqmc_wrap_green_workspace(n, count), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    print(f"[no-alloc] static wrap n={n} count={count}", flush=True)
    nat.check(lib.lqmc_wrap_green_static(0, n, count, a[0].data_ptr(), a[1].data_ptr(), a[2].data_ptr(),
                                         a[3].data_ptr(), None), "lqmc_wrap_green_static")
    nat.check(lib.lqmc_wrap_green(0, n, count, b[0].data_ptr(), b[1].data_ptr(), b[2].data_ptr(), b[3].data_ptr(),
                                  ws.data_ptr(), ws.numel(), None), "lqmc_wrap_green")
    torch.cuda.synchronize()
    got, ref = a[3].cpu().numpy(), b[3].cpu().numpy()
    assert np.array_equal(got, ref)
    assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()


def test_static_wrap_equals_workspace_wrap(lib):
    _run(lib, 64, 3, 1)
    _run(lib, 16, 2, 2)   # smaller second call on the same static workspace


def test_static_wrap_refuses_oversized(lib):
    import torch
    t = torch.zeros(1, dtype=torch.float64, device="cuda")
    p = t.data_ptr()
    rc = lib.lqmc_wrap_green_static(0, 1024, 2, p, p, p, p, None)
    assert rc == -1 and b"16 MiB" in lib.lqmc_last_error()
    assert lib.lqmc_wrap_green_static(0, 0, 1, p, p, p, p, None) == -1
