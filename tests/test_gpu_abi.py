"""C-ABI behaviour on the GPU: device-to-device Green's function copies, sweeps on a
caller stream, the per-recompute deviation record and the precision flag."""
import numpy as np
import pytest

from gpu_cases import spec_from, square_case
from latticeqmc_amd._native import LqmcError
from latticeqmc_amd.engine import SweepEngine

pytestmark = pytest.mark.gpu


def _engine(precision="fp64", **kw):
    k, f = square_case(4, 16, 2.0, 4.0, 3)
    return SweepEngine([spec_from(k, 2.0, 16, 4.0, f, seed=5)], mode="dqmc", precision=precision, delay=4,
                       n_stab=8, rng="philox", **kw)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_green_dev_equals_host_copy(precision):
    print(f"[green_dev] {precision}", flush=True)
    import torch
    eng = _engine(precision, max_dg=-1.0)  # FP32 at n_stab = 8 drifts past the default flag bound
    eng.sweep(3, measure=True)
    eng.sync()
    out = torch.full((2, 16, 16), float("nan"), dtype=torch.float64, device="cuda")
    for which in (0, 1, 2):
        eng.green_dev(0, out.data_ptr(), which)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), eng.green(0, which)), which
    eng.close()


def test_sweep_on_caller_stream_is_seen_by_getters():
    print("[caller stream]", flush=True)
    import torch
    a, b = _engine(), _engine()
    s = torch.cuda.Stream()
    for _ in range(3):
        a.sweep(1, measure=True, stream=s.cuda_stream)
        b.sweep(1, measure=True)
    # no explicit synchronisation: the getters wait on the context stream, which waits
    # for the sweep on the caller stream
    assert np.array_equal(a.green(0, which=2), b.green(0, which=2))
    assert np.array_equal(a.field(0), b.field(0))
    a.sync()
    b.sync()
    a.close()
    b.close()


def test_recompute_record_and_precision_flag():
    print("[precision flag]", flush=True)
    k, f = square_case(8, 40, 4.0, 4.0, 0)
    spec = spec_from(k, 4.0, 40, 4.0, f, seed=1)
    strict = SweepEngine([spec], mode="dqmc", precision="fp32", delay=8, n_stab=10, rng="philox", max_dg=1e-12)
    strict.sweep(1, measure=True)
    with pytest.raises(LqmcError, match="0x20"):
        strict.sync()
    rec = strict.recompute_errors(0)
    assert rec.shape == (4,) and rec.max() > 1e-12 and rec.max() == strict.stats(0)["max_err"]
    strict.close()
    off = SweepEngine([spec], mode="dqmc", precision="fp32", delay=8, n_stab=10, rng="philox", max_dg=-1.0)
    off.sweep(1, measure=True)
    off.sync()
    assert off.stats(0)["flags"] == 0
    off.reset_stats()
    assert not off.recompute_errors(0).any()
    off.close()


@pytest.mark.parametrize("n,count", [(16, 2), (64, 3), (100, 2), (256, 4)])
def test_wrap_green_matches_numpy(n, count):
    """lqmc_wrap_green against B^-1 G B in numpy (B = diag(v) E_K) on random G; the E_K of a
    square lattice at dtau = 0.1 and exp(-+lambda) diagonals of a random field."""
    print(f"[wrap] n={n} count={count}", flush=True)
    import torch
    from latticeqmc_amd.engine import wrap_green
    from oracle import model
    side = int(round(n ** 0.5))
    k = model.square_k(side, side)
    _, lam, ek, eki = model.params(k, 8.0, 80, 4.0)
    rs = np.random.RandomState(n + count)
    v = np.where(rs.rand(count, n) < 0.5, np.exp(-lam), np.exp(lam))
    g = rs.rand(count, n, n)
    want = np.array([np.linalg.solve(v[m][:, None] * ek, g[m] @ (v[m][:, None] * ek)) for m in range(count)])
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (ek, eki, v, g)]
    wrap_green(*dev)
    torch.cuda.synchronize()
    got = dev[3].cpu().numpy()
    assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()
