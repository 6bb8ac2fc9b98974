"""Job-layer parity on the GPU: ``measure`` (the chain pool of lqmc/multiprocessing.py:252-289),
``measure_betas`` (one job per beta, multiprocessing.py:292-341) and ``measure_scan`` (the
U x beta workflow of comp.py:19-47) run their chains in one batched launch; each result
must equal, bitwise in FP64, the same chains run one at a time through the engine with the
seeds the job layer assigns."""
import numpy as np
import pytest

from latticeqmc_amd import HubbardModel, measure, measure_betas, measure_scan
from latticeqmc_amd.batched import _model_spec
from latticeqmc_amd.engine import ChainSpec, SweepEngine

pytestmark = pytest.mark.gpu

KW = dict(n_stab=8, delay=4)


def _alone(spec, warmup, sweeps):
    eng = SweepEngine([spec], mode="dqmc", precision="fp64", rng="philox", **KW)
    eng.sweep(warmup, measure=False)
    eng.sweep(sweeps, measure=True)
    eng.sync()
    g = eng.green(0, which=1)
    eng.close()
    return g


def _model(u=4.0, mu=0.0):
    m = HubbardModel(u=u, mu=mu)
    m.build_square(4)
    return m


def test_measure_betas_equals_chains_alone(tmp_path):
    print("[jobs parity] measure_betas", flush=True)
    m = _model()
    betas = [1.0, 2.0, 3.0]
    up, dn = measure_betas(m, betas, 16, 5, 12, seed=7, tmp_file=str(tmp_path / "tmp.npz"), **KW)
    for i, b in enumerate(betas):
        g = _alone(_model_spec(m, b, 16, 7 + i), 5, 12)
        assert np.array_equal(up[i], g[0]) and np.array_equal(dn[i], g[1]), b


def test_measure_pool_equals_mean_of_chains():
    print("[jobs parity] measure", flush=True)
    m = _model()
    gu, gd = measure(m, 2.0, 16, 5, 16, cores=4, seed=3, **KW)
    gs = [_alone(_model_spec(m, 2.0, 16, 3 + c), 5, 4) for c in range(4)]
    want = np.mean(gs, axis=0)
    assert np.array_equal(gu, want[0]) and np.array_equal(gd, want[1])


def test_measure_scan_equals_chains_alone(tmp_path):
    print("[jobs parity] measure_scan", flush=True)
    m = _model()
    u_values, betas = [2.0, 4.0], [1.0, 2.0]
    paths = measure_scan(m, u_values, betas, 16, 5, 10, folder=str(tmp_path), seed=11, **KW)
    for j, u in enumerate(u_values):
        m2 = _model()
        m2.set_params(u=u)  # t = 1, mu = U/2, as comp.py sets each U
        k = m2.ham_kinetic()
        with np.load(paths[u], allow_pickle=False) as d:
            for i, b in enumerate(betas):
                s = 11 + j * len(betas) + i
                rs = np.random.RandomState(s)
                field = (2 * rs.randint(0, 2, size=(16, 16)) - 1).astype(np.int8)
                g = _alone(ChainSpec(k, b, 16, u, field, seed=s), 5, 10)
                assert np.array_equal(d["gf_up"][i], g[0]) and np.array_equal(d["gf_dn"][i], g[1]), (u, b)


def test_process_managers_equal_chains_alone(tmp_path, monkeypatch):
    """ParallelProcessManager: sweeps // procs per chain and the remainder on chain 0, the
    unweighted mean of the chain means (multiprocessing.py:252-267); SerialProcessManager:
    one chain per beta, equal to measure_betas (multiprocessing.py:292-341)."""
    print("[jobs parity] process managers", flush=True)
    from latticeqmc_amd import ParallelProcessManager, SerialProcessManager
    monkeypatch.chdir(tmp_path)
    m = _model()
    pm = ParallelProcessManager(m, 2.0, 16, 5, procs=4, seed=2, **KW)
    pm.set_jobs(18)
    pm.run()
    gs = [_alone(_model_spec(m, 2.0, 16, 2 + c), 5, 6 if c == 0 else 4) for c in range(4)]
    assert np.array_equal(pm.get_result(), np.sum(gs, axis=0) / 4)
    sm = SerialProcessManager(m, 16, 5, 8, procs=2, seed=9, **KW)
    sm.set_jobs([1.0, 2.0])
    sm.run()
    res = sm.get_result()
    assert res.shape == (2, 2, 16, 16) and not (tmp_path / "tmp_gf_series.npz").exists()
    for i, b in enumerate([1.0, 2.0]):
        assert np.array_equal(res[i], _alone(_model_spec(m, b, 16, 9 + i), 5, 8))
