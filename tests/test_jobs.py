"""Job layer on the CPU with stub runners in place of the GPU: the temporary-results file
of measure_betas (keys beta, data) with resume after an interruption, and the U x beta
scan's u=<U>.npz files with skip-on-exist (lqmc/multiprocessing.py:292-341, comp.py:19-47)."""
import os

import numpy as np
import pytest

from latticeqmc_amd import HubbardModel, get_datadir, get_datapath, measure_betas, measure_scan


def _model():
    m = HubbardModel(u=4.0)
    m.build_square(2)
    return m


def test_measure_betas_caches_and_resumes(tmp_path):
    m = _model()
    betas = np.array([0.5, 1.0, 2.0, 4.0])
    tmp = str(tmp_path / "tmp_gf_series.npz")
    calls = []

    def stub(model, bs, L, warmup, sweeps, seeds, fail_after=None):
        calls.append(list(bs))
        if fail_after is not None and len(calls) > fail_after:
            raise RuntimeError("interrupted")
        return np.array([np.full((2, 4, 4), b) + np.arange(2)[:, None, None] for b in bs])

    with pytest.raises(RuntimeError):
        measure_betas(m, betas, 8, 1, 1, batch=2, tmp_file=tmp,
                      runner=lambda *a, **k: stub(*a, fail_after=1))
    with np.load(tmp, allow_pickle=False) as d:
        assert np.array_equal(d["beta"], betas)
        assert not np.isnan(d["data"][:2]).any() and np.isnan(d["data"][2:]).all()
    calls.clear()
    up, dn = measure_betas(m, betas, 8, 1, 1, batch=2, tmp_file=tmp, runner=stub)
    assert calls == [[2.0, 4.0]]                       # only the pending betas ran
    assert not os.path.exists(tmp)                     # removed once complete
    assert up.shape == (4, 4, 4) and dn.shape == (4, 4, 4)
    assert np.allclose(up[:, 0, 0], betas) and np.allclose(dn[:, 0, 0], betas + 1)


def test_measure_betas_ignores_pickled_cache(tmp_path):
    m = _model()
    tmp = str(tmp_path / "tmp_gf_series.npz")
    np.savez(tmp, beta=np.array([1.0]), data=np.array([None], dtype=object))
    up, dn = measure_betas(m, [1.0], 8, 1, 1, tmp_file=tmp,
                           runner=lambda model, bs, *a, **k: np.zeros((len(bs), 2, 4, 4)))
    assert up.shape == (1, 4, 4)


def test_measure_scan_writes_and_skips(tmp_path):
    m = _model()
    betas = [1.0, 2.0]
    seen = []

    seeds = {}

    def stub(specs, warmup, sweeps):
        seen.append([(s.u, s.beta) for s in specs])
        for s in specs:
            seeds[(s.u, s.beta)] = s.seed
        return np.array([np.full((2, 4, 4), 10 * s.u + s.beta) for s in specs])

    folder = str(tmp_path / "scan")
    paths = measure_scan(m, [1.0, 4.0], betas, 8, 2, 3, folder=folder, runner=stub)
    assert seen == [[(1.0, 1.0), (1.0, 2.0), (4.0, 1.0), (4.0, 2.0)]]   # one batched run
    with np.load(paths[4.0], allow_pickle=False) as d:
        assert set(d.files) == {"beta", "gf_up", "gf_dn"}
        assert np.allclose(d["gf_up"][:, 0, 0], [41.0, 42.0])
    assert os.path.basename(paths[1.0]) == "u=1.0.npz"
    seen.clear()
    measure_scan(m, [1.0, 4.0, 6.0], betas, 8, 2, 3, folder=folder, runner=stub)
    assert seen == [[(6.0, 1.0), (6.0, 2.0)]]          # existing U files are skipped
    # each (U, beta) pair keeps the seed of its place in the whole scan, so a resumed scan
    # runs the same chains as one uninterrupted scan and no two pairs share a stream
    assert seeds == {(1.0, 1.0): 0, (1.0, 2.0): 1, (4.0, 1.0): 2, (4.0, 2.0): 3, (6.0, 1.0): 4, (6.0, 2.0): 5}
    assert m.u == 6.0 and m.mu == 3.0 and m.t == 1.0   # set_params(u=U) as comp.py


def test_data_paths(tmp_path):
    m = _model()
    d = get_datadir(m.lattice.shape, 50, 500, 5000, root=str(tmp_path))
    assert d.endswith(os.path.join("", "2x2_nt=50_warm=500_meas=5000")) and os.path.isdir(d)
    p = get_datapath("gf", m, post="_tmp", root=str(tmp_path), beta=2.0)
    assert p.endswith(os.path.join("u=4.0_t=1_mu=2.0_w=2_h=2", "gf beta=2.0_tmp.npz"))
