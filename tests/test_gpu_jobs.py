"""Job layer on the GPU: measure_betas and the U x beta scan run their chains in one
batched launch.  The interaction is decoupled in the particle-hole symmetric form
U (n_up - 1/2)(n_dn - 1/2) while K carries -mu, so mu = 0 is half filling (filling 1/2 per
spin on the bipartite square lattice); the reference's default mu = U/2 is not."""
import os
import sys

import numpy as np
import pytest

from latticeqmc_amd import HubbardModel, measure, measure_betas, measure_scan

pytestmark = pytest.mark.gpu


def test_measure_functions_on_gpu(tmp_path):
    print("[jobs]", flush=True)
    sys.stdout.flush()
    m = HubbardModel(u=4.0, mu=0.0)
    m.build_square(2)
    tmp = str(tmp_path / "tmp_gf_series.npz")
    up, dn = measure_betas(m, [1.0, 2.0], 16, 20, 100, tmp_file=tmp, n_stab=8)
    assert up.shape == (2, 4, 4) and not os.path.exists(tmp)
    assert np.all(np.isfinite(up)) and np.all(np.isfinite(dn))
    assert abs(np.mean(1 - np.diagonal(up, axis1=1, axis2=2)) - 0.5) < 0.05
    gu, gd = measure(m, 2.0, 16, 20, 160, cores=8, n_stab=8)
    assert gu.shape == (4, 4) and abs(np.mean(1 - np.diag(gd)) - 0.5) < 0.05
    paths = measure_scan(m, [2.0, 4.0], [1.0, 2.0], 16, 20, 100, folder=str(tmp_path / "scan"), n_stab=8)
    with np.load(paths[2.0], allow_pickle=False) as d:
        assert d["gf_up"].shape == (2, 4, 4) and np.all(np.isfinite(d["gf_dn"]))
