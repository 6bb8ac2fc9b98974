"""The other workload shapes at full size, through the size-independent property the
stabilised sweep offers: every FP64 recompute agrees with the propagated G to 1e-6 (the
gate of the throughput metric), no device flags, and every chain attempts N*L flips per
sweep.  24x24 (N = 576, U = 6, beta = 10, L = 100) is the largest lattice the QR panel
supports; the batch of 256 uneven chains mixes 6x6, 8x8 and 10x10 lattices at three
temperatures.  Both use a recompute interval of 6 slices: at 8 the 24x24 U=6 chain drifts
to ~2e-5 and the U<=8 batch to ~4e-6 between recomputes.""" 
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from latticeqmc_amd.engine import SweepEngine  # noqa: E402

pytestmark = pytest.mark.gpu


def _check(specs, sweeps, **kw):
    eng = SweepEngine(specs, mode="dqmc", precision="fp64", rng="philox", graph=False, **kw)
    eng.sweep(sweeps, measure=True)
    eng.sync()
    st = eng.stats()
    per = [eng.stats(i)["attempted"] for i in range(len(specs))]
    eng.close()
    assert st["flags"] == 0
    assert st["max_err"] <= 1e-6, st["max_err"]
    assert per == [s.n_sites * s.time_steps * sweeps for s in specs]


def test_24x24_chain_stable():
    print("[24x24]", flush=True)
    cfg = dict(bench.CONFIGS["c5"])
    for key in ("sides", "u", "beta", "L", "seeds"):
        cfg[key] = cfg[key][:1]
    _check(bench.build_specs(cfg), 2, delay=16, n_stab=6)


def test_256_uneven_chains_stable():
    print("[256 chains]", flush=True)
    _check(bench.build_specs(bench.c3_config()), 3, delay=8, n_stab=6)
