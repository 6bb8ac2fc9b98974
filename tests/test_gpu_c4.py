"""The throughput workload's own shape on the GPU: 16x16 periodic lattice (N = 256), U = 4,
beta = 8, dtau = 0.1 (L = 80), delay 16 and a recompute every 8 slices -- the form the
benchmark times, whose flip runs on four workgroups per chain (512 threads each).

- One chain against the oracle (oracle/dqmc.py) on replayed uniforms: accept bits equal
  away from ties |ratio - u| < 1e-12, the same field, and G within 1e-8 (lqmc.py:301-347
  made consistent, as the oracle states it).
- All 64 chains in one launch, replayed from the captured hipGraph, through the
  size-independent properties: no device flags, N*L attempts per chain and sweep, every
  recompute within 1e-6 of the propagated G (the metric's gate), and chain 0 of the batch
  bitwise equal to the same chain run alone.
"""
import os
import sys

import numpy as np
import pytest

from gpu_cases import masked_equal, spec_from
from latticeqmc_amd.engine import SweepEngine
from oracle import dqmc, model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("L,beta,groups,sweeps,delay", [(16, 2.0, 0, 2, 16), (80, 8.0, 4, 1, 16), (16, 2.0, 0, 1, 1)])
def test_c4_chain_matches_oracle(L, beta, groups, sweeps, delay):
    """delay 16: the benchmarked form; delay 1: the plain form it is compared with, whose
    rank-1 flip also runs on four workgroups per chain at this size."""
    _say(f"[c4 chain] 16x16 L={L} beta={beta} delay={delay} n_stab=8 flip_groups={groups}")
    side, u, n_stab = 16, 4.0, 8
    n = side * side
    k = model.square_k(side, side)
    f = model.random_field(n, L, 0)
    o = dqmc.DQMCOracle(k, beta, L, u, f, n_stab=n_stab)
    eng = SweepEngine([spec_from(k, beta, L, u, f)], mode="dqmc", precision="fp64", delay=delay,
                      n_stab=n_stab, rng="replay", trace=True, flip_groups=groups)
    g0 = eng.green(0, which=0)
    assert max(np.abs(g0[s] - o.g[s]).max() for s in range(2)) < 1e-8
    rs = np.random.RandomState(100)
    for s in range(sweeps):
        uni = rs.rand(L * n)
        res = o.sweep(uni)
        eng.set_uniforms(0, uni)
        eng.sweep(1, measure=True)
        eng.sync()
        ratio, acc = eng.trace(0)
        ok, ties = masked_equal(acc, res["accept"], res["ratio"], uni)
        assert ok, f"sweep {s}: accept bits differ ({ties} ties masked)"
        assert np.abs(ratio - res["ratio"]).max() <= 1e-8 * max(1.0, np.abs(res["ratio"]).max())
        assert np.array_equal(eng.field(0), o.field)
        gl = eng.green(0, which=2)
        gc = eng.green(0, which=0)
        err = max(max(np.abs(gl[sp] - res["g_meas"][sp]).max() for sp in range(2)),
                  max(np.abs(gc[sp] - o.g[sp]).max() for sp in range(2)))
        _say(f"  sweep {s}: acceptance {acc.mean():.3f}, max|G - G_oracle| {err:.2e}")
        assert err < 1e-8
    st = eng.stats(0)
    assert st["flags"] == 0 and st["max_err"] < 1e-8
    eng.close()


def test_c4_batch_properties():
    _say("[c4 batch] 64 chains 16x16 U=4 beta=8 L=80, delay 16, n_stab 8, hipGraph")
    specs = bench.build_specs(bench.CONFIGS["c4"])
    eng = SweepEngine(specs, mode="dqmc", precision="fp64", delay=16, n_stab=8, rng="philox", graph=True)
    sweeps = 2
    eng.sweep(1, measure=False)
    eng.sweep(sweeps - 1, measure=True)
    eng.sync()
    st = eng.stats()
    assert st["flags"] == 0
    assert st["max_err"] <= 1e-6, st["max_err"]
    for i, s in enumerate(specs):
        assert eng.stats(i)["attempted"] == s.n_sites * s.time_steps * sweeps
    rec = np.array([eng.recompute_errors(i) for i in range(len(specs))])
    assert rec.shape == (64, 10) and rec.max() <= 1e-6 and rec.max() == st["max_err"]
    _say(f"  max|dG| {st['max_err']:.2e}, acceptance {st['accepted'] / st['attempted']:.3f}")
    alone = SweepEngine(specs[:1], mode="dqmc", precision="fp64", delay=16, n_stab=8, rng="philox", graph=True)
    alone.sweep(1, measure=False)
    alone.sweep(sweeps - 1, measure=True)
    alone.sync()
    assert np.array_equal(alone.green(0, which=0), eng.green(0, which=0))
    assert np.array_equal(alone.green(0, which=1), eng.green(0, which=1))
    assert np.array_equal(alone.field(0), eng.field(0))
    alone.close()
    eng.close()
