"""Reduced-precision forms on the GPU against FP64, in the configured shapes of the workloads.

- 8x8 U=4 beta=4 L=40, one chain, FP32 G/wraps/updates (delay 8, a recompute every 8
  slices): the deviation max|G_FP32 - G_FP64| at each of the five recomputes of a sweep
  against the FP32 restatement of the oracle (oracle/dqmc.py, dtype=float32) on the same
  replayed uniforms.  Both propagate in FP32 and recompute in FP64, so their deviations sit
  at the same scale (well above the FP64 floor, well below O(1)); the summation orders of
  the matrix cores and of numpy's FP32 products differ, so the values agree in magnitude,
  not bit for bit.  The first slice's accept bits agree away from near-ties.
- 256 uneven chains (sides 6/8/10, U in [2, 8], beta in {2, 4, 6}) in the configured form:
  FP32 updates, a recompute every 10 slices, hipGraph replay; with the FP64 form of the
  same interval beside it.  Both are run and their deviations recorded; the precision flag
  (LQMC_FLAG_PRECISION, threshold 1e-3) is raised exactly when a recompute deviates by more.
- 32 chains 24x24 U=6 beta=10 L=100, a recompute every 8 slices: FP32 rank-k products
  against split-BF16 (three BF16 products, FP32 accumulation) with FP32 G, each against the
  FP64 recompute, and the FP64 form of the same interval.
"""
import os
import sys

import numpy as np
import pytest

from gpu_cases import masked_equal, spec_from
from latticeqmc_amd._native import LqmcError
from latticeqmc_amd.engine import SweepEngine
from oracle import dqmc, model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_c2_fp32_recompute_deviations_match_fp32_oracle():
    _say("[c2 fp32] 8x8 U=4 beta=4 L=40, delay 8, n_stab 8, FP32 vs the FP32 oracle")
    side, beta, L, u, n_stab = 8, 4.0, 40, 4.0, 8
    n = side * side
    k = model.square_k(side, side)
    f = model.random_field(n, L, 0)
    o = dqmc.DQMCOracle(k, beta, L, u, f, n_stab=n_stab, delay=8, dtype=np.float32)
    eng = SweepEngine([spec_from(k, beta, L, u, f)], mode="dqmc", precision="fp32", delay=8, n_stab=n_stab,
                      rng="replay", trace=True, max_dg=-1.0)
    uni = np.random.RandomState(5).rand(L * n)
    res = o.sweep(uni)
    eng.set_uniforms(0, uni)
    eng.sweep(1, measure=True)
    eng.sync()
    gpu = eng.recompute_errors(0)
    ref = np.asarray(res["errors"])
    fmt = {"float_kind": lambda x: f"{x:.2e}"}
    _say(f"  per-recompute max|dG| GPU    {np.array2string(gpu, formatter=fmt)}")
    _say(f"  per-recompute max|dG| oracle {np.array2string(ref, formatter=fmt)}")
    assert gpu.shape == ref.shape == (L // n_stab,)
    # FP32 propagation: far above the FP64 floor (~1e-13 here), far below O(1)
    assert gpu.min() > 1e-8 and ref.min() > 1e-8
    assert gpu.max() < 1e-2 and ref.max() < 1e-2
    # the same scale at every recompute
    ratio = gpu / ref
    assert np.all(ratio > 1 / 30) and np.all(ratio < 30), ratio
    # the first slice: identical starting G up to FP32 rounding, so the decisions agree
    ratio_gpu, acc_gpu = eng.trace(0)
    ok, ties = masked_equal(acc_gpu[:n], res["accept"][:n], res["ratio"][:n], uni[:n], tol=1e-4)
    assert ok, f"first slice accept bits differ ({ties} near-ties masked)"
    st = eng.stats(0)
    assert st["flags"] == 0 and st["attempted"] == n * L
    eng.close()


def _run(specs, sweeps, **kw):
    """Run the sweeps; returns (stats, max per-recompute deviation, whether the sync raised).
    A device flag makes the sync raise; the stats are read after it either way."""
    eng = SweepEngine(specs, mode="dqmc", rng="philox", **kw)
    eng.sweep(1, measure=False)
    eng.sweep(sweeps - 1, measure=True)
    try:
        eng.sync()
        raised = False
    except LqmcError:
        raised = True
    st = eng.stats()
    per = [eng.stats(i)["attempted"] for i in range(len(specs))]
    rec = [eng.recompute_errors(i).max() for i in range(len(specs))]
    eng.close()
    assert per == [s.n_sites * s.time_steps * sweeps for s in specs]
    assert raised == (st["flags"] != 0)
    return st, float(np.max(rec))


def test_c3_configured_form_fp32_n10_graph():
    _say("[c3 configured] 256 chains, FP32 updates, n_stab 10, hipGraph; FP64 n_stab 10 beside it")
    specs = bench.build_specs(bench.c3_config())
    st32, e32 = _run(specs, 2, precision="fp32", delay=8, n_stab=10, graph=True)
    st64, e64 = _run(specs, 2, precision="fp64", delay=8, n_stab=10, graph=True)
    _say(f"  max|dG| vs FP64 recompute: fp32 {st32['max_err']:.2e} (flags {st32['flags']}), "
         f"fp64 {st64['max_err']:.2e} (flags {st64['flags']})")
    assert np.isfinite(st32["max_err"]) and np.isfinite(st64["max_err"])
    assert st32["max_err"] == e32 and st64["max_err"] == e64
    # only the precision flag may be raised, and exactly when a recompute deviates by > 1e-3
    for st in (st32, st64):
        assert (st["flags"] & ~32) == 0
        assert bool(st["flags"] & 32) == (st["max_err"] > 1e-3)
    assert st64["max_err"] < st32["max_err"]


def test_c5_configured_form_fp32_vs_bf16x3():
    _say("[c5 configured] 32 chains 24x24 U=6 beta=10 L=100, n_stab 8: fp32, bf16x3, fp64")
    specs = bench.build_specs(bench.CONFIGS["c5"])
    err = {}
    for prec in ("fp32", "bf16x3", "fp64"):
        st, e = _run(specs, 2, precision=prec, delay=16, n_stab=8, graph=True, max_dg=-1.0)
        assert st["flags"] == 0 and np.isfinite(st["max_err"]) and st["max_err"] == e
        err[prec] = st["max_err"]
    _say(f"  max|dG| vs FP64 recompute: fp32 {err['fp32']:.2e}, bf16x3 {err['bf16x3']:.2e}, "
         f"fp64 {err['fp64']:.2e}")
    # the reduced-precision products sit far above the FP64 form's drift
    assert err["fp32"] > 10 * err["fp64"] and err["bf16x3"] > 10 * err["fp64"]
