"""The autotuner on the GPU: a small search over (k, n_s) and the flip workgroup count returns
an entry that passes the 1e-6 recompute gate with its margin, is the fastest passing
candidate, and is cached under the workload's key so the benchmark can look it up; the
committed cache holds a passing entry for the throughput workload (64 chains 16x16 U=4
beta=8 L=80)."""
import json
import os
import sys

import pytest

from latticeqmc_amd import autotune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

pytestmark = pytest.mark.gpu


def test_tune_returns_passing_entry(tmp_path):
    print("[autotune] 8 chains 8x8 U=4 beta=2 L=20, k in (8, 16), n_s in (8, 4)", flush=True)
    cfg = dict(desc="8 chains 8x8", sides=[8] * 8, u=[4.0] * 8, beta=[2.0] * 8, L=[20] * 8, seeds=list(range(8)))
    specs = bench.build_specs(cfg)
    path = str(tmp_path / "cache.json")
    entry = autotune.tune(specs, delays=(8, 16), n_stabs=(8, 4), sweeps=1, path=path, log=print)
    assert entry["passed"] and entry["max_err"] <= entry["margin"] * entry["gate"]
    passing = [r for r in entry["candidates"] if r["max_err"] <= 0.5e-6 and r["flags"] == 0]
    assert entry["ms_per_sweep"] == min(r["ms_per_sweep"] for r in passing)
    # the second stage tried every workgroup count the split flip accepts for the winner
    ws = {r["flip_groups"] for r in entry["candidates"]
          if (r["delay"], r["n_stab"]) == (entry["delay"], entry["n_stab"])}
    assert set(autotune.group_candidates(specs, entry["delay"])) <= ws
    with open(path) as f:
        assert json.load(f)[autotune.workload_key(specs, "fp64")]["delay"] == entry["delay"]
    assert autotune.lookup(specs, "fp64", path=path)["n_stab"] == entry["n_stab"]


def test_committed_cache_has_the_throughput_workload():
    specs = bench.build_specs(bench.CONFIGS["c4"])
    entry = autotune.lookup(specs, "fp64")
    assert entry is not None, "no cached entry for the 64-chain 16x16 workload"
    assert entry["passed"] and entry["max_err"] <= 1e-6
