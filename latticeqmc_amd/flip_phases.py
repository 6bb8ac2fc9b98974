"""Phase clocks of the delayed flip at the throughput shape (64 chains of 16 x 16, U = 4,
beta = 8, L = 80, k = 16, n_s = 8, four workgroups per chain).

    python -m latticeqmc_amd.build --probe     # on the build host: liblqmc_hip_probe.so
    python -m latticeqmc_amd.flip_phases       # on the GPU

The probe library is the product library with the flip's phase clocks compiled in (built with
-DLQMC_FLIP_PROBE; see flip.hip): thread 0 of every workgroup sums the wall-clock ticks (100 MHz)
spent in each phase of its 16-site blocks and writes the sums over the slice's trace ratios when
the launch ends.  One sweep runs launch by launch with tracing on (the per-site trace is not
written in this build, so the sites run as in the product), and the means over chains, slices
and workgroups are printed per launch and per block.
"""
import os
import sys
import time

import numpy as np

from . import _native
from .build import PROBE_LIB

PHASES = [(0, "chain barrier"), (1, "uniforms + sync"), (2, "diagonal-block loads (wave 0)"),
          (3, "sites (wave 0; deferred strips and row staging on the helper waves)"),
          (4, "wait for the other waves"), (12, "tables"), (13, "inverses"), (5, "scatter"),
          (6, "Z tiles"), (7, "strip update and copies"), (8, "tail")]


def main(argv=None):
    if not os.path.exists(PROBE_LIB):
        raise SystemExit(f"{PROBE_LIB} is missing: build it with `python -m latticeqmc_amd.build --probe`")
    _native.load(PROBE_LIB)
    from .engine import SweepEngine
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    groups = 4
    specs = bench.build_specs(bench.CONFIGS["c4"])
    eng = SweepEngine(specs, mode="dqmc", precision="fp64", delay=16, n_stab=8, rng="philox", trace=True,
                      graph=False, flip_groups=groups)
    try:
        eng.sweep(1, measure=False)
        eng.sync()
        t0 = time.time()
        eng.sweep(1, measure=True)
        eng.sync()
        dt = time.time() - t0
        flags = eng.stats()["flags"]
        rows = []
        for c in range(len(specs)):
            r, _ = eng.trace(c)
            L, N = specs[c].time_steps, specs[c].n_sites
            rows.append(r.reshape(L, N)[:, :16 * groups].reshape(L, groups, 16))
    finally:
        eng.close()
    a = np.array(rows)  # chain, slice, workgroup, slot
    nb = float(a[0, 0, 0, 14])
    print(f"delayed flip, c4, k = 16, {groups} workgroups per chain: sweep {1e3 * dt:.1f} ms (launch by launch, "
          f"tracing on), flags {flags}, {nb:.0f} blocks per launch")
    print("| phase | us per launch | us per block |")
    print("|---|---:|---:|")
    for p, name in PHASES:
        us = a[..., p].mean() / 100.0
        print(f"| {name} | {us:.1f} | {us / nb:.2f} |")
    print(f"| total per launch (workgroup clock) | {a[..., 15].mean() / 100.0:.1f} | |")


if __name__ == "__main__":
    main()
