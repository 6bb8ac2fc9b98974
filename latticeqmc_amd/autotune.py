"""Autotuner for the delayed form: picks the delay k and the recompute interval n_s.

Each candidate (k, n_s) runs the real workload for a few sweeps; the fastest
candidate whose max|G_propagated - G_recomputed| stays within ``margin * gate``
wins.  Results are cached as JSON keyed by the workload (chain sizes, beta, U,
precision) and the device architecture, so a later run can reuse them without
tuning.
"""
import json
import os
import time

from .engine import SweepEngine

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_cache.json")


def workload_key(specs, precision, arch="gfx950"):
    sizes = sorted({(s.n_sites, s.time_steps, round(float(s.beta), 6), round(float(s.u), 6)) for s in specs})
    return f"{arch}|{precision}|chains={len(specs)}|" + ";".join(f"N{n}L{l}b{b}U{u}" for n, l, b, u in sizes)


def load_cache(path=DEFAULT_CACHE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def lookup(specs, precision, path=DEFAULT_CACHE):
    return load_cache(path).get(workload_key(specs, precision))


def measure(specs, precision, delay, n_stab, sweeps=3, warmup=1, graph=True):
    eng = SweepEngine(specs, mode="dqmc", precision=precision, delay=delay, n_stab=n_stab, rng="philox",
                      graph=graph)
    upd = sum(s.n_sites * s.time_steps for s in specs)
    eng.sweep(warmup)
    eng.sync()
    eng.reset_stats()
    t0 = time.perf_counter()
    eng.sweep(sweeps, measure=True)
    eng.sync()
    dt = time.perf_counter() - t0
    st = eng.stats()
    eng.close()
    return {"delay": delay, "n_stab": n_stab, "ups": upd * sweeps / dt, "ms_per_sweep": 1e3 * dt / sweeps,
            "max_err": st["max_err"]}


def tune(specs, precision="fp64", delays=(8, 16, 32), n_stabs=(10, 8, 6), gate=1e-6, margin=0.5, sweeps=3,
         path=DEFAULT_CACHE, log=print):
    rows = []
    for ns in n_stabs:
        for k in delays:
            r = measure(specs, precision, k, ns, sweeps=sweeps)
            rows.append(r)
            if log:
                log(f"[autotune] k={k} n_s={ns}: {r['ms_per_sweep']:.1f} ms/sweep, max|dG| {r['max_err']:.2e}")
    ok = [r for r in rows if r["max_err"] <= margin * gate]
    best = max(ok, key=lambda r: r["ups"]) if ok else min(rows, key=lambda r: r["max_err"])
    entry = {"delay": best["delay"], "n_stab": best["n_stab"], "gate": gate, "margin": margin,
             "candidates": rows}
    if path:
        cache = load_cache(path)
        cache[workload_key(specs, precision)] = entry
        with open(path, "w") as f:
            json.dump(cache, f, indent=1)
    return entry
