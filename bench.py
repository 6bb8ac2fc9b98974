"""Throughput benchmark: spin-flip updates/sec of the DQMC sweep on one MI355X.

Default workload (the configuration the metric is quoted on): 64 independent
chains of a 16x16 periodic square lattice (N=256), t=1, U=4, mu=0, beta=8,
dtau=0.1 (L=80), i.i.d. +-1 fields from seeds 0..63, FP64 QR/UDV recompute every
n_s slices (autotuned, default 8).  One step = one sweep of every chain = 64*256*80 = 1,310,720
attempted single-spin flips.

Forms timed:
  chosen  batched chains, submatrix-delayed rank-k updates (k=16) with the
          rank-k products on the FP64 matrix cores, FP64 wraps, FP64 QR/UDV
          recompute every n_s slices, whole sweep replayed from a hipGraph
  plain   the same chains batched, FP64, rank-1 Sherman-Morrison per accepted
          flip, no graph
  fp32    the chosen form with FP32 G, wraps and updates, timed on one sweep
          and reported with its max|G - G_FP64|, which fails the 1e-6 gate at
          this beta (errors grow ~x2-3 per FP32 wrap)
value = chosen-form updates/sec over exactly --steps sweeps (inputs resident in
HBM, device synchronised before and after).  The JSON line also carries the
plain-form rate, their ratio, max|G - G_FP64| over the FP64 recomputes, the
roofline record of the profiled kernel and the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"hbm": 8000.0, "fp64": 78.6, "fp32": 157.3}  # GB/s, TFLOP/s (MI355X dense)

CONFIGS = {
    "c4": dict(desc="64 chains 16x16 U=4 beta=8 dtau=0.1 L=80", sides=[16] * 64, u=[4.0] * 64,
               beta=[8.0] * 64, L=[80] * 64, seeds=list(range(64))),
    "c2": dict(desc="1 chain 8x8 U=4 beta=4 dtau=0.1 L=40", sides=[8], u=[4.0], beta=[4.0], L=[40], seeds=[0]),
    "c5": dict(desc="32 chains 24x24 U=6 beta=10 dtau=0.1 L=100", sides=[24] * 32, u=[6.0] * 32,
               beta=[10.0] * 32, L=[100] * 32, seeds=list(range(32))),
}


def c3_config():
    rs = np.random.RandomState(0)
    sides = rs.choice([6, 8, 10], size=256).tolist()
    u = rs.uniform(2.0, 8.0, size=256).tolist()
    beta = rs.choice([2.0, 4.0, 6.0], size=256).tolist()
    L = [int(round(b / 0.1)) for b in beta]
    return dict(desc="256 chains side in {6,8,10}, U~U[2,8], beta in {2,4,6}, dtau=0.1", sides=sides, u=u,
                beta=beta, L=L, seeds=list(range(256)))


CONFIGS["c3"] = None


def build_specs(cfg):
    from latticeqmc_amd.engine import ChainSpec
    from latticeqmc_amd.hubbard import HubbardModel
    specs = []
    kcache = {}
    for side, u, beta, L, seed in zip(cfg["sides"], cfg["u"], cfg["beta"], cfg["L"], cfg["seeds"]):
        if side not in kcache:
            m = HubbardModel(u=u, t=1.0, mu=0.0)
            m.build_square(side)
            kcache[side] = m.ham_kinetic()
        k = kcache[side]
        rs = np.random.RandomState(seed)
        field = (2 * rs.randint(0, 2, size=(side * side, L)) - 1).astype(np.int8)
        specs.append(ChainSpec(k, beta, L, u, field, seed=seed))
    return specs


def sync_all():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def time_form(specs, steps, warmup, profile_kind=None, **kw):
    from latticeqmc_amd.engine import SweepEngine
    eng = SweepEngine(specs, rng="philox", **kw)
    upd = sum(s.n_sites * s.time_steps for s in specs)
    # the warmup ends with a measured sweep, so both sweep graphs (with and without the
    # measurement) are captured and instantiated before the timed region
    if warmup > 1:
        eng.sweep(warmup - 1, measure=False)
    eng.sweep(1, measure=True)
    eng.sync()
    eng.reset_stats()
    sync_all()
    t0 = time.perf_counter()
    eng.sweep(steps, measure=True)
    eng.sync()
    sync_all()
    dt = time.perf_counter() - t0
    st = eng.stats()
    rec = np.max([eng.recompute_errors(i) for i in range(len(specs))], axis=0).tolist() if len(
        {s.time_steps for s in specs}) == 1 else None
    prof = None
    if profile_kind:
        eng.use_graph(False)
        eng.profile(profile_kind)
        eng.sweep(max(1, min(steps, 2)), measure=True)
        eng.sync()
        prof = eng.profile_result()
        eng.profile("off")
    splits = [eng.ek_split(i) for i in range(len(specs))]
    eng.close()
    return dict(ms_per_step=1e3 * dt / steps, ups=upd * steps / dt, stats=st, prof=prof, upd_per_step=upd, rec=rec,
                splits=splits)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# Bounded samples per process (whole slices of one chain of the workload, at the reference's
# ~20 upd/s per core the faithful form needs ~13 s for one 16x16 slice).
CPU_SAMPLES = {"faithful": 2, "outer": 64, "dqmc": 32}


def cpu_baseline(cfg, procs=None):
    """CPU sweeps of the workload on the host cores, one chain per process like the
    reference's chain pool (lqmc/multiprocessing.py:252-289), each on a bounded sample of
    whole slices, with one BLAS thread per process:
      faithful  lqmc.py:301-347 as written -- the pure-Python j,k update loop and the
                per-attempt mean/var of _debug (lqmc.py:217-232); the headline baseline;
      outer     the same sweep with the rank-1 update as np.outer;
      dqmc      the oracle's stabilised textbook sweep (the physics the GPU form computes).
    The pool holds min(16, cpu_count) processes: the host share of one GPU on the
    benchmark machines is 16 cores."""
    import multiprocessing as mp
    ncpu = os.cpu_count() or 1
    procs = procs or min(16, ncpu)
    n = cfg["sides"][0] ** 2
    ctx = mp.get_context("fork")
    out = {}
    with ctx.Pool(procs, initializer=_pool_init) as pool:
        for kind, slices in CPU_SAMPLES.items():
            t0 = time.perf_counter()
            counts = pool.map(_cpu_worker, [(kind, cfg["sides"][c % len(cfg["sides"])], cfg["u"][0], cfg["beta"][0],
                                             cfg["L"][0], c, slices) for c in range(procs)])
            dt = time.perf_counter() - t0
            out[kind] = dict(value=sum(counts) / dt, attempts=sum(counts), seconds=dt, slices_per_process=slices)
    f = out["faithful"]
    return dict(value=f["value"], unit="spin-flip updates/s", cores=procs, kind="port",
                sample=f"faithful restatement of lqmc.py:301-347 (Python j,k loop, per-attempt mean/var), "
                       f"{CPU_SAMPLES['faithful']} slice x {n} sites per process incl. the sweep-start inversion, "
                       f"{procs} processes, {f['attempts']} attempts in {f['seconds']:.1f} s",
                cpu_model=cpu_model(), cpu_count=ncpu, variants=out)


def _pool_init():
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(1)
    except Exception:
        pass


def _cpu_worker(args):
    kind, side, u, beta, L, seed, slices = args
    from oracle import dqmc, mode_r, model
    k = model.square_k(side, side)
    _, lam, ek, _ = model.params(k, beta, L, u)
    f = model.random_field(side * side, L, seed)
    uni = np.random.RandomState(seed + 1000).rand(side * side * L)
    if kind == "faithful":
        return mode_r.sweep_faithful(f, ek, lam, uni, max_slices=slices)
    if kind == "outer":
        mode_r.sweep(f, ek, lam, uni, max_slices=slices)
        return slices * side * side
    o = dqmc.DQMCOracle(k, beta, L, u, f, n_stab=8)
    o.sweep(uni, max_slices=slices)
    return slices * side * side


def step_roofline(specs, n_stab, delay, esz, acc, ms_per_step, splits=None):
    """Whole-sweep roofline of the stabilised delayed form: T_roof = F64/P64 + F32/P32 +
    bytes/BW with the algorithmic work per chain and spin
      left stack   (L - n_s) slice products E_K X + (nb - 1) (QR 8/3 N^3 incl. Q, V update 2N^3)
      wraps        L x 2 products with E_K^-1 and E_K   -- in the form's precision
      right chain  L products E_K X + nb (QR 8/3 N^3 + V update 2N^3)
      recompute    nb x (X1 2N^3 + X2 2N^3 + QR(X) 8/3 N^3 + Z 2N^3 + solve N^3 + G 2N^3)
      flips        L x 2aN^3 (the rank-k products)        -- in the form's precision
    and the flips' G traffic, 4 N^2 s / k bytes per attempted update (both spins' G read and
    written once per block of k sites).  An E_K product is a dense 2N^3 GEMM, or, for a chain
    whose E_K factors as Ex (x) Ey (splits[i] = (n1, n2); FP64 products), 2 N^2 (n1 + n2) flops
    and 2 N^2 s bytes (operand read, product written), priced at the larger of flops/P64 and
    bytes/BW.  Diagonal scalings, ratios and copies are not counted."""
    f64 = f_form = byts = t_ek = ek_bytes = 0.0
    for i, sp in enumerate(specs):
        n, L = float(sp.n_sites), sp.time_steps
        n1, n2 = splits[i] if splits else (0, 0)
        n3 = n ** 3
        nb = -(-L // n_stab)
        stab = ((nb - 1) * (8 / 3 + 2) + nb * (8 / 3 + 2) + nb * (2 + 2 + 8 / 3 + 2 + 1 + 2)) * n3
        ek64 = (L - n_stab) + L          # E_K products of the stack and the right chain (FP64)
        ekf = 2 * L                      # the wraps' two products (form precision)
        form = L * 2 * acc * n3
        if n1:
            nk = ek64 + (ekf if esz == 8 else 0)
            if esz != 8:
                form += ekf * 2 * n3
            per_f, per_b = 2 * n * n * (n1 + n2), 2 * n * n * 8
            t_ek += 2 * nk * max(per_f / (PEAK["fp64"] * 1e12), per_b / (PEAK["hbm"] * 1e9))
            ek_bytes += 2 * nk * per_b
        else:
            stab += ek64 * 2 * n3
            form += ekf * 2 * n3
        f64 += 2 * stab
        f_form += 2 * form
        byts += 4 * n * n * esz / max(delay, 1) * n * L if delay > 1 else 4 * acc * n * n * esz * n * L
    p_form = PEAK["fp32"] if esz == 4 else PEAK["fp64"]
    if esz == 8:
        f64, f_form = f64 + f_form, 0.0
    t = (f64 / (PEAK["fp64"] * 1e12) + f_form / (p_form * 1e12) + byts / (PEAK["hbm"] * 1e9)) * 1e3 + t_ek * 1e3
    return {"t_roof_ms": t, "ms_per_step": ms_per_step, "frac": t / ms_per_step, "flops_fp64": f64,
            "flops_form": f_form, "flip_bytes": byts, "ek_factored_bytes": ek_bytes, "ek_factored_ms": t_ek * 1e3,
            "model": "T = F64/78.6 TF + F_form/P_form + flip bytes/8 TB/s + factored E_K products "
                     "(max of flops/78.6 TF and bytes/8 TB/s each), algorithmic counts per sweep"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c4")
    ap.add_argument("--delay", type=int, default=0, help="0: autotuned/cached choice")
    ap.add_argument("--n-stab", type=int, default=0, help="0: autotuned/cached choice")
    ap.add_argument("--flip-groups", type=int, default=-1,
                    help="workgroups per chain in the flip; -1: autotuned/cached choice, 0: automatic")
    ap.add_argument("--autotune", action="store_true", help="tune (k, n_s) under the 1e-6 gate first")
    ap.add_argument("--tune-cache", default=None, help="autotune JSON cache path")
    ap.add_argument("--precision", default="fp64", help="fp64 | fp32 | bf16x3 (FP32 G, split-BF16 rank-k products)")
    ap.add_argument("--max-dg", type=float, default=0.0,
                    help="recompute deviation that raises the device precision flag (0: 1e-3; negative: off)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-kind", default="flip",
                    help="kernel kind timed with HIP events: flip|gemm|qr|solve|kron")
    ap.add_argument("--traffic", default=None,
                    help="profiles/ summary (JSON) holding the measured HBM bytes per launch of the profiled "
                         "kernel (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device once (the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 8}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs on single-GPU boxes; the chains of the "
              "workload shard across GPUs without communication until the final reduction. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    cfg = c3_config() if args.config == "c3" else CONFIGS[args.config]
    specs = build_specs(cfg)
    log(f"[bench] workload: {cfg['desc']}, {len(specs)} chains")
    from latticeqmc_amd import autotune
    cache = args.tune_cache or autotune.DEFAULT_CACHE
    if args.autotune:
        entry = autotune.tune(specs, precision=args.precision, path=cache, log=log)
    else:
        entry = autotune.lookup(specs, args.precision, path=cache)
    tuned = entry is not None
    # untuned defaults per workload: the recompute interval keeps max|dG| <= 1e-6 with margin
    # (24x24 U=6 beta=10 and the U<=8 batch need n_s=6; n_s=8 drifts to 1e-5 / 4e-6 there)
    dflt = {"c4": (16, 8), "c2": (8, 8), "c3": (8, 6), "c5": (16, 6)}.get(args.config, (16, 6))
    if not args.delay:
        args.delay = entry["delay"] if tuned else dflt[0]
    if not args.n_stab:
        args.n_stab = entry["n_stab"] if tuned else dflt[1]
    if args.flip_groups < 0:
        args.flip_groups = entry.get("flip_groups", 0) if tuned else 0
    # a cache entry applies only to the form it was tuned for
    tuned = tuned and (args.delay, args.n_stab, args.flip_groups) == (
        entry["delay"], entry["n_stab"], entry.get("flip_groups", 0))
    log(f"[bench] chosen form: {args.precision} delay={args.delay} n_stab={args.n_stab} "
        f"flip_groups={args.flip_groups or 'auto'} graph={not args.no_graph} (autotuned: {tuned})")
    chosen = time_form(specs, args.steps, args.warmup, profile_kind=args.profile_kind, mode="dqmc",
                       precision=args.precision, delay=args.delay, n_stab=args.n_stab, graph=not args.no_graph,
                       max_dg=args.max_dg, flip_groups=args.flip_groups)
    log(f"[bench] chosen: {chosen['ms_per_step']:.1f} ms/sweep, max|dG| {chosen['stats']['max_err']:.2e}")
    plain = None
    if not args.no_plain:
        plain = time_form(specs, args.steps, args.warmup, profile_kind="flip", mode="dqmc", precision="fp64",
                          delay=1, n_stab=args.n_stab, graph=False)
        log(f"[bench] plain: {plain['ms_per_step']:.1f} ms/sweep, max|dG| {plain['stats']['max_err']:.2e}")
    fp32 = None
    if not args.no_fp32 and args.precision != "fp32":
        fp32 = time_form(specs, 1, 1, profile_kind=None, mode="dqmc", precision="fp32", delay=args.delay,
                         n_stab=args.n_stab, graph=not args.no_graph, max_dg=-1.0)
        log(f"[bench] fp32 form: {fp32['ms_per_step']:.1f} ms/sweep, max|dG| {fp32['stats']['max_err']:.2e}")
    st = chosen["stats"]
    out = {
        "metric": "spin-flip updates/sec, chosen form vs FP64 plain on 1 GPU, max|dG|<=1e-6",
        "value": chosen["ups"],
        "unit": "spin-flip updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": chosen["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (i.i.d. +-1 HS fields from seeds, Philox uniforms)",
        "config": {"workload": cfg["desc"], "chains": len(specs), "updates_per_step": chosen["upd_per_step"],
                   "form": f"batched, delayed rank-k (k={args.delay}), {args.precision} wraps/updates, "
                           f"FP64 QR/UDV recompute every {args.n_stab} slices, "
                           + ("hipGraph" if not args.no_graph else "no graph"),
                   "delay": args.delay, "n_stab": args.n_stab, "flip_groups": args.flip_groups,
                   "autotuned": tuned},
        "max_dG_vs_fp64": st["max_err"],
        "max_dG_per_recompute": chosen.get("rec"),
        "gate_1e-6": bool(st["max_err"] <= 1e-6),
        "acceptance": st["accepted"] / max(st["attempted"], 1),
    }
    esz = 8 if args.precision == "fp64" else 4
    acc = st["accepted"] / max(st["attempted"], 1)
    out["config"]["ek_split"] = chosen["splits"][0] if len(set(map(tuple, chosen["splits"]))) == 1 else "mixed"
    out["step_roofline"] = step_roofline(specs, args.n_stab, args.delay, esz, acc, chosen["ms_per_step"],
                                         chosen["splits"])
    if plain:
        out["plain_fp64_value"] = plain["ups"]
        out["plain_ms_per_step"] = plain["ms_per_step"]
        out["plain_max_dG"] = plain["stats"]["max_err"]
        out["speedup_vs_plain"] = chosen["ups"] / plain["ups"]
        if plain.get("prof"):
            pargs = argparse.Namespace(**vars(args))
            pargs.delay, pargs.precision = 1, "fp64"
            out["plain_roofline"] = roofline("flip", *plain["prof"], specs, pargs, plain)
        out["plain_step_roofline"] = step_roofline(specs, args.n_stab, 1, 8, plain["stats"]["accepted"] /
                                                   max(plain["stats"]["attempted"], 1), plain["ms_per_step"],
                                                   plain["splits"])
    if fp32:
        out["fp32_form"] = {"value": fp32["ups"], "max_dG_vs_fp64": fp32["stats"]["max_err"],
                            "gate_1e-6": bool(fp32["stats"]["max_err"] <= 1e-6)}
    prof = chosen.get("prof")
    if prof:
        avg_ms, launches, work = prof
        out["roofline"] = roofline(args.profile_kind, avg_ms, launches, work, specs, args, chosen)
        tpath = args.traffic or os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (args.profile_kind, args.config))
        if os.path.exists(tpath):
            with open(tpath) as f:
                tr = json.load(f)
            if tr.get("delay") in (None, out["config"]["delay"]):
                out["roofline"]["traffic"] = tr["bytes_per_launch"]
                out["roofline"]["traffic_source"] = os.path.relpath(tpath, ROOT)
    if not args.no_cpu:
        log("[bench] CPU baseline (bounded sample)")
        out["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(out), flush=True)


def roofline(kind, avg_ms, launches, work, specs, args, form):
    """Roofline record of the profiled kernel kind: algorithmic work per launch over the
    mean launch duration (HIP events on the engine stream).

    flip, delayed form: each block of k sites reads and writes both spins' G once
    (G += P Z), 2 * 2 * N^2 * s bytes per block = 4 N^2 s / k bytes per attempted flip
    (s = element size); with 4 a N^2 flops per attempt the arithmetic intensity is
    a k / s flop/byte (~1.3 at a = 0.65, k = 16, FP64), far below the FP64 ridge point
    (78.6 TFLOP/s / 8 TB/s ~ 10), so the bound is memory.  The flop rate is reported
    beside it.
    flip, rank-1 form: 4 a N^2 s bytes per attempt (G streamed per accepted flip).
    kron: the factored E_K products, memory-bound: each N x N product reads its operand and
    writes its result once, 2 N^2 * 8 bytes; the engine records 2 N^2 (n1 + n2) flops per
    product, so the bytes are flops * 8 / (n1 + n2).
    gemm/qr/solve: the flops recorded by the engine per launch, against the FP64 peak.
    """
    st = form["stats"]
    acc = st["accepted"] / max(st["attempted"], 1)
    n = specs[0].n_sites
    esz = 8 if args.precision == "fp64" else 4
    peak_mfma = PEAK["fp64"] if args.precision == "fp64" else PEAK["fp32"]
    if kind == "flip":
        flips_launch = work / max(launches, 1)
        if args.delay > 1:
            byts = 4.0 * n * n * esz / args.delay * flips_launch
            flops = 4.0 * acc * n * n * flips_launch
            achieved = byts / (avg_ms * 1e-3) / 1e9
            return {"bound": "hbm", "kernel": "k_flip_delayed (delayed rank-k, k=%d)" % args.delay,
                    "achieved": achieved, "peak": PEAK["hbm"], "unit": "GB/s", "frac": achieved / PEAK["hbm"],
                    "traffic": None, "avg_ms": avg_ms, "launches": launches, "bytes_per_launch": byts,
                    "per_unit": "4*N^2*s/k bytes per attempted flip (G read+write once per block of k sites)",
                    "flops_per_launch": flops, "tflops": flops / (avg_ms * 1e-3) / 1e12,
                    "tflops_frac_of_mfma_peak": flops / (avg_ms * 1e-3) / 1e12 / peak_mfma}
        byts = 4.0 * acc * n * n * esz * flips_launch
        achieved = byts / (avg_ms * 1e-3) / 1e9
        return {"bound": "hbm", "kernel": "k_flip_rank1", "achieved": achieved, "peak": PEAK["hbm"], "unit": "GB/s",
                "frac": achieved / PEAK["hbm"], "traffic": None, "avg_ms": avg_ms, "launches": launches,
                "bytes_per_launch": byts, "per_unit": "4*a*N^2*s bytes per attempted flip"}
    flops = work / max(launches, 1)
    if kind == "kron":
        n1, n2 = form["splits"][0]
        byts = flops * 8.0 / max(n1 + n2, 1)
        achieved = byts / (avg_ms * 1e-3) / 1e9
        return {"bound": "hbm", "kernel": "k_kron (E_K products through Ex (x) Ey)", "achieved": achieved,
                "peak": PEAK["hbm"], "unit": "GB/s", "frac": achieved / PEAK["hbm"], "traffic": None,
                "avg_ms": avg_ms, "launches": launches, "bytes_per_launch": byts,
                "per_unit": "2*N^2*8 bytes per product (operand read, result written once)",
                "flops_per_launch": flops, "tflops": flops / (avg_ms * 1e-3) / 1e12}
    achieved = flops / (avg_ms * 1e-3) / 1e12
    peak = peak_mfma if kind == "gemm" else PEAK["fp64"]
    return {"bound": "mfma", "kernel": kind, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": None, "avg_ms": avg_ms, "launches": launches,
            "flops_per_launch": flops}


if __name__ == "__main__":
    main()
