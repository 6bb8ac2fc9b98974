"""Many independent Markov chains of uneven size in one batched GPU launch.

Replaces the reference's process pool (lqmc/multiprocessing.py:252-341): every
chain of a ``BatchedDQMC`` advances in the same launches (one workgroup per
chain for the flip phase, batched GEMM/QR/solve for the rest), and the
results of ``measure``/``measure_betas`` keep the reference's return shapes
(lqmc/__init__.py:17-96).
"""
import numpy as np

from .engine import ChainSpec, SweepEngine


class BatchedDQMC:
    def __init__(self, chains, **engine_kw):
        engine_kw.setdefault("rng", "philox")
        self.specs = list(chains)
        self.engine = SweepEngine(self.specs, **engine_kw)

    def run(self, warmup, sweeps, chunk=None):
        chunk = chunk or max(1, max(warmup, sweeps))
        done = 0
        while done < warmup:
            k = min(chunk, warmup - done)
            self.engine.sweep(k, measure=False)
            done += k
        done = 0
        while done < sweeps:
            k = min(chunk, sweeps - done)
            self.engine.sweep(k, measure=True)
            done += k
        self.engine.sync()
        return [self.engine.green(i, which=1) for i in range(len(self.specs))]

    def observables(self, i):
        return self.engine.observables(i)

    def stats(self, i=-1):
        return self.engine.stats(i)

    def close(self):
        self.engine.close()


def _model_spec(model, beta, time_steps, seed):
    rs = np.random.RandomState(seed)
    field = (2 * rs.randint(0, 2, size=(model.n_sites, time_steps)) - 1).astype(np.int8)
    return ChainSpec(model.ham_kinetic(), beta, time_steps, model.u, field, seed=seed)


def measure(model, beta, time_steps, warmup, sweeps, cores=None, det_mode=False, seed=0, **engine_kw):
    """Mean (G_up, G_dn) over ``cores`` independent chains (lqmc/__init__.py:17-54).

    cores == 1 runs the single-chain drop-in; otherwise the measurement sweeps
    are split over ``cores`` chains (default 8) run in one batched launch, each
    with the full warmup, and the per-chain means are averaged unweighted, as
    the reference's ParallelProcessManager does (multiprocessing.py:260-267).
    """
    if det_mode:
        raise NotImplementedError("det_mode is not provided by the GPU engine")
    if cores is not None and cores == 1:
        from .lqmc import LatticeQMC
        solver = LatticeQMC(model, beta, time_steps, warmup, sweeps, **engine_kw)
        g = solver.run()
        return g[0], g[1]
    n = 8 if cores in (None, 0) else abs(int(cores))
    per = int(np.ceil(sweeps / n))
    specs = [_model_spec(model, beta, time_steps, seed + c) for c in range(n)]
    b = BatchedDQMC(specs, **engine_kw)
    gs = b.run(warmup, per)
    b.close()
    g = np.mean(gs, axis=0)
    return g[0], g[1]


def measure_betas(model, betas, time_steps, warmup=500, sweeps=5000, cores=-1, caching=True,
                  det_mode=False, seed=0, **engine_kw):
    """(G_up, G_dn) of shape (M, N, N), one chain per beta in one launch (lqmc/__init__.py:57-96)."""
    if det_mode:
        raise NotImplementedError("det_mode is not provided by the GPU engine")
    specs = [_model_spec(model, float(b), time_steps, seed + c) for c, b in enumerate(betas)]
    bd = BatchedDQMC(specs, **engine_kw)
    gs = np.asarray(bd.run(warmup, sweeps))
    bd.close()
    return gs[:, 0], gs[:, 1]
