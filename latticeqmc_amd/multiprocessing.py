"""The reference's chain-pool managers (lqmc/multiprocessing.py:55-341) on the batched GPU engine.

``ParallelProcessManager`` splits the measurement sweeps of one (model, beta) over ``procs``
independent chains, each with the full warmup, and averages the per-chain means unweighted
(multiprocessing.py:252-267); ``SerialProcessManager`` runs one chain per beta with a
resumable ``tmp_gf_series.npz`` cache (292-341).  Same constructors, ``set_jobs``, ``run`` and
``get_result``; the jobs run as chains of one batched launch instead of forked processes, so
there is no pipe to deadlock on large lattices and no per-process progress table.  Chain c
draws its initial field from ``RandomState(seed + c)`` and its Metropolis uniforms from the
on-device Philox stream of that seed.
"""
import os

import numpy as np

from .batched import BatchedDQMC, _model_spec, measure_betas


def _n_procs(procs):
    cpu = os.cpu_count() or 1
    if procs is None or procs == 0:
        return cpu
    return cpu + procs if procs < 0 else int(procs)


class ProcessManager:
    """Common part of the managers (multiprocessing.py:55-246): ``procs`` chains (None or 0:
    the CPU count, negative: that many fewer), the static keyword arguments of every job and
    the per-job variable ones set by ``set_jobs``."""

    def __init__(self, procs=None, seed=0, **default_kwargs):
        self.max_procs = _n_procs(procs)
        self.default_kwargs = default_kwargs
        self.var_kwargs = {}
        self.total = 0
        self.result = None
        self.seed = seed
        self.engine_kw = {}

    @property
    def model(self):
        return self.default_kwargs["model"]

    @property
    def all_done(self):
        return self.result is not None and all(r is not None for r in self.result)

    def get_result(self):
        return np.asarray(self.result)

    def __str__(self):
        return (f"{self.__class__.__name__}(Vars: {list(self.var_kwargs.keys())}, Jobs: {self.total}, "
                f"Processes: {self.max_procs})")


class ParallelProcessManager(ProcessManager):
    """One (model, beta) measured by ``procs`` chains (multiprocessing.py:252-289)."""

    def __init__(self, model, beta, time_steps, warmup=300, det_mode=False, procs=None, seed=0, **engine_kw):
        super().__init__(procs, seed=seed, model=model, beta=beta, time_steps=time_steps, warmup=warmup,
                         det_mode=det_mode)
        self.engine_kw = dict(engine_kw)
        if det_mode:
            self.engine_kw.update(mode="det", precision="fp64", delay=1)

    def set_jobs(self, sweeps):
        """sweeps // procs measurement sweeps per chain, the remainder on chain 0
        (multiprocessing.py:260-263)."""
        sweeplist = np.full(self.max_procs, sweeps // self.max_procs, dtype=np.int64)
        sweeplist[0] += sweeps - int(np.sum(sweeplist))
        self.var_kwargs = {"sweeps": sweeplist}
        self.total = self.max_procs
        self.result = [None] * self.total

    def run(self, sleep=None):
        kw = self.default_kwargs
        specs = [_model_spec(kw["model"], kw["beta"], kw["time_steps"], self.seed + c) for c in range(self.max_procs)]
        sweeplist = self.var_kwargs["sweeps"]
        base = int(sweeplist[1:].min()) if self.max_procs > 1 else int(sweeplist[0])
        bd = BatchedDQMC(specs, **self.engine_kw)
        try:
            gs = bd.run(kw["warmup"], base)
            self.result = [np.asarray(g) for g in gs]
            extra = int(sweeplist[0]) - base
            if extra > 0:  # chain 0's remainder; the other chains' means are already taken
                bd.engine.sweep(extra, measure=True)
                bd.engine.sync()
                self.result[0] = bd.engine.green(0, which=1)
        finally:
            bd.close()

    def get_result(self):
        """Unweighted mean of the per-chain means (multiprocessing.py:265-267)."""
        return np.sum(np.asarray(self.result), axis=0) / self.max_procs


class SerialProcessManager(ProcessManager):
    """One chain per beta with the full warmup and ``sweeps`` measurement sweeps
    (multiprocessing.py:292-341), resumable from ``tmp_gf_series.npz``."""

    def __init__(self, model, time_steps, warmup=300, sweeps=2000, det_mode=False, procs=None, caching=True,
                 seed=0, **engine_kw):
        super().__init__(procs, seed=seed, model=model, time_steps=time_steps, warmup=warmup, sweeps=sweeps,
                         det_mode=det_mode)
        self.caching = caching
        self._tmp_file = "tmp_gf_series.npz"
        self.engine_kw = dict(engine_kw)

    def set_jobs(self, betas):
        self.var_kwargs = {"beta": np.asarray(betas, dtype=np.float64)}
        self.total = len(betas)
        self.result = [None] * self.total

    def run(self, sleep=None):
        kw = self.default_kwargs
        up, dn = measure_betas(kw["model"], self.var_kwargs["beta"], kw["time_steps"], kw["warmup"], kw["sweeps"],
                               caching=self.caching, det_mode=kw["det_mode"], seed=self.seed,
                               tmp_file=self._tmp_file, **self.engine_kw)
        self.result = [np.stack([up[i], dn[i]]) for i in range(self.total)]

    def delete_cache(self):
        if os.path.isfile(self._tmp_file):
            os.remove(self._tmp_file)
