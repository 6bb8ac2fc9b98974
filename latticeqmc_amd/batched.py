"""Many independent Markov chains of uneven size in one batched GPU launch.

Replaces the reference's process pool (lqmc/multiprocessing.py:252-341): every
chain of a ``BatchedDQMC`` advances in the same launches (one workgroup per
chain for the flip phase, batched GEMM/QR/solve for the rest), and the
results of ``measure``/``measure_betas`` keep the reference's return shapes
(lqmc/__init__.py:17-96).
"""
import os

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

    def green_tau(self, i):
        """Unequal-time G(tau, 0) of chain i (engine option ``time_displaced=True``), mean over
        the measured sweeps at the stabilisation points: (tau, G) with tau = l * dtau
        decreasing to 0 and G of shape (ntau, 2, N, N)."""
        ls, g = self.engine.green_tau(i)
        c = self.specs[i]
        return ls * (c.beta / c.time_steps), g

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
        engine_kw.update(mode="det", precision="fp64", delay=1)
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


TMP_FILE = "tmp_gf_series.npz"


def gpu_runner(model, betas, time_steps, warmup, sweeps, seeds, **engine_kw):
    """One chain per beta of ``model`` in one batched launch; (M, 2, N, N) mean G."""
    specs = [_model_spec(model, float(b), time_steps, int(s)) for b, s in zip(betas, seeds)]
    bd = BatchedDQMC(specs, **engine_kw)
    try:
        return np.asarray(bd.run(warmup, sweeps))
    finally:
        bd.close()


def _load_cache(path, betas, shape):
    """Results cached by an earlier, interrupted run; NaN rows are still pending.  Files
    holding Python objects (the reference's own cache) are not unpickled and are ignored."""
    try:
        with np.load(path, allow_pickle=False) as d:
            if np.array_equal(d["beta"], betas) and d["data"].shape == shape:
                return np.array(d["data"], dtype=np.float64)
    except (ValueError, KeyError, OSError):
        pass
    print(f"Ignoring incompatible cache {path}")
    return None


def measure_betas(model, betas, time_steps, warmup=500, sweeps=5000, cores=-1, caching=True,
                  det_mode=False, seed=0, batch=None, tmp_file=TMP_FILE, runner=None, **engine_kw):
    """(G_up, G_dn) of shape (M, N, N) for M temperatures (lqmc/__init__.py:57-96).

    One chain per beta with the full warmup and ``sweeps`` measurement sweeps, as the
    reference's SerialProcessManager runs one job per beta (multiprocessing.py:292-341), but
    ``batch`` betas at a time in one batched launch (default: all).  With ``caching`` the
    results so far are kept in ``tmp_file`` (keys ``beta`` and ``data`` (M, 2, N, N), pending
    rows NaN) after every batch; a rerun with the same betas resumes from it, and the file is
    removed when all betas are done.  ``cores`` is accepted for signature compatibility.
    """
    if det_mode:
        engine_kw.update(mode="det", precision="fp64", delay=1)
    runner = runner or gpu_runner
    betas = np.asarray(betas, dtype=np.float64)
    n = model.n_sites
    shape = (len(betas), 2, n, n)
    data = None
    if caching and os.path.isfile(tmp_file):
        data = _load_cache(tmp_file, betas, shape)
        if data is not None:
            done = int(np.sum(~np.isnan(data).any(axis=(1, 2, 3))))
            print(f"Found temporary data. Continuing with {len(betas) - done} of {len(betas)} jobs...")
    if data is None:
        data = np.full(shape, np.nan)
    pending = [i for i in range(len(betas)) if np.isnan(data[i]).any()]
    step = batch or max(1, len(pending))
    for c0 in range(0, len(pending), step):
        idx = pending[c0:c0 + step]
        data[idx] = runner(model, betas[idx], time_steps, warmup, sweeps, [seed + i for i in idx], **engine_kw)
        if caching:
            np.savez(tmp_file, beta=betas, data=data)
    if caching and os.path.isfile(tmp_file):
        os.remove(tmp_file)
    return data[:, 0], data[:, 1]


def get_datadir(shape, time_steps, warmup, sweeps, root="data"):
    """Result folder of a U x beta scan, ``<root>/<w>x<h>_nt=<L>_warm=<W>_meas=<S>`` (the
    reference's comp.py:11-16 layout)."""
    w, h = shape[0], shape[1] if len(shape) > 1 else 1
    folder = os.path.abspath(os.path.join(root, f"{w}x{h}_nt={time_steps}_warm={warmup}_meas={sweeps}"))
    os.makedirs(folder, exist_ok=True)
    return folder


def get_datapath(filename, model, post="", mkdir=True, root="data", **kwargs):
    """Path ``<root>/u=.._t=.._mu=.._w=.._h=../<filename> <k=v_...><post>.npz`` (lqmc/tools.py:11-38)."""
    w, h = model.lattice.shape[0], model.lattice.shape[1]
    folder = os.path.join(root, f"u={model.u}_t={model.t}_mu={model.mu}_w={w}_h={h}")
    if mkdir and not os.path.isdir(folder):
        os.makedirs(folder)
    kwargstr = "_".join([f"{key}={val}" for key, val in kwargs.items()])
    return os.path.abspath(os.path.join(folder, f"{filename} {kwargstr}{post}.npz"))


def measure_scan(model, u_values, betas, time_steps, warmup=500, sweeps=5000, folder=None, seed=0,
                 runner=None, **engine_kw):
    """U x beta scan written as ``u=<U>.npz`` files (keys ``beta``, ``gf_up``, ``gf_dn``),
    the workflow of the reference's comp.py:19-47 -- but every pending (U, beta) pair runs in
    ONE batched launch instead of one process pool per U.  U values whose file already exists
    are skipped (resume).  As in comp.py, ``model.set_params(u=U)`` sets each U (t = 1,
    mu = U/2), and the model is left with the last one.  ``runner(specs, warmup, sweeps)``
    replaces the GPU run (used by the CPU tests).  Returns {U: path}."""
    betas = np.asarray(betas, dtype=np.float64)
    folder = folder or get_datadir(model.lattice.shape, time_steps, warmup, sweeps)
    os.makedirs(folder, exist_ok=True)
    paths, todo = {}, []
    for u in u_values:
        model.set_params(u=u)
        path = os.path.join(folder, f"u={model.u}.npz")
        paths[u] = path
        if not os.path.isfile(path):
            todo.append(u)
    if not todo:
        return paths
    specs = []
    u_index = {u: j for j, u in enumerate(u_values)}
    for u in todo:
        model.set_params(u=u)
        kmat = model.ham_kinetic()
        for i, b in enumerate(betas):
            # one seed per (U, beta) pair of the whole scan, however it is split or resumed
            s = seed + u_index[u] * len(betas) + i
            rs = np.random.RandomState(s)
            field = (2 * rs.randint(0, 2, size=(model.n_sites, time_steps)) - 1).astype(np.int8)
            specs.append(ChainSpec(kmat, float(b), time_steps, model.u, field, seed=s))
    if runner is not None:
        gs = np.asarray(runner(specs, warmup, sweeps))
    else:
        bd = BatchedDQMC(specs, **engine_kw)
        try:
            gs = np.asarray(bd.run(warmup, sweeps))
        finally:
            bd.close()
    for j, u in enumerate(todo):
        g = gs[j * len(betas):(j + 1) * len(betas)]
        np.savez(paths[u], beta=betas, gf_up=g[:, 0], gf_dn=g[:, 1])
    return paths
