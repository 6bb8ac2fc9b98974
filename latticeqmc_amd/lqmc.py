"""Drop-in ``LatticeQMC`` running the sweep on the GPU.

Same constructor, attributes and methods as the reference's ``lqmc.lqmc.LatticeQMC``
(lqmc/lqmc.py:14-408): ``model, n_sites, time_steps, config, warm_sweeps,
meas_sweeps, exp_k, lamb, dtau, beta, it, status``; ``set_beta``,
``set_temperature``, ``get_exp_v``, ``get_m``, ``iter_sweeps``, ``warmup_loop``,
``measure_loop``, ``run_lqmc``, ``run``.  ``iter_sweeps`` is still called once per
sweep, so subclasses that override it (as lqmc/multiprocessing.py:39-43 does)
keep working.

Extra keyword-only options select the GPU form:
  mode="dqmc"       textbook stabilised DQMC (default); "reference" reproduces the
                    reference's own fast-sweep math (lqmc.py:301-347), which is
                    only consistent at U = 0
  precision, delay, n_stab, rng, seed, device, graph -- see ``SweepEngine``.
  delay=None (default) picks the delayed form where it pays: 16-site blocks for
  N >= 64, 8 for N >= 16, the rank-1 update below; graph=None replays whole
  sweeps from a captured hipGraph except in determinant mode.
  rng="numpy" (default) draws the uniforms from numpy's global RNG in the
  reference's order, one per attempted flip, so a seeded run consumes the same
  stream as the reference; rng="philox" draws them on the device.

``det_mode=True`` runs the reference's brute-force determinant sweep
(lqmc.py:236-299) on the device (lattices of at most 32 sites): every attempted
flip recomputes det(I + prod E_K V) of both spins, and the measurement is the
inverse of M(0) after each sweep.
"""
import logging
import time

import numpy as np
from scipy.linalg import expm

from .configuration import Configuration
from .engine import ChainSpec, SweepEngine

log = logging.getLogger("latticeqmc_amd")


class LatticeQMC:
    def __init__(self, model, beta, time_steps, warmup=300, sweeps=2000, det_mode=False, log_lvl=None,
                 *, mode="dqmc", precision="fp64", delay=None, n_stab=10, rng="numpy", seed=None,
                 device=0, graph=None, console=True, time_displaced=False):
        if log_lvl is not None:
            log.setLevel(log_lvl)
        self.model = model
        self.n_sites = model.n_sites
        self.time_steps = time_steps
        self.config = Configuration(self.n_sites, time_steps)
        self.warm_sweeps = warmup
        self.meas_sweeps = sweeps
        self.det_mode = det_mode
        self.status = ""
        self.it = 0
        self.ratio = 0.0
        self.acc = False
        self.ham_kin = self.model.ham_kinetic()
        self.beta = 0.0
        self.dtau = 0.0
        self.lamb = 0.0
        self.exp_k = None
        self.exp_k_inv = None
        if delay is None:
            delay = 1 if mode != "dqmc" else (16 if self.n_sites >= 64 else 8 if self.n_sites >= 16 else 1)
        self.mode, self.precision, self.delay, self.n_stab = mode, precision, delay, n_stab
        if det_mode:  # the reference's determinant sweep (lqmc.py:236-299) on the device
            self.mode, self.precision, self.delay = "det", "fp64", 1
        if graph is None:
            graph = self.mode != "det"
        self.rng = rng
        self.seed = seed if seed is not None else int(np.random.randint(0, 2**31 - 1)) if rng == "philox" else 0
        self.device = device
        self.graph = graph
        self.console = console
        self.time_displaced = bool(time_displaced)
        self._engine = None
        self.set_beta(beta)

    # -- parameters (lqmc.py:93-128) ------------------------------------------------------------
    def set_beta(self, beta):
        self.dtau = beta / self.time_steps
        self.beta = beta
        self.lamb = np.arccosh(np.exp(self.model.u * self.dtau / 2.0)) if self.model.u else 0
        self.exp_k = expm(-1 * self.dtau * self.ham_kin)
        self.exp_k_inv = expm(self.dtau * self.ham_kin)
        check_val = self.model.u * self.model.t * self.dtau ** 2
        if check_val >= 0.1:
            log.warning(f"Check-value {check_val:.2} should be smaller than 0.1!")
        self._engine = None

    def set_temperature(self, temp):
        self.set_beta(1 / temp)

    # -- host helpers with the reference's semantics (lqmc.py:132-185) --------------------------
    def get_exp_v(self, l, sigma):
        return np.diagflat(np.exp(-1 * sigma * self.lamb * self.config[:, l]))

    def get_m(self, l0, sigma):
        l0 = l0 % self.time_steps
        indices = list(reversed(range(self.time_steps)))
        order = indices[-l0:] + indices[:-l0]
        b_prod = np.eye(self.n_sites)
        for l in order:
            b_prod = b_prod @ (self.exp_k @ self.get_exp_v(l, sigma))
        return np.eye(self.n_sites) + b_prod

    # -- engine ----------------------------------------------------------------------------------
    def _get_engine(self):
        if self._engine is None:
            spec = ChainSpec(self.ham_kin, self.beta, self.time_steps, self.model.u, self.config.config,
                             seed=self.seed, expk=self.exp_k, expk_inv=self.exp_k_inv, lam=float(self.lamb))
            self._engine = SweepEngine([spec], mode=self.mode, precision=self.precision, delay=self.delay,
                                       n_stab=self.n_stab, rng="philox" if self.rng == "philox" else "replay",
                                       device=self.device, graph=self.graph,
                                       time_displaced=self.time_displaced and self.mode == "dqmc")
        return self._engine

    def _sweep(self, measure):
        eng = self._get_engine()
        if self.rng != "philox":
            eng.set_uniforms(0, np.random.rand(self.time_steps * self.n_sites))
        eng.sweep(1, measure=measure)

    def _pull_config(self):
        if self._engine is not None:
            self.config.config = self._engine.field(0)

    def iter_sweeps(self, n, console_updates=200):
        print_interval = max(1, int(n / console_updates)) if console_updates else 0
        for it in range(n):
            if self.console and print_interval and (it + 1) % print_interval == 0:
                print(f"\r{self.status} Sweep {it + 1} ({100 * (it + 1) / n:.1f}%)", end="", flush=True)
            self.it = it
            yield it
        if self.console and print_interval:
            print(f"\r{self.status} Sweep {n} (100.0%)")

    def _update_step(self):
        """One sweep; returns the end-of-sweep (G_up, G_dn) like lqmc.py:301-347."""
        self._sweep(measure=True)
        g = self._engine.green(0, which=2)
        return g[0], g[1]

    def warmup_loop(self):
        self.status = "Warmup"
        for _ in self.iter_sweeps(self.warm_sweeps):
            self._sweep(measure=False)
        self._engine.sync() if self._engine else None
        self._pull_config()

    def measure_loop(self):
        self.status = "Measurement"
        eng = self._get_engine()
        before = eng.stats(0)["measured"]
        g_before = eng.green(0, which=1) * before if before else 0.0
        for _ in self.iter_sweeps(self.meas_sweeps):
            self._sweep(measure=True)
        eng.sync()
        self._pull_config()
        n = eng.stats(0)["measured"]
        total = eng.green(0, which=1) * n - g_before
        return total / max(self.meas_sweeps, 1)

    # -- determinant sweep entry points (lqmc.py:261-299) ----------------------------------------
    def _use_det(self):
        """Switch this solver to the device determinant sweep, keeping the field."""
        if self.mode != "det":
            self._pull_config()
            if self._engine is not None:
                self._engine.close()
            self._engine = None
            self.mode, self.precision, self.delay = "det", "fp64", 1

    def warmup_loop_det(self):
        self._use_det()
        self.warmup_loop()

    def measure_loop_det(self):
        self._use_det()
        return self.measure_loop()

    def run_lqmc(self):
        self.warmup_loop()
        return self.measure_loop()

    def run(self):
        t0 = time.time()
        gf = self.run_lqmc()
        mins, secs = divmod(time.time() - t0, 60)
        log.info(f"Total time: {int(mins):0>2}:{int(secs):0>2} min")
        return gf

    def stats(self):
        return self._get_engine().stats(0)

    def observables(self):
        return self._get_engine().observables(0)

    def green_tau(self):
        """Unequal-time G(tau, 0) (``time_displaced=True``, Mode D): (tau, G) with tau = l dtau at
        the stabilisation points and G (ntau, 2, N, N), the mean over every measured sweep of
        this solver.  No counterpart in lqmc/lqmc.py; at U = 0 it is minus the free-fermion
        ``tools.compute_pole_gf_tau`` (tools.py:182-200) at those tau."""
        ls, g = self._get_engine().green_tau(0)
        return ls * self.dtau, g
