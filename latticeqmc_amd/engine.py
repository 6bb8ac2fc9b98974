"""Python handle on the native sweep engine (one context = many chains on one GPU).

``ChainSpec`` carries what the reference's ``LatticeQMC.set_beta`` derives
(lqmc/lqmc.py:93-117): dtau = beta/L, lambda = arccosh(exp(U dtau/2)) (0 for
U = 0), E_K = expm(-dtau K); E_K^-1 = expm(+dtau K) is uploaded as well so the
wraps need no device inversion.
"""
import ctypes
from dataclasses import dataclass, field as dc_field

import numpy as np
from scipy.linalg import expm

from . import _native as nat

_MODES = {"dqmc": nat.MODE_DQMC, "reference": nat.MODE_REFERENCE, "det": nat.MODE_DET}
_PREC = {"fp64": nat.FP64, "fp32": nat.FP32, "bf16x3": nat.BF16X3}
_RNG = {"replay": nat.RNG_REPLAY, "philox": nat.RNG_PHILOX}


def slice_params(kmat, beta, time_steps, u):
    dtau = beta / time_steps
    lam = float(np.arccosh(np.exp(u * dtau / 2.0))) if u else 0.0
    return dtau, lam, expm(-dtau * kmat), expm(dtau * kmat)


@dataclass
class ChainSpec:
    kmat: np.ndarray            # (N, N) kinetic matrix K = -mu I - t A
    beta: float
    time_steps: int
    u: float
    field: np.ndarray           # (N, L) int8 +-1, reference layout
    seed: int = 0
    expk: np.ndarray = dc_field(default=None, repr=False)
    expk_inv: np.ndarray = dc_field(default=None, repr=False)
    lam: float = None

    def __post_init__(self):
        self.kmat = np.ascontiguousarray(self.kmat, dtype=np.float64)
        self.field = np.ascontiguousarray(self.field, dtype=np.int8)
        n = self.kmat.shape[0]
        if self.kmat.shape != (n, n) or self.field.shape != (n, self.time_steps):
            raise ValueError("kmat must be (N, N) and field (N, time_steps)")
        if self.expk is None or self.lam is None:
            _, lam, ek, eki = slice_params(self.kmat, self.beta, self.time_steps, self.u)
            self.expk = ek if self.expk is None else self.expk
            self.expk_inv = eki if self.expk_inv is None else self.expk_inv
            self.lam = lam if self.lam is None else self.lam

    @property
    def n_sites(self):
        return self.kmat.shape[0]


class SweepEngine:
    """Batched DQMC sweeps of independent chains on one GPU.

    mode      "dqmc" (textbook, QR/UDV-stabilised), "det" (the reference's brute-force
              determinant sweep, lqmc.py:236-299, N <= 32) or "reference" (literal
              lqmc/lqmc.py:301-347 math, FP64 rank-1, unstabilised)
    precision "fp64" or "fp32" (storage of G, wraps and flip updates; the
              stabilisation always runs in FP64), or "bf16x3": the FP32 form with the
              delayed rank-k products as three split-BF16 matrix-core products
    delay     1: rank-1 per accepted flip; k > 1: submatrix blocks of k sites
    n_stab    slices between FP64 recomputes
    rng       "replay" (uniforms supplied per sweep) or "philox" (on device)
    flip_groups  workgroups per chain in the delayed flip (delay 16 or 32; 0 = automatic,
              filling the compute units; results do not depend on it)
    time_displaced  Mode D: accumulate G(tau, 0) at the stabilisation points of measured
              sweeps (``green_tau``); adds two GEMMs and a triangular solve per point
    max_dg    Mode D: a recompute deviating from the propagated G by more than this raises
              a device flag, so ``sync`` fails (0: 1e-3; negative: off)
    dense_ek  run every product with E_K as a dense GEMM; by default a chain whose E_K is
              exactly Ex (x) Ey (square and cubic lattices, n1, n2 <= 16) applies the factors
              in its FP64 products (``ek_split``)
    """

    def __init__(self, chains, mode="dqmc", precision="fp64", delay=1, n_stab=10, rng="replay",
                 trace=False, device=0, obs_capacity=4096, flip_threads=0, flip_groups=0, graph=False,
                 time_displaced=False, max_dg=0.0, dense_ek=False):
        self.lib = nat.load()
        self.chains = list(chains)
        if not self.chains:
            raise ValueError("at least one chain is required")
        self.mode, self.precision, self.rng = mode, precision, rng
        self.delay, self.n_stab, self.trace_on = int(delay), int(n_stab), bool(trace)
        opt = nat.Options(_MODES[mode], _PREC[precision], int(delay), int(n_stab), _RNG[rng],
                          int(bool(trace)), int(obs_capacity), int(flip_threads), int(flip_groups),
                          int(bool(time_displaced)), float(max_dg), int(bool(dense_ek)))
        ns = (ctypes.c_int32 * len(self.chains))(*[c.n_sites for c in self.chains])
        ls = (ctypes.c_int32 * len(self.chains))(*[c.time_steps for c in self.chains])
        h = ctypes.c_void_p()
        nat.check(self.lib.lqmc_create(ctypes.byref(h), int(device), len(self.chains), ns, ls,
                                       ctypes.byref(opt)), "lqmc_create")
        self.h = h
        for i, c in enumerate(self.chains):
            self._set_chain(i, c)
        nat.check(self.lib.lqmc_init(self.h), "lqmc_init")
        if graph:
            nat.check(self.lib.lqmc_use_graph(self.h, 1), "lqmc_use_graph")

    @staticmethod
    def _dp(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def _set_chain(self, i, c):
        ek = np.ascontiguousarray(c.expk, dtype=np.float64)
        eki = np.ascontiguousarray(c.expk_inv, dtype=np.float64)
        km = np.ascontiguousarray(c.kmat, dtype=np.float64)
        f = np.ascontiguousarray(c.field, dtype=np.int8)
        nat.check(self.lib.lqmc_set_chain(self.h, i, self._dp(ek), self._dp(eki), self._dp(km),
                                          float(c.lam), f.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                                          int(c.seed) & 0xFFFFFFFFFFFFFFFF), "lqmc_set_chain")

    # -- control ---------------------------------------------------------------------------------
    def set_uniforms(self, i, u):
        c = self.chains[i]
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, c.n_sites * c.time_steps)
        nat.check(self.lib.lqmc_set_uniforms(self.h, i, self._dp(u), u.shape[0]), "lqmc_set_uniforms")

    def sweep(self, n=1, measure=False, stream=None):
        nat.check(self.lib.lqmc_sweep(self.h, int(n), int(bool(measure)), stream), "lqmc_sweep")

    def sync(self):
        nat.check(self.lib.lqmc_sync(self.h), "lqmc_sync")

    def use_graph(self, on=True):
        nat.check(self.lib.lqmc_use_graph(self.h, int(bool(on))), "lqmc_use_graph")

    # -- results ---------------------------------------------------------------------------------
    def green(self, i, which=0):
        n = self.chains[i].n_sites
        out = np.empty((2, n, n))
        nat.check(self.lib.lqmc_get_green(self.h, i, self._dp(out), int(which)), "lqmc_get_green")
        return out

    def spin_correlation(self, i):
        """Mean over the measured sweeps of <S^z_i S^z_j> (S^z = n_up - n_dn), (N, N)."""
        n = self.chains[i].n_sites
        out = np.empty((n, n), dtype=np.float64)
        nat.check(self.lib.lqmc_get_spin_corr(self.h, i, self._dp(out)), "lqmc_get_spin_corr")
        return out

    def green_tau(self, i):
        """Unequal-time G(tau, 0) = (B(tau,0)^-1 + B(beta,tau))^-1 of chain i, mean over the
        measured sweeps, at the stabilisation points tau = l dtau, l = L-n_s, L-2n_s, ..., 0.
        Returns (l, G) with l int32 (ntau,) and G (ntau, 2, N, N) [tau][spin][i][j]; at U = 0
        G equals -tools.compute_pole_gf_tau at those tau (tools.py:182-200)."""
        c = self.chains[i]
        nt = ctypes.c_int32()
        nat.check(self.lib.lqmc_get_green_tau(self.h, i, None, None, ctypes.byref(nt)), "lqmc_get_green_tau")
        ls = np.empty(nt.value, dtype=np.int32)
        out = np.empty((nt.value, 2, c.n_sites, c.n_sites))
        nat.check(self.lib.lqmc_get_green_tau(self.h, i, self._dp(out),
                                              ls.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), None),
                  "lqmc_get_green_tau")
        return ls, out

    def recompute_errors(self, i):
        """Mode D: max |G_propagated - G_recomputed| at each recompute point of the sweep
        (ceil(L/n_s) values, in sweep order) since creation or ``reset_stats``."""
        n = ctypes.c_int32()
        nat.check(self.lib.lqmc_get_recompute_err(self.h, i, None, ctypes.byref(n)), "lqmc_get_recompute_err")
        out = np.empty(n.value)
        nat.check(self.lib.lqmc_get_recompute_err(self.h, i, self._dp(out), None), "lqmc_get_recompute_err")
        return out

    def green_dev(self, i, out_ptr, which=0, stream=None):
        """Copy G of chain i (float64, (2, N, N)) into device memory at out_ptr (e.g. a torch
        tensor's data_ptr()) on ``stream`` (None: the engine stream); asynchronous."""
        nat.check(self.lib.lqmc_get_green_dev(self.h, i, ctypes.c_void_p(out_ptr), int(which), stream),
                  "lqmc_get_green_dev")

    def field(self, i):
        c = self.chains[i]
        out = np.empty((c.n_sites, c.time_steps), dtype=np.int8)
        nat.check(self.lib.lqmc_get_field(self.h, i, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))),
                  "lqmc_get_field")
        return out

    def stats(self, i=-1):
        s = nat.Stats()
        nat.check(self.lib.lqmc_get_stats(self.h, int(i), ctypes.byref(s)), "lqmc_get_stats")
        return {k: getattr(s, k) for k, _ in nat.Stats._fields_ if k != "pad"}

    def reset_stats(self):
        nat.check(self.lib.lqmc_reset_stats(self.h), "lqmc_reset_stats")

    def trace(self, i):
        c = self.chains[i]
        nl = c.n_sites * c.time_steps
        r = np.empty(nl)
        a = np.empty(nl, dtype=np.uint8)
        nat.check(self.lib.lqmc_get_trace(self.h, i, self._dp(r), a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))),
                  "lqmc_get_trace")
        return r, a.astype(bool)

    def observables(self, i, max_rows=1 << 20):
        st = self.stats(i)
        rows_max = int(min(max_rows, max(st["measured"], 1)))
        out = np.empty((rows_max, 3))
        rows = ctypes.c_int()
        nat.check(self.lib.lqmc_get_observables(self.h, i, self._dp(out), rows_max, ctypes.byref(rows)),
                  "lqmc_get_observables")
        return out[: rows.value]

    def ek_split(self, i):
        """(n1, n2) of chain i's Kronecker-factored E_K products, or (0, 0) where they are dense."""
        n1, n2 = ctypes.c_int32(), ctypes.c_int32()
        nat.check(self.lib.lqmc_get_ek_split(self.h, int(i), ctypes.byref(n1), ctypes.byref(n2)), "lqmc_get_ek_split")
        return n1.value, n2.value

    def profile(self, kind):
        nat.check(self.lib.lqmc_profile(self.h, {"off": -1, "flip": 0, "gemm": 1, "qr": 2, "solve": 3,
                                                 "kron": 4}[kind]),
                  "lqmc_profile")

    def profile_result(self):
        avg = ctypes.c_double()
        n = ctypes.c_int64()
        w = ctypes.c_double()
        nat.check(self.lib.lqmc_profile_result(self.h, ctypes.byref(avg), ctypes.byref(n), ctypes.byref(w)),
                  "lqmc_profile_result")
        return avg.value, n.value, w.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.lqmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def wrap_green(expk, expk_inv, v, g, workspace=None, stream=None, device=0):
    """Slice wrap G_m <- B_m^-1 G_m B_m, B_m = diag(v_m) E_K, of caller-held Green's functions
    on the device (``lqmc_wrap_green``; the consistent form of lqmc/lqmc.py:338-345).

    expk, expk_inv: (N, N) float64 device tensors; v: (count, N); g: (count, N, N), updated
    in place.  ``workspace`` (a uint8 device tensor of at least
    ``lqmc_wrap_green_workspace(N, count)`` bytes) is allocated with torch when omitted.
    Asynchronous on ``stream`` (a raw hipStream_t; None: the null stream)."""
    import torch
    lib = nat.load()
    count, n = int(g.shape[0]), int(g.shape[1])
    for t, shape in ((expk, (n, n)), (expk_inv, (n, n)), (v, (count, n)), (g, (count, n, n))):
        if tuple(t.shape) != shape or t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"expected a contiguous float64 device tensor of shape {shape}")
    need = lib.lqmc_wrap_green_workspace(n, count)
    if workspace is None:
        workspace = torch.empty(need, dtype=torch.uint8, device=g.device)
    nat.check(lib.lqmc_wrap_green(int(device), n, count, expk.data_ptr(), expk_inv.data_ptr(), v.data_ptr(),
                                  g.data_ptr(), workspace.data_ptr(), int(workspace.numel()), stream),
              "lqmc_wrap_green")
    return g
