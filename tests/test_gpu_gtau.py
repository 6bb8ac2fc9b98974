"""Unequal-time G(tau, 0) on the GPU (time_displaced option) against the oracle.

The oracle's sweep records G(tau, 0) = (S^-1 + R)^-1 at every stabilisation point
(oracle/dqmc.py stable_gtd, pinned on the CPU against explicit slice products and the
U = 0 free-fermion tools).  Bars: per-point means over the measured sweeps within 1e-8
of the oracle's (relative to max(1, |G|)), with the flips replayed bit for bit; sweeps that
do not measure add nothing; at U = 0 the means equal -tools.compute_pole_gf_tau; the
option is refused outside Mode D and the getter without it.
"""
import sys

import numpy as np
import pytest

from gpu_cases import masked_equal, spec_from, square_case
from latticeqmc_amd import tools
from latticeqmc_amd.engine import SweepEngine
from oracle import dqmc, model

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("side,L,beta,n_stab,delay", [
    (4, 16, 2.0, 5, 1), (6, 20, 2.0, 10, 16), (8, 40, 4.0, 10, 8),
])
def test_green_tau_matches_oracle(side, L, beta, n_stab, delay):
    _say(f"[gtau] side={side} L={L} beta={beta} n_stab={n_stab} delay={delay}")
    k, f = square_case(side, L, beta, 4.0, 5)
    n = side * side
    o = dqmc.DQMCOracle(k, beta, L, 4.0, f, n_stab=n_stab)
    eng = SweepEngine([spec_from(k, beta, L, 4.0, f)], mode="dqmc", delay=delay, n_stab=n_stab,
                      rng="replay", trace=True, time_displaced=True)
    rs = np.random.RandomState(7)
    acc_sum = None
    for s, measure in enumerate((False, True, True)):
        uni = rs.rand(L * n)
        res = o.sweep(uni)
        eng.set_uniforms(0, uni)
        eng.sweep(1, measure=measure)
        eng.sync()
        ratio, acc = eng.trace(0)
        ok, _ = masked_equal(acc, res["accept"], res["ratio"], uni)
        assert ok, f"sweep {s}: accept bits differ"
        ls, gm = eng.green_tau(0)
        assert list(ls) == [l for l, _ in res["gtd"]]
        if not measure:
            assert not gm.any()  # nothing accumulated outside measured sweeps
            continue
        g = np.array([[gu, gd] for _, (gu, gd) in res["gtd"]])
        acc_sum = g if acc_sum is None else acc_sum + g
    ls, gm = eng.green_tau(0)
    want = acc_sum / 2
    assert np.abs(gm - want).max() <= 1e-8 * max(1.0, np.abs(want).max())
    assert eng.stats(0)["flags"] == 0
    eng.close()


def test_green_tau_u0_equals_free_fermion_tools():
    _say("[gtau u0]")
    k = model.square_k(4, 4) - 0.3 * np.eye(16)
    L, beta, ns = 16, 2.0, 6
    from latticeqmc_amd.batched import BatchedDQMC
    bd = BatchedDQMC([spec_from(k, beta, L, 0.0, model.random_field(16, L, 0))], mode="dqmc", delay=4,
                     n_stab=ns, graph=True, time_displaced=True)
    bd.run(1, 2)
    tau, gm = bd.green_tau(0)
    assert np.allclose(tau, np.array([10, 4, 0]) * beta / L)
    tgrid, gt = tools.compute_pole_gf_tau(k, beta, n_tau=L + 1)  # tau_j = j dtau
    for j, l in enumerate([10, 4, 0]):
        assert np.isclose(tgrid[l], tau[j])
        for s in range(2):
            assert np.abs(gm[j, s] + gt[:, :, l]).max() < 1e-10
    bd.close()


def test_green_tau_option_checks():
    _say("[gtau checks]")
    from latticeqmc_amd._native import LqmcError
    k, f = square_case(2, 8, 1.0, 4.0, 0)
    with pytest.raises(LqmcError, match="Mode D"):
        SweepEngine([spec_from(k, 1.0, 8, 4.0, f)], mode="reference", rng="replay", time_displaced=True)
    eng = SweepEngine([spec_from(k, 1.0, 8, 4.0, f)], mode="dqmc", n_stab=4, rng="philox")
    with pytest.raises(LqmcError, match="time_displaced"):
        eng.green_tau(0)
    eng.close()


def test_green_tau_uneven_chains():
    """Two chains of different sizes and slice counts in one context: each keeps its own
    stabilisation points (ceil(L/n_s) of them) and matches its own oracle."""
    _say("[gtau uneven]")
    n_stab = 5
    cases = [(4, 16, 2.0, 11), (3, 12, 1.5, 12)]
    specs, oracles = [], []
    for side, L, beta, seed in cases:
        k, f = square_case(side, L, beta, 4.0, seed)
        specs.append(spec_from(k, beta, L, 4.0, f))
        oracles.append(dqmc.DQMCOracle(k, beta, L, 4.0, f, n_stab=n_stab))
    eng = SweepEngine(specs, mode="dqmc", delay=4, n_stab=n_stab, rng="replay", time_displaced=True)
    rs = np.random.RandomState(9)
    sums = [None, None]
    for _ in range(2):
        for c, ((side, L, _, _), o) in enumerate(zip(cases, oracles)):
            uni = rs.rand(L * side * side)
            res = o.sweep(uni)
            eng.set_uniforms(c, uni)
            g = np.array([[gu, gd] for _, (gu, gd) in res["gtd"]])
            sums[c] = g if sums[c] is None else sums[c] + g
        eng.sweep(1, measure=True)
        eng.sync()
    for c, ((side, L, _, _), o) in enumerate(zip(cases, oracles)):
        ls, gm = eng.green_tau(c)
        assert list(ls) == [L - 1 - min((j + 1) * n_stab - 1, L - 1) for j in range(-(-L // n_stab))]
        assert np.array_equal(eng.field(c), o.field)
        want = sums[c] / 2
        assert np.abs(gm - want).max() <= 1e-8 * max(1.0, np.abs(want).max())
    eng.close()


def test_drop_in_green_tau_u0():
    """LatticeQMC(time_displaced=True): after run(), green_tau() at U = 0 is minus the
    free-fermion pole Green's function of the model's kinetic matrix at tau = l dtau."""
    _say("[gtau drop-in]")
    from latticeqmc_amd import HubbardModel, LatticeQMC
    m = HubbardModel(u=0, t=1, mu=0.4)
    m.build_square(4)
    q = LatticeQMC(m, beta=2.0, time_steps=16, warmup=1, sweeps=2, n_stab=6, rng="philox", seed=3,
                   console=False, time_displaced=True)
    q.run()
    tau, g = q.green_tau()
    tgrid, gt = tools.compute_pole_gf_tau(q.ham_kin, 2.0, n_tau=17)
    assert np.allclose(tau, tgrid[[10, 4, 0]])
    for j, l in enumerate([10, 4, 0]):
        for s in range(2):
            assert np.abs(g[j, s] + gt[:, :, l]).max() < 1e-10
