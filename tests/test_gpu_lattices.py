"""The sweep on the other lattices the geometry layer builds (the reference's prefabs,
lqmc/lattice.py:310-421, and rectangular builds, 914-992): open and periodic chains, a
periodic rectangle, the open two-atom hexagonal lattice and the periodic cubic lattice,
each against the oracle on replayed uniforms -- accept bits equal away from ties, the same
field, G within 1e-8.  The kernels take K as a dense matrix, so these check that nothing
in them assumes the square lattice's structure (site count, degree, bipartite sign of K)."""
import numpy as np
import pytest

from gpu_cases import masked_equal, spec_from
from latticeqmc_amd.hubbard import HubbardModel
from latticeqmc_amd.lattice import Lattice
from latticeqmc_amd.engine import SweepEngine
from oracle import dqmc, model

pytestmark = pytest.mark.gpu


def _kmat(kind, mu):
    m = HubbardModel(u=4.0, t=1.0, mu=mu)
    if kind == "chain10_open":
        m.lattice = Lattice.chain(size=10)
    elif kind == "chain12_pbc":
        m.build(12, 1, cycling=0)
    elif kind == "rect3x5_pbc":
        m.build(3, 5, cycling=(0, 1))
    elif kind == "hex3x2_open":
        m.lattice = Lattice.hexagonal(shape=(3, 2))
    elif kind == "cubic3_pbc":
        m.lattice = Lattice.cubic(shape=(3, 3, 3))
        m.lattice.set_periodic_boundary((0, 1, 2))
    else:
        raise ValueError(kind)
    return m.ham_kinetic()


@pytest.mark.parametrize("kind,mu,delay", [
    ("chain10_open", 0.0, 1), ("chain12_pbc", 0.0, 4), ("rect3x5_pbc", 0.3, 8),
    ("hex3x2_open", 0.0, 4), ("cubic3_pbc", 0.0, 16),
])
def test_lattice_sweep_matches_oracle(kind, mu, delay):
    print(f"[lattice] {kind} mu={mu} delay={delay}", flush=True)
    k = _kmat(kind, mu)
    n = k.shape[0]
    L, beta, u, n_stab = 16, 2.0, 4.0, 8
    f = model.random_field(n, L, 4)
    o = dqmc.DQMCOracle(k, beta, L, u, f, n_stab=n_stab)
    eng = SweepEngine([spec_from(k, beta, L, u, f)], mode="dqmc", precision="fp64", delay=delay, n_stab=n_stab,
                      rng="replay", trace=True)
    g0 = eng.green(0, which=0)
    assert max(np.abs(g0[s] - o.g[s]).max() for s in range(2)) < 1e-8
    rs = np.random.RandomState(11)
    for s in range(2):
        uni = rs.rand(L * n)
        res = o.sweep(uni)
        eng.set_uniforms(0, uni)
        eng.sweep(1, measure=True)
        eng.sync()
        ratio, acc = eng.trace(0)
        ok, ties = masked_equal(acc, res["accept"], res["ratio"], uni)
        assert ok, f"sweep {s}: accept bits differ ({ties} ties)"
        assert np.array_equal(eng.field(0), o.field)
        gl = eng.green(0, which=2)
        assert max(np.abs(gl[sp] - res["g_meas"][sp]).max() for sp in range(2)) < 1e-8
        assert max(np.abs(eng.green(0, which=0)[sp] - o.g[sp]).max() for sp in range(2)) < 1e-8
    assert eng.stats(0)["flags"] == 0
    eng.close()
