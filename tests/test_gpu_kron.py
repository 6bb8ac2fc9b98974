"""E_K products through the lattice's Kronecker factors against the dense products.

A chain whose E_K = expm(-dtau K) is exactly Ex (x) Ey (the hopping matrix a Kronecker sum:
square, rectangular and cubic lattices, a chemical potential included) runs its FP64 E_K
products -- the wraps, the right chain, the left stack, and Mode R's products -- through the
factors; ``dense_ek=True`` forces the dense GEMMs.  The split found by the host check is
read back, and the two paths are run side by side on the same fields and Philox streams: the
same accept decisions (fields equal) and G equal to rounding.  The oracle comparisons of the
other GPU tests (square lattices, the 3x5 rectangle, the 3x3x3 cubic lattice) run through
the factored path as well, since it is the default.
"""
import numpy as np
import pytest

from gpu_cases import spec_from
from latticeqmc_amd.engine import SweepEngine
from latticeqmc_amd.hubbard import HubbardModel
from latticeqmc_amd.lattice import Lattice
from oracle import model

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _k(kind):
    if kind.startswith("square"):
        side = int(kind[6:])
        return model.square_k(side, side)
    m = HubbardModel(u=4.0, t=1.0, mu=0.3)
    if kind == "rect3x5":
        m.build(3, 5, cycling=(0, 1))
    elif kind == "cubic3":
        m.lattice = Lattice.cubic(shape=(3, 3, 3))
        m.lattice.set_periodic_boundary((0, 1, 2))
    elif kind == "hex3x2":
        m.lattice = Lattice.hexagonal(shape=(3, 2))
    elif kind == "chain12":
        m.build(12, 1, cycling=0)
    else:
        raise ValueError(kind)
    return m.ham_kinetic()


@pytest.mark.parametrize("kind,split", [
    ("square16", (16, 16)), ("square6", (6, 6)), ("rect3x5", (3, 5)), ("cubic3", (3, 9)),
    ("hex3x2", (0, 0)), ("chain12", (0, 0)),
])
def test_split_found(kind, split):
    _say(f"[kron split] {kind}")
    k = _k(kind)
    n = k.shape[0]
    L = 8
    spec = spec_from(k, 2.0, L, 4.0, model.random_field(n, L, 1))
    eng = SweepEngine([spec], mode="dqmc", delay=4, n_stab=4, rng="philox")
    dense = SweepEngine([spec], mode="dqmc", delay=4, n_stab=4, rng="philox", dense_ek=True)
    try:
        assert eng.ek_split(0) == split
        assert dense.ek_split(0) == (0, 0)
    finally:
        eng.close()
        dense.close()


@pytest.mark.parametrize("mode,kinds,delay", [
    ("dqmc", ("square16", "square6", "rect3x5", "cubic3", "hex3x2"), 16),
    ("dqmc", ("square6", "rect3x5"), 1),
    ("reference", ("square6", "rect3x5"), 1),
])
def test_factored_equals_dense(mode, kinds, delay):
    _say(f"[kron vs dense] {mode} delay {delay} {kinds}")
    specs = []
    for i, kind in enumerate(kinds):
        k = _k(kind)
        n = k.shape[0]
        L = 40 if mode == "dqmc" else 16
        beta = 4.0 if mode == "dqmc" else 1.0
        specs.append(spec_from(k, beta, L, 4.0, model.random_field(n, L, 10 + i), seed=100 + i))
    kw = dict(mode=mode, delay=delay, n_stab=8, rng="philox")
    a = SweepEngine(specs, **kw)
    b = SweepEngine(specs, dense_ek=True, **kw)
    try:
        splits = [a.ek_split(i) for i in range(len(specs))]
        _say(f"  splits {splits}")
        a.sweep(2, measure=True)
        b.sweep(2, measure=True)
        a.sync()
        b.sync()
        for i in range(len(specs)):
            fa, fb = a.field(i), b.field(i)
            assert np.array_equal(fa, fb), f"chain {i} ({kinds[i]}): fields differ"
            for w in (0, 1):
                ga, gb = a.green(i, which=w), b.green(i, which=w)
                d = float(np.abs(ga - gb).max())
                _say(f"  chain {i} {kinds[i]} spin {w}: max|G_kron - G_dense| {d:.2e}")
                assert d < 1e-9
            sa, sb = a.stats(i), b.stats(i)
            assert sa["accepted"] == sb["accepted"]
            if mode == "dqmc":
                assert sa["max_err"] < 1e-6
    finally:
        a.close()
        b.close()
