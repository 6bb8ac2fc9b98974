"""GPU sweep vs the oracle, through the C ABI (ctypes).

Bars: accept bits identical except ties |ratio - u| < 1e-12; FP64 G within 1e-8
(absolute, Mode D) or 1e-8 * max(1, max|G|) (Mode R, whose G reaches 1e2..1e3).
"""
import sys

import numpy as np
import pytest

from gpu_cases import masked_equal, spec_from, square_case
from latticeqmc_amd.engine import SweepEngine
from oracle import dqmc, exact, model, philox

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["mode_r_2x2", "mode_r_4x4", "mode_r_4x4_u0"])
def test_mode_r_matches_reference_fixture(golden_dir, name):
    _say(f"[mode_r] {name}")
    g = np.load(f"{golden_dir}/{name}.npz")
    L = int(g["L"])
    spec = spec_from(g["K"], float(g["beta"]), L, float(g["u"]), g["field0"])
    eng = SweepEngine([spec], mode="reference", precision="fp64", rng="replay", trace=True)
    for s in range(g["uniforms"].shape[0]):
        eng.set_uniforms(0, g["uniforms"][s])
        eng.sweep(1, measure=True)
        eng.sync()
        ratio, acc = eng.trace(0)
        ok, ties = masked_equal(acc, g["acc"][s], g["ratio"][s], g["uniforms"][s])
        assert ok, f"sweep {s}: accept bits differ"
        assert np.abs(ratio - g["ratio"][s]).max() <= 1e-8 * max(1.0, np.abs(g["ratio"][s]).max())
        assert np.array_equal(eng.field(0), g["field"][s])
        gl = eng.green(0, which=2)
        scale = max(1.0, np.abs(g["g_up"][s]).max(), np.abs(g["g_dn"][s]).max())
        assert np.abs(gl[0] - g["g_up"][s]).max() <= 1e-8 * scale
        assert np.abs(gl[1] - g["g_dn"][s]).max() <= 1e-8 * scale
    assert np.abs(eng.green(0, which=1) - g["meas_g"]).max() <= 1e-8 * max(1.0, np.abs(g["meas_g"]).max())
    eng.close()


def _run_dqmc_vs_oracle(side, L, beta, u, n_stab, delay, sweeps, seed=0, precision="fp64", flip_groups=0):
    k, f = square_case(side, L, beta, u, seed)
    n = side * side
    o = dqmc.DQMCOracle(k, beta, L, u, f, n_stab=n_stab)
    eng = SweepEngine([spec_from(k, beta, L, u, f)], mode="dqmc", precision=precision, delay=delay,
                      n_stab=n_stab, rng="replay", trace=True, flip_groups=flip_groups)
    g0 = eng.green(0, which=0)
    init_err = max(np.abs(g0[s] - o.g[s]).max() for s in range(2))
    rs = np.random.RandomState(100 + seed)
    worst = 0.0
    s_sum = np.zeros((n, n))
    obs_ref = []
    hop = k - np.diag(np.diag(k))
    for s in range(sweeps):
        uni = rs.rand(L * n)
        res = o.sweep(uni)
        s_sum += dqmc.spin_correlation(*res["g_meas"])
        ob = dqmc.observables(res["g_meas"][0], res["g_meas"][1], hop)
        obs_ref.append([ob["density"], ob["docc"], ob["ekin"]])
        eng.set_uniforms(0, uni)
        eng.sweep(1, measure=True)
        eng.sync()
        ratio, acc = eng.trace(0)
        ok, ties = masked_equal(acc, res["accept"], res["ratio"], uni)
        if precision == "fp64":
            assert ok, f"sweep {s}: accept bits differ"
            assert np.array_equal(eng.field(0), o.field)
        gl = eng.green(0, which=2)
        gc = eng.green(0, which=0)
        worst = max(worst, max(np.abs(gl[sp] - res["g_meas"][sp]).max() for sp in range(2)),
                    max(np.abs(gc[sp] - o.g[sp]).max() for sp in range(2)))
    if precision == "fp64":
        # measurement layer: per-sweep estimators and the spin-correlation accumulator
        worst = max(worst, np.abs(eng.spin_correlation(0) - s_sum / sweeps).max(),
                    np.abs(eng.observables(0) - np.array(obs_ref)).max())
    st = eng.stats(0)
    eng.close()
    return init_err, worst, st


@pytest.mark.parametrize("side,L,beta,n_stab,delay", [
    (2, 16, 2.0, 8, 1), (4, 16, 2.0, 8, 1), (4, 16, 2.0, 8, 4), (4, 16, 2.0, 5, 16),
    (8, 40, 4.0, 10, 1), (8, 40, 4.0, 10, 8), (6, 20, 2.0, 10, 32),
])
def test_dqmc_fp64_matches_oracle(side, L, beta, n_stab, delay):
    _say(f"[dqmc fp64] side={side} L={L} beta={beta} n_stab={n_stab} delay={delay}")
    init_err, worst, st = _run_dqmc_vs_oracle(side, L, beta, 4.0, n_stab, delay, sweeps=2)
    assert init_err < 1e-8
    assert worst < 1e-8
    assert st["flags"] == 0
    assert st["max_err"] < 1e-8


@pytest.mark.parametrize("side,L,beta,n_stab,delay,groups", [
    (8, 40, 4.0, 10, 16, 2), (10, 20, 2.0, 10, 16, 3), (6, 20, 2.0, 10, 32, 3),
    (8, 40, 4.0, 10, 1, 2), (10, 20, 2.0, 10, 1, 3),
])
def test_dqmc_split_flip_matches_oracle(side, L, beta, n_stab, delay, groups):
    """The delayed flip with several workgroups per chain (per-block barriers, deferred
    own-row strips), and the rank-1 flip split over several workgroups (a barrier per accepted
    flip, the pending row written after it), against the oracle, as the one-workgroup forms."""
    _say(f"[dqmc split] side={side} L={L} delay={delay} groups={groups}")
    init_err, worst, st = _run_dqmc_vs_oracle(side, L, beta, 4.0, n_stab, delay, sweeps=2, flip_groups=groups)
    assert init_err < 1e-8
    assert worst < 1e-8
    assert st["flags"] == 0
    assert st["max_err"] < 1e-8


@pytest.mark.parametrize("delay", [16, 1])
def test_split_flip_batched_equals_single_group(delay):
    """Uneven chains, 2 workgroups per chain: bitwise equal to one workgroup per chain."""
    _say(f"[split batched] delay={delay}")
    cases = [(8, 16, 2.0, 1), (6, 20, 2.0, 2), (7, 8, 1.0, 3), (10, 12, 1.5, 4)]
    specs = []
    for side, L, beta, seed in cases:
        k, f = square_case(side, L, beta, 4.0, seed)
        specs.append(spec_from(k, beta, L, 3.0 + seed, f, seed=seed))
    one = SweepEngine(specs, mode="dqmc", delay=delay, n_stab=5, rng="philox", flip_groups=1)
    two = SweepEngine(specs, mode="dqmc", delay=delay, n_stab=5, rng="philox", flip_groups=2)
    for e in (one, two):
        e.sweep(3, measure=True)
        e.sync()
    for i in range(len(specs)):
        assert np.array_equal(one.green(i, which=1), two.green(i, which=1)), i
        assert np.array_equal(one.green(i, which=0), two.green(i, which=0)), i
        assert np.array_equal(one.field(i), two.field(i)), i
        assert two.stats(i)["flags"] == 0
    one.close()
    two.close()


def test_dqmc_u0_closed_form():
    _say("[dqmc u0]")
    k = model.square_k(4, 4)
    f = model.random_field(16, 16, 0)
    eng = SweepEngine([spec_from(k, 2.0, 16, 0.0, f)], mode="dqmc", n_stab=8, rng="philox")
    eng.sweep(2, measure=True)
    eng.sync()
    free = exact.free_g(k, 2.0)
    g = eng.green(0, which=1)
    assert np.abs(g[0] - free).max() < 1e-10 and np.abs(g[1] - free).max() < 1e-10
    st = eng.stats(0)
    assert st["accepted"] == st["attempted"] == 2 * 16 * 16
    eng.close()


def test_philox_stream_matches_oracle():
    _say("[philox]")
    side, L, beta = 4, 16, 2.0
    k, f = square_case(side, L, beta, 4.0, 1)
    seed = 987654321
    o = dqmc.DQMCOracle(k, beta, L, 4.0, f, n_stab=8)
    spec = spec_from(k, beta, L, 4.0, f, seed=seed)
    eng = SweepEngine([spec], mode="dqmc", n_stab=8, rng="philox", trace=True)
    for s in range(2):
        uni = philox.uniforms(seed, s * L * 16, L * 16)
        res = o.sweep(uni)
        eng.sweep(1)
        eng.sync()
        ratio, acc = eng.trace(0)
        ok, _ = masked_equal(acc, res["accept"], res["ratio"], uni)
        assert ok
    assert np.array_equal(eng.field(0), o.field)
    eng.close()


def test_batched_uneven_equals_alone():
    _say("[batched]")
    cases = [(4, 16, 2.0, 1), (6, 20, 2.0, 2), (2, 8, 1.0, 3), (4, 12, 1.5, 4)]
    specs = []
    for side, L, beta, seed in cases:
        k, f = square_case(side, L, beta, 4.0, seed)
        specs.append(spec_from(k, beta, L, 3.0 + seed, f, seed=seed))
    batched = SweepEngine(specs, mode="dqmc", n_stab=5, rng="philox")
    batched.sweep(3, measure=True)
    batched.sync()
    for i, sp in enumerate(specs):
        alone = SweepEngine([sp], mode="dqmc", n_stab=5, rng="philox")
        alone.sweep(3, measure=True)
        alone.sync()
        assert np.array_equal(alone.green(0, which=1), batched.green(i, which=1)), i
        assert np.array_equal(alone.green(0, which=0), batched.green(i, which=0)), i
        assert np.array_equal(alone.field(0), batched.field(i)), i
        assert alone.stats(0)["accepted"] == batched.stats(i)["accepted"]
        alone.close()
    batched.close()


def test_dqmc_fp32_delayed_tracks_fp64():
    _say("[dqmc fp32 delayed]")
    side, L, beta, n_stab = 4, 16, 2.0, 4
    k, f = square_case(side, L, beta, 4.0, 2)
    n = side * side
    o = dqmc.DQMCOracle(k, beta, L, 4.0, f, n_stab=n_stab)
    eng = SweepEngine([spec_from(k, beta, L, 4.0, f)], mode="dqmc", precision="fp32", delay=8,
                      n_stab=n_stab, rng="replay", trace=True)
    rs = np.random.RandomState(7)
    uni = rs.rand(L * n)
    res = o.sweep(uni)
    eng.set_uniforms(0, uni)
    eng.sweep(1, measure=True)
    eng.sync()
    ratio, acc = eng.trace(0)
    # FP32 G deviates ~1e-6 from FP64: accept bits agree away from near-ties
    ok, ties = masked_equal(acc, res["accept"], res["ratio"], uni, tol=1e-3)
    assert ok
    st = eng.stats(0)
    assert 0 < st["max_err"] < 1e-3 and st["flags"] == 0
    eng.close()


def test_engine_refuses_oversized_lds():
    _say("[lds limit]")
    from latticeqmc_amd._native import LqmcError
    k, f = square_case(16, 8, 1.0, 4.0, 0)
    with pytest.raises(LqmcError, match="delay"):
        SweepEngine([spec_from(k, 1.0, 8, 4.0, f)], mode="dqmc", delay=64, n_stab=4, rng="philox")
    # 28x28 sites: the QR column panel (8 columns of 784 rows plus workspace) exceeds 64 KB.
    k, f = square_case(28, 4, 1.0, 4.0, 0)
    with pytest.raises(LqmcError, match="LDS"):
        SweepEngine([spec_from(k, 1.0, 4, 4.0, f)], mode="dqmc", delay=16, n_stab=4, rng="philox")


def test_physics_2x2_matches_trotter_exact():
    """Mode D sampling on the GPU reproduces the Trotterised exact 2x2 (4-site ring)
    values at U=4, beta=2, L=16: double occupancy 0.084755 and kinetic energy per site
    -0.727136 (oracle/exact.py), within 4 standard errors over 32 independent chains."""
    _say("[physics 2x2]")
    from latticeqmc_amd import tools
    k = model.square_k(2, 2)
    ex = exact.Hubbard(k, 4.0).trotter_expect(2.0, 16)
    specs = [spec_from(k, 2.0, 16, 4.0, model.random_field(4, 16, c), seed=1000 + c) for c in range(32)]
    eng = SweepEngine(specs, mode="dqmc", n_stab=8, rng="philox", graph=True)
    eng.sweep(100, measure=False)
    eng.sweep(400, measure=True)
    eng.sync()
    per_chain = np.array([eng.observables(c)[:, 1:].mean(axis=0) for c in range(32)])  # docc, ekin
    eng.close()
    for col, name in ((0, "docc"), (1, "ekin")):
        est, err = tools.jackknife(lambda x: x, per_chain[:, col], n_bins=32)
        assert abs(est - ex[name]) < 4 * err + 1e-4, (name, est, err, ex[name])


@pytest.mark.parametrize("side,delay", [(4, None), (8, None), (4, 1)])
def test_drop_in_numpy_stream_matches_oracle(side, delay):
    """LatticeQMC(mode="dqmc", rng="numpy").run() -- by default the delayed form (8-site blocks
    at 4x4, 16 at 8x8) replayed from captured hipGraphs -- consumes numpy's global stream in the
    reference's order (the field at construction, then N*L uniforms per sweep, lqmc.py:309-317);
    the oracle replays the same stream: the same final field and the measured mean G within 1e-8
    (the return value of measure_loop, lqmc.py:356-375)."""
    _say(f"[drop-in numpy stream] side={side} delay={delay}")
    from latticeqmc_amd import HubbardModel, LatticeQMC
    m = HubbardModel(u=4, t=1, mu=0)
    m.build_square(side)
    np.random.seed(21)
    q = LatticeQMC(m, beta=2.0, time_steps=16, warmup=2, sweeps=3, n_stab=8, delay=delay, console=False)
    assert q.graph and q.delay == ({4: 8, 8: 16}[side] if delay is None else delay)
    f = q.config.config.copy()
    state = np.random.get_state()
    gf = q.run()
    np.random.set_state(state)
    o = dqmc.DQMCOracle(q.ham_kin, 2.0, 16, 4.0, f, n_stab=8, delay=q.delay)
    n = q.n_sites
    for _ in range(2):
        o.sweep(np.random.rand(16 * n))
    total = np.zeros((2, n, n))
    for _ in range(3):
        res = o.sweep(np.random.rand(16 * n))
        total += np.array(res["g_meas"])
    assert np.array_equal(q.config.config, o.field)
    assert np.abs(gf - total / 3).max() < 1e-8


def test_dqmc_bf16x3_tracks_fp64():
    """The split-BF16 form (FP32 G, rank-k products as three BF16 matrix-core products) on
    one replayed sweep: accept bits equal to the FP64 oracle's away from |ratio - u| < 1e-2,
    every recompute within 1e-2 of the FP64 G, and a larger deviation than the FP32 form's on
    the same trajectory (the products carry ~2^-16 relative error against FP32's 2^-24)."""
    _say("[dqmc bf16x3]")
    side, L, beta, n_stab = 4, 16, 2.0, 2
    k, f = square_case(side, L, beta, 4.0, 2)
    n = side * side
    o = dqmc.DQMCOracle(k, beta, L, 4.0, f, n_stab=n_stab)
    uni = np.random.RandomState(7).rand(L * n)
    res = o.sweep(uni)
    errs = {}
    for prec in ("fp32", "bf16x3"):
        eng = SweepEngine([spec_from(k, beta, L, 4.0, f)], mode="dqmc", precision=prec, delay=8,
                          n_stab=n_stab, rng="replay", trace=True)
        eng.set_uniforms(0, uni)
        eng.sweep(1, measure=True)
        eng.sync()
        ratio, acc = eng.trace(0)
        ok, ties = masked_equal(acc, res["accept"], res["ratio"], uni, tol=1e-2)
        assert ok, prec
        st = eng.stats(0)
        assert st["flags"] == 0 and 0 < st["max_err"] < 1e-2, (prec, st["max_err"])
        errs[prec] = st["max_err"]
        eng.close()
    _say(f"  max|dG| fp32 {errs['fp32']:.2e}, bf16x3 {errs['bf16x3']:.2e}")
    assert errs["bf16x3"] > 2 * errs["fp32"]
