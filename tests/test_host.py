"""Host-side pieces: Configuration RNG semantics and the C-ABI library exports (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from latticeqmc_amd.configuration import Configuration
from latticeqmc_amd.tools import binned_mean, local_moment

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_configuration_draw_matches_reference_fixture(golden_dir):
    g = np.load(f"{golden_dir}/mode_r_4x4.npz")
    np.random.seed(int(g["seed"]))
    c = Configuration(16, 16)
    assert c.config.dtype == np.int8
    assert np.array_equal(c.config, g["field0"])
    c.update(3, 5)
    assert c[3, 5] == -g["field0"][3, 5]


def test_uniform_stream_vectorised_equals_scalar_draws():
    a = np.random.RandomState(7)
    b = np.random.RandomState(7)
    assert np.array_equal([a.rand() for _ in range(257)], b.rand(257))


def test_local_moment_and_binning():
    g = np.diag([0.25, 0.75])
    assert np.allclose(local_moment(g, g), [0.75 + 0.75 - 2 * 0.5625, 0.25 + 0.25 - 2 * 0.0625])
    m, e = binned_mean(np.arange(100.0), 10)
    assert m == 49.5 and e > 0


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "lqmc_hip.h")).read()
    return sorted(set(re.findall(r"\b(lqmc_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from latticeqmc_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("liblqmc_hip.so is not built; run python -m latticeqmc_amd.build")
    lib = ctypes.CDLL(_native.LIB_PATH)
    decl = _declared_symbols()
    assert len(decl) >= 18
    for name in decl:
        assert hasattr(lib, name), name
    assert set(_native.EXPORTS) == set(decl)


def test_native_load_declares_prototypes():
    from latticeqmc_amd import _native
    lib = _native.load()
    assert lib.lqmc_last_error() is not None


def test_jackknife_and_structure_factor():
    from latticeqmc_amd import tools
    rs = np.random.RandomState(0)
    a = 2.0 + rs.randn(4000)
    b = 4.0 + rs.randn(4000)
    est, err = tools.jackknife(lambda x, y: x / y, a, b, n_bins=40)
    assert abs(est - 0.5) < 4 * err and 0.002 < err < 0.02
    m, e = tools.jackknife(lambda x: x, a, n_bins=40)
    bm, be = tools.binned_mean(a, n_bins=40)
    assert abs(m - bm) < 1e-12 and abs(e - be) < 1e-12      # linear f: jackknife == binning
    # Neel pattern on a 4x4 square: S(pi, pi) = N, S(0, 0) = 0
    pos = np.array([(x, y) for x in range(4) for y in range(4)], dtype=float)
    s = np.array([1.0 if (x + y) % 2 == 0 else -1.0 for x, y in pos])
    c = np.outer(s, s)
    assert abs(tools.structure_factor(c, pos, (np.pi, np.pi)) - 16.0) < 1e-12
    assert abs(tools.structure_factor(c, pos, (0.0, 0.0))) < 1e-12


def test_det_mode_rejects_fp32_and_delay_before_device_use():
    """Option checks run before any device call, so they hold on a host without a GPU."""
    from latticeqmc_amd._native import LqmcError
    from latticeqmc_amd.engine import ChainSpec, SweepEngine
    k = np.zeros((4, 4))
    spec = ChainSpec(k, 1.0, 4, 4.0, np.ones((4, 4), dtype=np.int8), expk=np.eye(4), expk_inv=np.eye(4), lam=0.5)
    for kw in ({"precision": "fp32"}, {"delay": 4}):
        with pytest.raises(LqmcError, match="determinant mode"):
            SweepEngine([spec], mode="det", rng="replay", **kw)


def test_free_gf_tools_against_closed_forms():
    """compute_pole_gf_tau at tau = 0 and beta against (I + e^{-beta K})^-1 (the U = 0 equal-time
    G), reconstruct(decompose(K)) = K, the Fermi function's symmetry and the Matsubara grid."""
    from latticeqmc_amd import tools
    from oracle import exact, model
    k = model.square_k(3, 3) - 0.7 * np.eye(9)
    beta = 4.0
    tau, g = tools.compute_pole_gf_tau(k, beta, n_tau=33)
    g_eq = exact.free_g(k, beta)
    assert np.abs(g[:, :, 0] + g_eq).max() < 1e-12
    assert np.abs(g[:, :, -1] - (g_eq - np.eye(9))).max() < 1e-12  # -f(K) = G - I
    # a single pole: G(tau) = -exp(-tau e) (1 - f(e))
    e = 1.3
    t = np.array([0.0, 0.5, 2.0])
    want = -np.exp(-t * e) / (1.0 + np.exp(-beta * e))
    assert np.allclose(tools.pole_gf_tau(t, e, 1.0, beta), want, rtol=1e-13, atol=0)
    rv, xi, rvi = tools.decompose(k)
    assert np.abs(tools.reconstruct(rv, xi, rvi) - k).max() < 1e-12
    assert np.abs(tools.reconstruct(rv, xi, rvi, diag=True) - np.diag(k)).max() < 1e-12
    x = np.linspace(-3, 3, 7)
    assert np.allclose(tools.fermi_fct(x, 2.0) + tools.fermi_fct(-x, 2.0), 1.0)
    assert np.allclose(tools.matsubara_frequencies(np.arange(3), 2.0), 1j * np.pi * np.array([1, 3, 5]) / 2.0)
    with pytest.raises(ValueError):
        tools.pole_gf_tau(-0.1, e, 1.0, beta)
