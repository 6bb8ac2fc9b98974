"""Mode R oracle == the reference's _update_step, bit for bit, on the reference-generated fixtures."""
import numpy as np
import pytest

from oracle import mode_r, model


@pytest.mark.parametrize("name", ["mode_r_2x2", "mode_r_4x4", "mode_r_4x4_u0"])
def test_mode_r_matches_reference_trajectory(golden_dir, name):
    g = np.load(f"{golden_dir}/{name}.npz")
    _, lam, ek, _ = model.params(g["K"], float(g["beta"]), int(g["L"]), float(g["u"]))
    f = g["field0"].copy()
    meas = np.zeros_like(g["meas_g"])
    for s in range(g["uniforms"].shape[0]):
        gu, gd, r, a = mode_r.sweep(f, ek, lam, g["uniforms"][s])
        assert np.array_equal(a, g["acc"][s])
        assert np.array_equal(r, g["ratio"][s])
        assert np.array_equal(f, g["field"][s])
        assert np.array_equal(gu, g["g_up"][s]) and np.array_equal(gd, g["g_dn"][s])
        meas += np.asarray([gu, gd])
    assert np.allclose(meas / g["uniforms"].shape[0], g["meas_g"], rtol=0, atol=1e-12)


def test_mode_r_equals_textbook_at_u0(golden_dir):
    """At U = 0 the reference sweep is exact: G = (I + e^{-beta K})^-1 (closed form)."""
    from oracle.exact import free_g
    g = np.load(f"{golden_dir}/mode_r_4x4_u0.npz")
    exact = free_g(g["K"], float(g["beta"]))
    # the end-of-sweep G is G(l=0) of the reference's K.V ordering, equal to the closed form at U=0
    assert np.abs(g["g_up"][0] - exact).max() < 1e-10
    assert np.abs(g["g_dn"][0] - exact).max() < 1e-10


def test_faithful_sweep_same_accepts(golden_dir):
    g = np.load(f"{golden_dir}/mode_r_2x2.npz")
    _, lam, ek, _ = model.params(g["K"], float(g["beta"]), int(g["L"]), float(g["u"]))
    f = g["field0"].copy()
    n = mode_r.sweep_faithful(f, ek, lam, g["uniforms"][0])
    assert n == g["uniforms"].shape[1]
    assert np.array_equal(f, g["field"][0])
