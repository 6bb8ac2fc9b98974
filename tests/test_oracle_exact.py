"""Exact-diagonalisation oracle pinned by free-fermion formulas and known 2x2 values."""
import numpy as np

from oracle import exact, model


def test_u0_ed_equals_free_fermions():
    k = model.square_k(2, 2, mu=0.3)
    h = exact.Hubbard(k, 0.0)
    for beta in (0.7, 2.0):
        ex = h.thermal_expect(beta)
        g = exact.free_g(k, beta)
        n = 1.0 - np.diag(g)
        assert np.isclose(ex["density"], 2 * n.mean(), atol=1e-12)
        assert np.isclose(ex["docc"], np.mean(n * n), atol=1e-12)
        hop = k - np.diag(np.diag(k))
        assert np.isclose(ex["ekin"], 2 * np.sum(hop * (np.eye(4) - g.T)) / 4, atol=1e-12)
    # Trotter exact at U=0 is exact for any L (H_U = 0 up to a constant)
    assert np.isclose(h.trotter_expect(2.0, 3)["docc"], h.thermal_expect(2.0)["docc"], atol=1e-12)


def test_2x2_u4_beta2_values():
    k = model.square_k(2, 2)
    h = exact.Hubbard(k, 4.0)
    tr = h.trotter_expect(2.0, 16)
    th = h.thermal_expect(2.0)
    assert abs(tr["docc"] - 0.084755) < 1e-6 and abs(tr["ekin"] + 0.727136) < 1e-6
    assert abs(th["docc"] - 0.087316) < 1e-6 and abs(th["ekin"] + 0.717419) < 1e-6
    assert abs(tr["density"] - 1.0) < 1e-12
    # Trotter error shrinks towards the continuous value
    assert abs(h.trotter_expect(2.0, 64)["docc"] - th["docc"]) < abs(tr["docc"] - th["docc"])


def test_two_site_atomic_limit():
    # t = 0: independent sites, D = e^{beta U/4} / (2 e^{beta U/4} + 2 e^{-beta U/4})
    k = np.zeros((2, 2))
    u, beta = 3.0, 1.3
    d = exact.Hubbard(k, u).thermal_expect(beta)["docc"]
    ref = np.exp(-beta * u / 4) / (2 * np.exp(-beta * u / 4) + 2 * np.exp(beta * u / 4))
    assert np.isclose(d, ref, atol=1e-12)
