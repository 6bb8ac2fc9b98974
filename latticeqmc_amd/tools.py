"""Observable helpers with the reference's semantics (lqmc/tools.py:41-124)."""
import numpy as np


def check_params(u, t, dtau):
    """Warn when U t dtau^2 >= 0.1 (lqmc/tools.py:41-60)."""
    check_val = u * t * dtau ** 2
    if check_val < 0.1:
        print(f"Check-value {check_val:.2} is smaller than 0.1!")
    else:
        print(f"Check-value {check_val:.2} should be smaller than 0.1!")


def filling(g_sigma, site=None, axis1=-2, axis2=-1):
    """n_sigma,i = 1 - G_sigma,ii (lqmc/tools.py:63-84)."""
    n = 1 - np.diagonal(g_sigma, axis1=axis1, axis2=axis2)
    if site is not None:
        n = n[..., site]
    return n


def local_moment(gf_up, gf_dn, site=None, axis1=-2, axis2=-1):
    """n_up + n_dn - 2 n_up n_dn from sweep-averaged G (lqmc/tools.py:87-109)."""
    n_up = filling(gf_up, site, axis1, axis2)
    n_dn = filling(gf_dn, site, axis1, axis2)
    return n_up + n_dn - 2 * n_up * n_dn


def local_gf(gf):
    return np.diagonal(gf, axis1=-2, axis2=-1)


def binned_mean(samples, n_bins=20):
    """Mean and binned standard error of a per-sweep series (axis 0)."""
    x = np.asarray(samples, dtype=np.float64)
    nb = max(1, min(n_bins, x.shape[0]))
    usable = (x.shape[0] // nb) * nb
    bins = x[:usable].reshape(nb, -1, *x.shape[1:]).mean(axis=1)
    mean = bins.mean(axis=0)
    err = bins.std(axis=0, ddof=1) / np.sqrt(nb) if nb > 1 else np.zeros_like(mean)
    return mean, err


def jackknife(func, *series, n_bins=20):
    """Jackknife mean and error of func(mean(series_1), mean(series_2), ...) over binned
    per-sweep series (axis 0 = sweeps), e.g. a ratio or a susceptibility of means.
    Returns (estimate, error); the estimate is bias-corrected, n f(all) - (n-1) mean_k f(-k)."""
    xs = [np.asarray(x, dtype=np.float64) for x in series]
    n_s = min(x.shape[0] for x in xs)
    nb = max(2, min(n_bins, n_s))
    usable = (n_s // nb) * nb
    bins = [x[:usable].reshape(nb, -1, *x.shape[1:]).mean(axis=1) for x in xs]
    total = [b.mean(axis=0) for b in bins]
    full = np.asarray(func(*total), dtype=np.float64)
    leave = np.array([func(*[(t * nb - b[k]) / (nb - 1) for t, b in zip(total, bins)]) for k in range(nb)])
    mean_leave = leave.mean(axis=0)
    est = nb * full - (nb - 1) * mean_leave
    err = np.sqrt((nb - 1) / nb * np.sum((leave - mean_leave) ** 2, axis=0))
    return est, err


def structure_factor(spin_corr, positions, q):
    """S(q) = (1/N) sum_ij e^{i q (r_i - r_j)} <S^z_i S^z_j> from an (N, N) equal-time spin
    correlation (e.g. SweepEngine.spin_correlation) and site positions (N, dim)."""
    r = np.asarray(positions, dtype=np.float64)
    phase = np.exp(1j * (r @ np.asarray(q, dtype=np.float64)))
    c = np.asarray(spin_corr, dtype=np.float64)
    return float(np.real(np.conj(phase) @ c @ phase) / c.shape[0])


def local_moment_from_docc(density, docc):
    """<m_z^2> = <n> - 2 <n_up n_dn> per site from the per-sweep density and double occupancy
    (the sweep-averaged estimator; the reference's local_moment, tools.py:87-109, uses
    products of averaged fillings instead)."""
    return np.asarray(density) - 2.0 * np.asarray(docc)


# ---- free-fermion Green's functions (the reference's GF tools, tools.py:127-200) ----------------

def matsubara_frequencies(points, beta):
    """Fermionic Matsubara frequencies i w_n = i pi (2 n + 1) / beta (tools.py:127-142)."""
    n = np.asarray(points)
    if not np.issubdtype(n.dtype, np.integer):
        raise TypeError("Matsubara indices must be integers")
    return 1j * np.pi * (2 * n + 1) / beta


def fermi_fct(eps, beta):
    """Fermi function 1 / (exp(beta eps) + 1), written as (1 - tanh(beta eps / 2)) / 2 so that
    large |beta eps| neither overflows nor loses the small tail (tools.py:145-167)."""
    return 0.5 * (1.0 - np.tanh(0.5 * beta * np.asarray(eps)))


def decompose(a):
    """Eigen-decomposition of a Hermitian matrix as (eigenvectors, eigenvalues, inverse
    eigenvectors) (tools.py:170-172)."""
    xi, rv = np.linalg.eigh(a)
    return rv, xi, np.linalg.inv(rv)


def reconstruct(rv, xi, rv_inv, diag=False):
    """rv diag(xi) rv_inv over the trailing axes of xi, or only its diagonal (tools.py:175-179)."""
    xi = np.asarray(xi)
    if diag:
        return np.einsum("ij,...j,ji->...i", rv, xi, rv_inv)
    return (rv * xi[..., np.newaxis, :]) @ rv_inv


def pole_gf_tau(tau, poles, weights, beta):
    """Imaginary-time Green's function of a sum of poles, 0 <= tau <= beta (tools.py:182-190):
    G(tau) = -sum_p w_p exp(-tau e_p) (1 - f(e_p)), evaluated as exp((beta - tau) e_p) f(e_p)
    for negative poles so that neither branch overflows."""
    poles, weights = np.atleast_1d(*np.broadcast_arrays(poles, weights))
    tau = np.asarray(tau, dtype=np.float64)
    if np.any((tau < 0.0) | (tau > beta)):
        raise ValueError("tau must lie in [0, beta]")
    t = tau.reshape(tau.shape + (1,) * poles.ndim)
    pos = poles.real >= 0
    expo = np.where(pos, -t * poles, (beta - t) * poles)
    occ = np.where(pos, fermi_fct(-poles, beta), fermi_fct(poles, beta))
    return -np.sum(weights * np.exp(expo) * occ, axis=-1)


def compute_pole_gf_tau(ham, beta, n_tau=2049):
    """Free Green's function of the one-body Hamiltonian ``ham`` on n_tau points of [0, beta];
    returns (tau, G) with G of shape (site, site, tau) (tools.py:193-200).  At tau = 0+ it is
    -(I + exp(-beta ham))^-1, minus the equal-time G that the sweep measures at U = 0."""
    rv, xi, rv_inv = decompose(ham)
    tau = np.linspace(0.0, beta, num=n_tau)
    diag = pole_gf_tau(tau, xi[..., np.newaxis], weights=1.0, beta=beta)
    gf = reconstruct(rv, diag, rv_inv)
    return tau, np.moveaxis(gf, 0, -1)
