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
