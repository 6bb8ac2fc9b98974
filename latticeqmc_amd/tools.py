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
