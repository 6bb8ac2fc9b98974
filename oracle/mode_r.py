"""Mode R: literal restatement of the reference fast sweep (test infrastructure only).

Follows lqmc/lqmc.py:301-347 step by step (the math is reproduced as is,
including the coefficients and the wrap that make it inconsistent for U != 0):

  G_s = inv(I + E_K V_s(L-1) ... E_K V_s(0))            (lqmc.py:303-307, 156-185)
  for l = L-1..0, i = 0..N-1:                          (309-311)
     arg = 2 lambda s[i,l]
     r = [1 + (1-Gu_ii)(e^{arg}-1)] [1 + (1-Gd_ii)(e^{-arg}-1)]   (313-316)
     accept iff u <= r                                 (317)
     on accept: c = -(e^{-+arg}-1) G[i,:]; c_i += (e^{-+arg}-1);
                e = G[:,i]/(1+c_i); G -= outer(e, c); s[i,l] *= -1   (320-333)
     if l > 0: W = V(l-1) E_K; G = (W G) inv(W)        (338-345)
  return (Gu, Gd)

``np.outer`` gives the same doubles as the reference's j,k loop (each entry is
one product and one subtraction, no fused multiply-add).  ``sweep_faithful``
keeps the pure-Python j,k loop and the per-attempt ``_debug`` mean/var cost
(lqmc.py:217-232) and is only used as the timed CPU baseline.
"""
import numpy as np


def expv(field, lam, l, sigma):
    """Diagonal of exp(V_sigma(l)) (lqmc.py:149-154)."""
    return np.exp(-1 * sigma * lam * field[:, l])


def get_m(field, expk, lam, l0, sigma):
    """M = I + prod_l E_K V(l) in the reference's cyclic order (lqmc.py:156-185)."""
    L = field.shape[1]
    l0 = l0 % L
    indices = list(reversed(range(L)))
    order = indices[-l0:] + indices[:-l0]
    b_prod = None
    for l in order:
        b = np.dot(expk, np.diagflat(expv(field, lam, l, sigma)))
        b_prod = b if b_prod is None else np.dot(b_prod, b)
    return np.eye(field.shape[0]) + b_prod


def sweep(field, expk, lam, uniforms, max_slices=None):
    """One Mode R sweep; mutates ``field``. Returns (g_up, g_dn, ratios, accepts).
    ``max_slices`` stops after that many slices (a bounded sample for timing)."""
    n, L = field.shape
    g_up = np.linalg.inv(get_m(field, expk, lam, 0, +1))
    g_dn = np.linalg.inv(get_m(field, expk, lam, 0, -1))
    ratios = np.empty(n * L)
    accs = np.zeros(n * L, dtype=bool)
    q = 0
    for l in reversed(range(L)):
        if max_slices is not None and L - 1 - l >= max_slices:
            break
        for i in range(n):
            arg = 2 * lam * field[i, l]
            d_up = 1 + (1 - g_up[i, i]) * (np.exp(+arg) - 1)
            d_dn = 1 + (1 - g_dn[i, i]) * (np.exp(-arg) - 1)
            ratio = d_up * d_dn
            acc = uniforms[q] <= ratio
            ratios[q] = ratio
            accs[q] = acc
            q += 1
            if acc:
                c_up = -(np.exp(-arg) - 1) * g_up[i, :]
                c_up[i] += (np.exp(-arg) - 1)
                c_dn = -(np.exp(+arg) - 1) * g_dn[i, :]
                c_dn[i] += (np.exp(+arg) - 1)
                e_up = g_up[:, i] / (1 + c_up[i])
                e_dn = g_dn[:, i] / (1 + c_dn[i])
                g_up = g_up - np.outer(e_up, c_up)
                g_dn = g_dn - np.outer(e_dn, c_dn)
                field[i, l] *= -1
        if l > 0:
            b_up = np.dot(np.diagflat(expv(field, lam, l - 1, +1)), expk)
            b_dn = np.dot(np.diagflat(expv(field, lam, l - 1, -1)), expk)
            g_up = np.dot(np.dot(b_up, g_up), np.linalg.inv(b_up))
            g_dn = np.dot(np.dot(b_dn, g_dn), np.linalg.inv(b_dn))
    return g_up, g_dn, ratios, accs


def sweep_faithful(field, expk, lam, uniforms, max_slices=None):
    """Reference-speed restatement: pure-Python j,k loop and per-attempt mean/var.

    Processes at most ``max_slices`` slices (a bounded sample for timing) and
    returns the number of attempted updates.
    """
    n, L = field.shape
    g_up = np.linalg.inv(get_m(field, expk, lam, 0, +1))
    g_dn = np.linalg.inv(get_m(field, expk, lam, 0, -1))
    q = 0
    done = 0
    log = ""
    for l in reversed(range(L)):
        if max_slices is not None and done >= max_slices:
            break
        for i in range(n):
            arg = 2 * lam * field[i, l]
            d_up = 1 + (1 - g_up[i, i]) * (np.exp(+arg) - 1)
            d_dn = 1 + (1 - g_dn[i, i]) * (np.exp(-arg) - 1)
            ratio = d_up * d_dn
            acc = uniforms[q % len(uniforms)] <= ratio
            q += 1
            if acc:
                c_up = -(np.exp(-arg) - 1) * g_up[i, :]
                c_up[i] += (np.exp(-arg) - 1)
                c_dn = -(np.exp(+arg) - 1) * g_dn[i, :]
                c_dn[i] += (np.exp(+arg) - 1)
                e_up = g_up[:, i] / (1 + c_up[i])
                e_dn = g_dn[:, i] / (1 + c_dn[i])
                for j in range(n):
                    for k in range(n):
                        g_up[j, k] = g_up[j, k] - e_up[j] * c_up[k]
                        g_dn[j, k] = g_dn[j, k] - e_dn[j] * c_dn[k]
                field[i, l] *= -1
            # per-attempt debug string cost of the reference (lqmc.py:230-232)
            log = f"{l:>2} {i:>2} -- {ratio:.1f} ({acc}) -- {np.mean(field):.3f} {np.var(field):.3f}"
        if l > 0:
            b_up = np.dot(np.diagflat(expv(field, lam, l - 1, +1)), expk)
            b_dn = np.dot(np.diagflat(expv(field, lam, l - 1, -1)), expk)
            g_up = np.dot(np.dot(b_up, g_up), np.linalg.inv(b_up))
            g_dn = np.dot(np.dot(b_dn, g_dn), np.linalg.inv(b_dn))
        done += 1
    del log
    return q


def det_sweep(field, expk, lam, uniforms, old_det, ratios=None):
    """Brute-force determinant sweep (lqmc.py:236-259). Returns (old_det, accepts); the
    acceptance ratios are written into ``ratios`` (length N L) when it is given."""
    n, L = field.shape
    accs = np.zeros(n * L, dtype=bool)
    q = 0
    for l in reversed(range(L)):
        for i in range(n):
            field[i, l] *= -1
            new_det = (np.linalg.det(get_m(field, expk, lam, l, +1))
                       * np.linalg.det(get_m(field, expk, lam, l, -1)))
            ratio = new_det / old_det
            acc = uniforms[q] <= ratio
            accs[q] = acc
            if ratios is not None:
                ratios[q] = ratio
            q += 1
            if acc:
                old_det = new_det
            else:
                field[i, l] *= -1
    return old_det, accs
