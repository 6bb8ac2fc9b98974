"""Oracle model quantities: K, E_K = expm(-dtau K), lambda (test infrastructure only).

K is rebuilt here from the closed form of the reference's square lattice
(site i = x*h + y, lattice.py:226-242; periodic bonds lattice.py:994-1033;
K = -mu I - t A, hubbard.py:91-111) independently of the product's builder.
Parameters follow lqmc.py:93-117: dtau = beta/L, lambda = arccosh(exp(U dtau/2))
(0 when U == 0), E_K = scipy expm(-dtau K).
"""
import numpy as np
from scipy.linalg import expm


def square_k(w, h, t=1.0, mu=0.0, periodic=(True, True)):
    n = w * h
    a = np.zeros((n, n))
    for x in range(w):
        for y in range(h):
            i = x * h + y
            for dx, dy, ax in ((1, 0, 0), (0, 1, 1)):
                ext = (w, h)[ax]
                nx, ny = x + dx, y + dy
                if (nx, ny)[ax] >= ext:
                    if not periodic[ax] or ext < 2:
                        continue
                    nx, ny = nx % w, ny % h
                j = nx * h + ny
                if j != i:
                    a[i, j] = a[j, i] = 1.0
    return -mu * np.eye(n) - t * a


def params(k, beta, L, u):
    dtau = beta / L
    lam = float(np.arccosh(np.exp(u * dtau / 2.0))) if u else 0.0
    expk = expm(-dtau * k)
    expk_inv = expm(dtau * k)
    return dtau, lam, expk, expk_inv


def random_field(n, L, seed):
    rs = np.random.RandomState(seed)
    return (2 * rs.randint(0, 2, size=(n, L)) - 1).astype(np.int8)
