"""Fock-space exact results for tiny Hubbard lattices (test infrastructure only).

The reference's exact.py (exact.py:27-54) only evaluates the U=0 pole formula
G(tau) (lqmc/tools.py:182-200); the interacting checks need their own exact
solver, written here from the definitions:

  H = sum_s sum_ij K_ij c+_is c_js + U sum_i (n_iu - 1/2)(n_id - 1/2)

with fermion modes m = s*N + i (s = 0 up, 1 down) and Jordan-Wigner signs.
``trotter_expect`` evaluates Tr[O (e^{-dtau H_U} e^{-dtau H_K})^L] / Tr[...],
which the discrete Hubbard-Stratonovich DQMC reproduces exactly at finite dtau;
``thermal_expect`` is the continuous-time value.
"""
import numpy as np
from scipy.linalg import expm


def _annihilators(n_modes):
    dim = 1 << n_modes
    ops = []
    states = np.arange(dim)
    for m in range(n_modes):
        c = np.zeros((dim, dim))
        occ = (states >> m) & 1
        lower = states & ((1 << m) - 1)
        sign = np.array([(-1) ** bin(x).count("1") for x in lower])
        src = states[occ == 1]
        c[src ^ (1 << m), src] = sign[occ == 1]
        ops.append(c)
    return ops


class Hubbard:
    def __init__(self, k, u):
        self.k = np.asarray(k, dtype=np.float64)
        self.n = self.k.shape[0]
        self.u = float(u)
        c = _annihilators(2 * self.n)
        self.c = c
        n = self.n
        self.num = [c[m].T @ c[m] for m in range(2 * n)]
        hk = np.zeros_like(c[0])
        for s in range(2):
            for i in range(n):
                for j in range(n):
                    if self.k[i, j] != 0.0:
                        hk += self.k[i, j] * (c[s * n + i].T @ c[s * n + j])
        self.hk = hk
        self.hu = sum(self.u * (self.num[i] - 0.5 * np.eye(len(hk))) @
                      (self.num[n + i] - 0.5 * np.eye(len(hk))) for i in range(n))
        hop = self.k - np.diag(np.diag(self.k))
        self.hop = np.zeros_like(hk)
        for s in range(2):
            for i in range(n):
                for j in range(n):
                    if hop[i, j] != 0.0:
                        self.hop += hop[i, j] * (c[s * n + i].T @ c[s * n + j])

    def operators(self):
        n = self.n
        docc = sum(self.num[i] @ self.num[n + i] for i in range(n)) / n
        dens = sum(self.num) / n
        return dict(docc=docc, ekin=self.hop / n, density=dens)

    def _expect(self, rho):
        z = np.trace(rho)
        return {name: float(np.trace(op @ rho) / z) for name, op in self.operators().items()}

    def trotter_expect(self, beta, L):
        dtau = beta / L
        t = np.diag(np.exp(-dtau * np.diag(self.hu))) @ expm(-dtau * self.hk)
        return self._expect(np.linalg.matrix_power(t, L))

    def thermal_expect(self, beta):
        w, v = np.linalg.eigh(self.hk + self.hu)
        rho = (v * np.exp(-beta * (w - w.min()))) @ v.T
        return self._expect(rho)


def free_g(k, beta):
    """U = 0 equal-time G = (I + e^{-beta K})^-1 via the eigenbasis of K."""
    w, v = np.linalg.eigh(k)
    return (v / (1.0 + np.exp(-beta * w))) @ v.T


def fock_one_body(c, a, modes):
    """Fock-space operator sum_ij a_ij c+_{modes[i]} c_{modes[j]} from annihilators c."""
    op = np.zeros_like(c[0])
    for i, mi in enumerate(modes):
        for j, mj in enumerate(modes):
            if a[i, j] != 0.0:
                op += a[i, j] * (c[mi].T @ c[mj])
    return op


def fock_product_expect(logs_up, logs_dn):
    """Many-body expectations for ONE auxiliary-field configuration, by brute force in
    Fock space.  The weight is B^ = prod_t e^{A^_t} (t in list order) for each spin, with
    A^_t = sum_ij (A_t)_ij c+_i c_j the one-body operator of the N x N matrix A_t.
    Returns (G_up, G_dn, S): G_s,ij = Tr[c_is c+_js B^]/Tr[B^] and
    S_ij = Tr[S^z_i S^z_j B^]/Tr[B^], S^z = n_up - n_dn.  Checks the Wick estimators of a
    non-symmetric G against their definition."""
    n = logs_up[0].shape[0]
    c = _annihilators(2 * n)
    b = np.eye(len(c[0]))
    for s, logs in enumerate((logs_up, logs_dn)):
        modes = [s * n + i for i in range(n)]
        for a in logs:
            b = b @ expm(fock_one_body(c, a, modes))
    z = np.trace(b)
    g = np.zeros((2, n, n))
    for s in range(2):
        for i in range(n):
            for j in range(n):
                g[s, i, j] = np.trace(c[s * n + i] @ c[s * n + j].T @ b) / z
    num = [c[m].T @ c[m] for m in range(2 * n)]
    sz = [num[i] - num[n + i] for i in range(n)]
    sc = np.array([[np.trace(sz[i] @ sz[j] @ b) / z for j in range(n)] for i in range(n)])
    return g[0], g[1], sc
