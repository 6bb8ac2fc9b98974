"""Mode D: textbook DQMC sweep with QR/UDV stabilisation (test infrastructure only).

Same field layout, sweep order and RNG consumption as the reference
(lqmc/lqmc.py:309-317: l = L-1..0 outer, i = 0..N-1 inner, one uniform per
attempt, accept iff u <= ratio, flip after the G update) and the same ratio
formula (lqmc.py:313-316); the slice matrices, update and wrap are the
consistent ones:

  B_s(l) = V_s(l) E_K,  V_s(l) = diag(exp(-s*lambda*field[:, l]))   (lqmc.py:149)
  G(l)   = (I + B(l) B(l-1) ... B(0) B(L-1) ... B(l+1))^-1
  alpha_up = e^{a}-1, alpha_dn = e^{-a}-1, a = 2 lambda field[i,l]
  ratio  = prod_s [1 + alpha_s (1 - G_s,ii)]
  accept: G_s -= alpha_s/(1+alpha_s(1-G_s,ii)) G_s[:,i] (e_i - G_s[i,:])^T
  wrap   : G_s <- B_s(l)^-1 G_s B_s(l)   (E_K^-1 = expm(+dtau K))

Stabilisation schedule (identical in the HIP engine): with recompute points
l_r(m) = L-1-m*n_s, the sweep keeps a left stack S_m = B(l_r(m))...B(0) in UDV
form (built at sweep start) and a right product R = B(L-1)...B(l+1) built from
the updated slices.  After every n_s slices (and after slice 0) G is replaced
by the stable (I + S R)^-1 and max|G_wrapped - G_stable| is recorded; the unequal-time
G(tau, 0) = (S^-1 + R)^-1 at that point (tau = l dtau, l slices in S) is recorded too.

UDV: W = (B...B U) diag(D); W P = Q R (pivoted Householder QR);
D' = |diag R|, V' = (D'^-1 R) P^T V.  Right products are kept transposed,
R = V^T diag(D) U^T, and extended with B^T = E_K V.
Stable solve:  with D = Db*Ds (Db = max(D,1), Ds = min(D,1))
  X = Db_L^-1 (U_L^T U_R) Db_R^-1 + Ds_L (V_L V_R^T) Ds_R
  G = U_R Db_R^-1 X^-1 Db_L^-1 U_L^T

``delay = k > 1`` processes each slice in blocks of k sites with the submatrix
update: the serial accept/reject runs on the k x k diagonal block of G only,
then G -= (G[:, A] Gamma^-1) (E_A - G[A, :]) with
Gamma = diag(1/alpha_A) + (I - G)[A, A] (Woodbury form of the accepted flips).
"""
import numpy as np
from scipy.linalg import qr as _qr


def udv_qr(w):
    """Pivoted QR of w -> (Q, D, T) with w = Q diag(D) T."""
    q, r, p = _qr(w, pivoting=True)
    d = np.abs(np.diag(r))
    t = np.empty_like(r)
    t[:, p] = r / d[:, None]
    return q, d, t


def stable_g(left, right):
    """(I + U_L D_L V_L V_R^T D_R U_R^T)^-1 from a left and a (transposed) right UDV."""
    ul, dl, vl = left
    ur, dr, vr = right
    dlb, dls = np.maximum(dl, 1.0), np.minimum(dl, 1.0)
    drb, drs = np.maximum(dr, 1.0), np.minimum(dr, 1.0)
    x = (ul.T @ ur) / dlb[:, None] / drb[None, :] + dls[:, None] * (vl @ vr.T) * drs[None, :]
    rhs = ul.T / dlb[:, None]
    y = np.linalg.solve(x, rhs)
    return (ur / drb[None, :]) @ y


def stable_gtd(left, right):
    """Unequal-time G(tau, 0) = (S^-1 + R)^-1 = G(tau) S for the left product S = U_L D_L V_L
    (slices below tau) and the transposed right product R = V_R^T D_R U_R^T (slices above):
    factoring V_L^-1 out on the left and U_R^T on the right leaves
    D_L^-1 (U_L^T U_R) + (V_L V_R^T) D_R = Ds_L^-1 X Db_R with the X of stable_g, so
    G(tau, 0) = U_R Db_R^-1 X^-1 Ds_L V_L.  No counterpart in lqmc/lqmc.py (its only
    imaginary-time function is the free-fermion one of lqmc/tools.py:182-200)."""
    ul, dl, vl = left
    ur, dr, vr = right
    dlb, dls = np.maximum(dl, 1.0), np.minimum(dl, 1.0)
    drb, drs = np.maximum(dr, 1.0), np.minimum(dr, 1.0)
    x = (ul.T @ ur) / dlb[:, None] / drb[None, :] + dls[:, None] * (vl @ vr.T) * drs[None, :]
    y = np.linalg.solve(x, dls[:, None] * vl)
    return (ur / drb[None, :]) @ y


def identity_udv(n):
    return (np.eye(n), np.ones(n), np.eye(n))


class DQMCOracle:
    """One Markov chain of the Mode D sweep in FP64."""

    def __init__(self, k, beta, L, u, field, n_stab=10, delay=1, expk=None, expk_inv=None, dtype=np.float64):
        """dtype=np.float32 keeps G, the wraps and the flip updates in FP32 (E_K rounded to FP32),
        with the ratios, coefficients and every recompute in FP64 -- the FP32 form of the engine;
        ``errors`` then records max|G_FP32 - G_FP64| at each recompute."""
        from .model import params
        self.dtype = np.dtype(dtype)
        self.k = np.asarray(k, dtype=np.float64)
        self.n = self.k.shape[0]
        self.L = int(L)
        self.n_stab = int(n_stab)
        self.delay = int(delay)
        self.dtau, self.lam, ek, eki = params(self.k, beta, L, u)
        self.expk = ek if expk is None else expk
        self.expk_inv = eki if expk_inv is None else expk_inv
        self.field = np.array(field, dtype=np.int8)
        # exp(+-lambda), exp(+-2 lambda) as the engine uploads them
        self.e1 = {+1: np.exp(-self.lam), -1: np.exp(self.lam)}   # exp(-lambda*s) for s=+1/-1
        self.e2p, self.e2m = np.exp(2 * self.lam), np.exp(-2 * self.lam)
        self.nblk = -(-self.L // self.n_stab)
        self.errors = []
        self.g = [None, None]
        self.init_g()
        if self.dtype != np.float64:
            self.g = [g.astype(self.dtype) for g in self.g]
            self._ek_f = self.expk.astype(self.dtype)
            self._eki_f = self.expk_inv.astype(self.dtype)

    # -- slice matrices ------------------------------------------------------------------
    def v(self, l, sigma):
        s = self.field[:, l]
        # exp(-sigma*lambda*s): for sigma*s = +1 -> exp(-lambda), -1 -> exp(+lambda)
        return np.where(sigma * s > 0, self.e1[+1], self.e1[-1])

    def bmul_left(self, l, sigma, x):
        """B(l) @ x = V(l) (E_K x)."""
        return self.v(l, sigma)[:, None] * (self.expk @ x)

    def btmul_left(self, l, sigma, x):
        """B(l)^T @ x = E_K (V(l) x)."""
        return self.expk @ (self.v(l, sigma)[:, None] * x)

    # -- stabilisation ---------------------------------------------------------------------
    def block_range(self, m):
        """Slices (lo..hi) of left-stack block m (hi = l_r(m))."""
        hi = self.L - 1 - m * self.n_stab
        lo = max(self.L - (m + 1) * self.n_stab, 0)
        return lo, hi

    def left_stack(self, sigma, m_min):
        """UDV of S_m = B(l_r(m))...B(0) for m = nblk-1 down to m_min."""
        st = {}
        u, d, vv = identity_udv(self.n)
        for m in range(self.nblk - 1, m_min - 1, -1):
            lo, hi = self.block_range(m)
            w = u
            for l in range(lo, hi + 1):
                w = self.bmul_left(l, sigma, w)
            w = w * d[None, :]
            q, dn, t = udv_qr(w)
            u, d, vv = q, dn, t @ vv
            st[m] = (u, d, vv)
        return st

    def init_g(self):
        for s_i, sigma in enumerate((+1, -1)):
            st = self.left_stack(sigma, 0)
            self.g[s_i] = stable_g(st[0], identity_udv(self.n))

    # -- updates ---------------------------------------------------------------------------
    def ratio(self, i, l):
        s = self.field[i, l]
        a_up = (self.e2p if s > 0 else self.e2m) - 1.0
        a_dn = (self.e2m if s > 0 else self.e2p) - 1.0
        gu, gd = self.g
        d_up = 1.0 + (1.0 - gu[i, i]) * a_up
        d_dn = 1.0 + (1.0 - gd[i, i]) * a_dn
        return d_up * d_dn, (a_up, a_dn), (d_up, d_dn)

    def rank1(self, i, alphas, dets):
        for s_i in range(2):
            g = self.g[s_i]
            coef = alphas[s_i] / dets[s_i]
            x = g[:, i].astype(np.float64)
            y = -g[i, :].astype(np.float64)
            y[i] += 1.0
            self.g[s_i] = (g - coef * np.outer(x, y)).astype(self.dtype)

    def slice_rank1(self, l, uni, rec):
        for i in range(self.n):
            r, al, de = self.ratio(i, l)
            acc = uni[i] <= r
            rec.append((r, acc))
            if acc:
                self.rank1(i, al, de)
                self.field[i, l] *= -1

    def slice_delayed(self, l, uni, rec):
        k = self.delay
        for b0 in range(0, self.n, k):
            blk = np.arange(b0, min(b0 + k, self.n))
            small = [self.g[0][np.ix_(blk, blk)].astype(np.float64), self.g[1][np.ix_(blk, blk)].astype(np.float64)]
            acc_idx, acc_alpha = [], [[], []]
            for jj, i in enumerate(blk):
                s = self.field[i, l]
                a_up = (self.e2p if s > 0 else self.e2m) - 1.0
                a_dn = (self.e2m if s > 0 else self.e2p) - 1.0
                d_up = 1.0 + (1.0 - small[0][jj, jj]) * a_up
                d_dn = 1.0 + (1.0 - small[1][jj, jj]) * a_dn
                r = d_up * d_dn
                acc = uni[i] <= r
                rec.append((r, acc))
                if acc:
                    for s_i, (al, de) in enumerate(((a_up, d_up), (a_dn, d_dn))):
                        gs = small[s_i]
                        x = gs[:, jj].copy()
                        y = -gs[jj, :].copy()
                        y[jj] += 1.0
                        small[s_i] = gs - (al / de) * np.outer(x, y)
                        acc_alpha[s_i].append(al)
                    acc_idx.append(i)
                    self.field[i, l] *= -1
            if acc_idx:
                a = np.asarray(acc_idx)
                for s_i in range(2):
                    g = self.g[s_i]
                    gam = np.diag(1.0 / np.asarray(acc_alpha[s_i])) - g[np.ix_(a, a)].astype(np.float64)
                    gam[np.diag_indices(len(a))] += 1.0
                    if self.dtype == np.float64:
                        p = g[:, a] @ np.linalg.inv(gam)
                        z = -g[a, :]
                        z[np.arange(len(a)), a] += 1.0
                        self.g[s_i] = g - p @ z
                    else:  # FP64 coefficients, FP32 operands and accumulation of the rank-k product
                        p = (g[:, a].astype(np.float64) @ np.linalg.inv(gam)).astype(self.dtype)
                        z = -g[a, :]
                        z[np.arange(len(a)), a] += self.dtype.type(1.0)
                        self.g[s_i] = g - p @ z

    def wrap(self, l):
        for s_i, sigma in enumerate((+1, -1)):
            v = self.v(l, sigma)
            if self.dtype == np.float64:
                gs = self.g[s_i] * (v[None, :] / v[:, None])      # V^-1 G V
                self.g[s_i] = self.expk_inv @ gs @ self.expk
            else:
                gs = self.g[s_i] * (v[None, :] / v[:, None]).astype(self.dtype)
                self.g[s_i] = self._eki_f @ gs @ self._ek_f

    # -- sweep -------------------------------------------------------------------------------
    def sweep(self, uniforms, max_slices=None):
        """One sweep; returns dict(accept, ratio, g_meas, errors).  ``max_slices`` stops after
        that many slices (a bounded sample for timing; the chain is then mid-sweep)."""
        L, n = self.L, self.n
        uniforms = np.asarray(uniforms, dtype=np.float64).reshape(L, n)
        stacks = [self.left_stack(+1, 1), self.left_stack(-1, 1)]
        right = [identity_udv(n), identity_udv(n)]
        wacc = [None, None]
        rec = []
        g_meas = None
        errs = []
        gtd = []  # (l, [G_up(l dtau, 0), G_dn(l dtau, 0)]) at each stabilisation point
        for t in range(L):
            if max_slices is not None and t >= max_slices:
                break
            l = L - 1 - t
            if self.delay > 1:
                self.slice_delayed(l, uniforms[t], rec)
            else:
                self.slice_rank1(l, uniforms[t], rec)
            if l == 0:
                g_meas = (self.g[0].astype(np.float64), self.g[1].astype(np.float64))
            self.wrap(l)
            for s_i, sigma in enumerate((+1, -1)):
                if wacc[s_i] is None:
                    wacc[s_i] = self.btmul_left(l, sigma, right[s_i][0])
                else:
                    wacc[s_i] = self.btmul_left(l, sigma, wacc[s_i])
            if (t + 1) % self.n_stab == 0 or t == L - 1:
                m_next = (t + 1) // self.n_stab
                err = 0.0
                for s_i in range(2):
                    ur, dr, vr = right[s_i]
                    q, dn, tt = udv_qr(wacc[s_i] * dr[None, :])
                    right[s_i] = (q, dn, tt @ vr)
                    wacc[s_i] = None
                    left = identity_udv(n) if t == L - 1 else stacks[s_i][m_next]
                    gst = stable_g(left, right[s_i])
                    err = max(err, float(np.max(np.abs(self.g[s_i].astype(np.float64) - gst))))
                    self.g[s_i] = gst.astype(self.dtype)
                    if s_i == 0:
                        gtd.append((l, []))
                    gtd[-1][1].append(stable_gtd(left, right[s_i]))
                errs.append(err)
        self.errors.extend(errs)
        ratios = np.array([r for r, _ in rec])
        accs = np.array([a for _, a in rec], dtype=bool)
        return dict(accept=accs, ratio=ratios, g_meas=g_meas, errors=errs, gtd=gtd)

    def exact_g(self, l, sigma):
        """Direct (unstabilised) G(l) for small beta checks."""
        n = self.n
        p = np.eye(n)
        order = [(l - j) % self.L for j in range(self.L)]  # B(l) B(l-1) ... B(l+1)
        for ll in reversed(order):
            p = self.bmul_left(ll, sigma, p)
        return np.linalg.inv(np.eye(n) + p)


def observables(g_up, g_dn, hop):
    """Per-configuration equal-time estimators from G_ij = <c_i c_j^+>.

    density n = sum_s (1 - G_s,ii)/N, double occupancy D = mean_i (1-Gu_ii)(1-Gd_ii),
    kinetic energy per site Ekin = sum_s sum_ij hop_ij (delta_ij - G_s,ji) / N.
    """
    n = g_up.shape[0]
    nu = 1.0 - np.diag(g_up)
    nd = 1.0 - np.diag(g_dn)
    dens = float(np.sum(nu + nd) / n)
    docc = float(np.mean(nu * nd))
    ekin = 0.0
    for g in (g_up, g_dn):
        ekin += float(np.sum(hop * (np.eye(n) - g.T)))
    return dict(density=dens, docc=docc, ekin=ekin / n)


def spin_correlation(g_up, g_dn):
    """Equal-time <S^z_i S^z_j>, S^z = n_up - n_dn, for one field configuration by Wick's
    theorem from G_ij = <c_i c_j^+>:  m_i m_j + sum_s (delta_ij - G_s,ji) G_s,ij  with
    m_i = G_dn,ii - G_up,ii.  (The reference measures G only, lqmc.py:371-375; this is
    the estimator behind its local-moment helper, tools.py:87-109, taken per sweep.)"""
    n = g_up.shape[0]
    m = np.diag(g_dn) - np.diag(g_up)
    c = np.outer(m, m)
    for g in (g_up, g_dn):
        c += (np.eye(n) - g.T) * g
    return c
