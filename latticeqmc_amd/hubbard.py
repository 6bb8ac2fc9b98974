"""Single-band Hubbard model parameters and kinetic matrix.

Same call surface as the reference's ``lqmc.hubbard.HubbardModel``
(lqmc/hubbard.py:12-111): ``HubbardModel(u=4, t=1, mu=None)`` with ``mu``
defaulting to ``u/2``, ``build``, ``build_square``, ``set_params``,
``ham_kinetic`` and ``n_sites``.  ``ham_kinetic`` returns the dense
K = -mu*I - t*A used by every slice matrix B_l = exp(V_l) exp(-dtau K).
"""
import numpy as np

from .lattice import Lattice


class HubbardModel:
    def __init__(self, u=4, t=1, mu=None):
        self.lattice = Lattice.square()
        self.u = u
        self.t = t
        self.mu = mu if mu is not None else u / 2

    def set_params(self, u, t=1.0, mu=None):
        self.u = u
        self.t = t
        self.mu = mu if mu is not None else u / 2

    def param_str(self):
        return f"u={self.u}_t={self.t}_mu={self.mu}"

    def __str__(self):
        return f"HubbardModel(u={self.u}, t={self.t}, mu={self.mu})"

    @property
    def n_sites(self):
        return self.lattice.n_sites

    def build(self, width, height=1, cycling=0):
        """Build a (width, height) lattice; ``cycling`` names the periodic axes (hubbard.py:60-76)."""
        self.lattice.build((width, height))
        if cycling is not None:
            self.lattice.set_periodic_boundary(cycling)

    def build_square(self, size, cycling=(0, 1)):
        self.build(size, size, cycling)

    def ham_kinetic(self):
        """K = -mu*I - t*A, A the nearest-neighbour adjacency (hubbard.py:91-111)."""
        n = self.lattice.n_sites
        ham = -self.mu * np.eye(n, dtype=np.float64)
        ham -= self.t * self.lattice.adjacency()
        return ham
