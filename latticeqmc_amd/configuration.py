"""Ising Hubbard-Stratonovich field s[i, l] in {-1, +1}.

Same layout and RNG semantics as the reference's ``lqmc.configuration.Configuration``
(lqmc/configuration.py:86-188): an int8 array of shape (n_sites, time_steps),
initialised by ``2*np.random.randint(0, 2, size=(N, L)) - 1`` from numpy's
global legacy RNG, so a seed set with ``np.random.seed`` gives the same field.
The device copy of the field is slice-major ([l][i]); the transposition happens
at upload.
"""
import numpy as np


class Configuration:
    dtype = np.int8

    def __init__(self, n_sites, time_steps, array=None):
        self.n_sites = n_sites
        self.time_steps = time_steps
        if array is not None:
            self.config = np.asarray(array, dtype=self.dtype)
        else:
            self.initialize()

    def copy(self):
        return Configuration(self.n_sites, self.time_steps, array=self.config.copy())

    def initialize(self, rng=None):
        """Draw an i.i.d. uniform +-1 field (configuration.py:120-124)."""
        src = np.random if rng is None else rng
        config = 2 * src.randint(0, 2, size=(self.n_sites, self.time_steps)) - 1
        self.config = config.astype(self.dtype)

    def update(self, i, t):
        self.config[i, t] *= -1

    def get(self, i, t):
        return self.config[i, t]

    def mean(self):
        return np.mean(self.config)

    def var(self):
        return np.var(self.config)

    def __eq__(self, other):
        return bool(np.all(self.config == other.config))

    def __getitem__(self, item):
        return self.config[item]

    def __str__(self):
        head = r"i\l  " + " ".join(f"{i:^3}" for i in range(self.time_steps))
        rows = [f"{s:<3} [" + " ".join(f"{x:^3}" for x in self.config[s, :]) + "]"
                for s in range(self.n_sites)]
        return head + "\n" + "\n".join(rows)
