"""Bravais-lattice geometry that feeds the kinetic matrix K of the Hubbard model.

Mirrors the call surface of the reference's ``lqmc.lattice.Lattice`` for the
parts the DQMC sweep uses (``square``/``chain`` prefabs, ``build``,
``set_periodic_boundary``, ``nearest_neighbours``, ``n_sites``, ``shape``,
``indices``), written from scratch for one-atom hypercubic cells.

Site order (lattice.py:226-242, ``vrange``): for a shape (w, h) site
``i = x*h + y`` -- the first axis is slowest.  Periodic bonds
(lattice.py:994-1033) join sites ``side-1`` apart along the chosen axis.  For an
axis of extent 2 those pairs are already nearest neighbours, so the 2x2
periodic square is a 4-site ring with single bonds, as in the reference.

Divergence: for sides whose ``side-1`` is the hypotenuse of an integer
triangle (6, 11, 14, 16, ...) the reference either raises (numpy >= 2.2) or adds
long-range bonds at e.g. (dx, dy) = (3, 4); this module builds the standard
nearest-neighbour periodic lattice there.
"""
import numpy as np


class Lattice:
    """One-atom hypercubic lattice with nearest-neighbour bonds."""

    def __init__(self, vectors):
        self.vectors = np.asarray(vectors, dtype=np.float64)
        self.dim = len(self.vectors)
        self.n_sites = 0
        self.shape = None
        self.indices = None
        self.neighbours = None
        self.distances = [1.0]

    @classmethod
    def square(cls, a=1.0, shape=None, **_):
        """Square-lattice prefab (lattice.py:337-362)."""
        latt = cls(np.eye(2) * a)
        if shape is not None:
            latt.build(shape)
        return latt

    @classmethod
    def chain(cls, a=1.0, size=None, **_):
        """Chain prefab (lattice.py:310-335): a square lattice built as (size, 1)."""
        latt = cls(np.eye(2) * a)
        if size is not None:
            latt.build((size, 1))
        return latt

    @property
    def n_dist(self):
        return len(self.distances)

    def build(self, shape):
        """Build the open lattice of the given shape (lattice.py:894-904)."""
        shape = np.atleast_1d(np.asarray(shape, dtype=np.int64))
        if shape.size != self.dim:
            shape = np.concatenate([shape, np.ones(self.dim - shape.size, dtype=np.int64)])
        grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
        coords = np.stack([g.reshape(-1) for g in grids], axis=1)  # first axis slowest
        self.shape = shape
        self.n_sites = int(coords.shape[0])
        self.indices = np.concatenate([coords, np.zeros((self.n_sites, 1), np.int64)], axis=1)
        self._coords = coords
        nbrs = [[] for _ in range(self.n_sites)]
        index_of = {tuple(c): i for i, c in enumerate(coords.tolist())}
        for i, c in enumerate(coords.tolist()):
            for ax in range(self.dim):
                for step in (-1, 1):
                    d = list(c)
                    d[ax] += step
                    j = index_of.get(tuple(d))
                    if j is not None:
                        nbrs[i].append(j)
        self.neighbours = [[sorted(n)] for n in nbrs]
        return self.indices, self.shape

    def set_periodic_boundary(self, axis=0):
        """Join the two ends of each given axis (lattice.py:994-1033)."""
        for ax in np.atleast_1d(axis):
            ax = int(ax)
            extent = int(self.shape[ax])
            if extent < 2:
                continue
            for i in range(self.n_sites):
                c = self._coords[i]
                if c[ax] != 0:
                    continue
                d = c.copy()
                d[ax] = extent - 1
                j = int(np.ravel_multi_index(tuple(d), tuple(self.shape)))
                for a, b in ((i, j), (j, i)):
                    if b not in self.neighbours[a][0]:
                        self.neighbours[a][0].append(b)
                        self.neighbours[a][0].sort()

    def nearest_neighbours(self, i):
        return self.neighbours[i][0]

    def position(self, i):
        return self.vectors.T @ self._coords[i].astype(np.float64)

    def adjacency(self):
        """Dense symmetric 0/1 nearest-neighbour matrix."""
        a = np.zeros((self.n_sites, self.n_sites))
        for i in range(self.n_sites):
            for j in self.neighbours[i][0]:
                a[i, j] = 1.0
        return a
