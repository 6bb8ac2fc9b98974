"""Bravais lattices with a basis: the geometry that feeds the kinetic matrix K.

Mirrors the call surface of the reference's ``lqmc.lattice.Lattice``
(lattice.py:283-1033) for what the Hubbard model uses: the ``chain``, ``square``,
``hexagonal`` and ``cubic`` prefabs, ``add_atom``, ``calculate_distances``,
``build``, ``set_periodic_boundary``, ``nearest_neighbours``, ``n_sites``,
``shape``, ``indices``, ``distances``, ``position``.  Plotting is not provided.

Geometry, as in the reference:
  * site position r = V n + a_alpha with the primitive vectors as the COLUMNS of V
    (lattice.py:204-223) and a_alpha the basis positions;
  * site order (``vrange``, lattice.py:226-242): the cell vectors of a meshgrid in
    'xy' indexing, Fortran-flattened and stably sorted by the first coordinate, the
    basis index fastest -- for 2D shapes (w, h) site i = x*h + y; in 3D y runs
    fastest, then z, then x;
  * neighbour shells: the ``n`` smallest distinct distances (rounded to 5 decimals)
    from the basis atoms of the origin cell to the sites of the cells in
    [-(n+1), n] along each axis (lattice.py:694-725); the neighbours of a site at shell
    d are the sites at exactly that rounded distance.

Periodic boundaries (lattice.py:994-1033): sites whose cell indices along the axis are
0 and extent-1 and whose displacement is parallel to that axis are joined.  For the
square and cubic lattices this is the reference's rule wherever the reference
completes.  Where ``extent - 1`` is also the length of a skewed displacement (square
sides 6, 11, 14, 16, ...; cubic 4x4x4; every skewed-cell lattice such as the
hexagonal one) the reference raises under numpy >= 2.2 (its axis test meets an empty
or multi-element array).  This module then keeps the standard axis-parallel pairing
for orthogonal cells and raises ``ValueError`` for skewed cells, whose periodic
boundary the reference does not define.
"""
import numpy as np

DIST_DECIMALS = 5


def vrange(axis_ranges):
    """Cell vectors of a box in the reference's order (lattice.py:226-242)."""
    axis = np.meshgrid(*[np.asarray(r) for r in axis_ranges])
    grid = np.stack([np.asarray(a).ravel(order="F") for a in axis], axis=1)
    return grid[np.argsort(grid[:, 0], kind="stable")]


class Lattice:
    """Bravais lattice (primitive vectors as matrix columns) with a basis of atoms."""

    def __init__(self, vectors):
        self.vectors = np.asarray(vectors, dtype=np.float64)
        self.dim = len(self.vectors)
        self.atoms = []
        self.atom_positions = []
        self.distances = []
        self.n_sites = 0
        self.shape = None
        self.indices = None
        self.neighbours = None
        self._pos = None

    # ---- prefabs (lattice.py:310-421) -----------------------------------------------------
    @classmethod
    def chain(cls, name="A", a=1.0, neighbour_dist=1, size=None):
        latt = cls(np.eye(2) * a)
        latt.add_atom(name)
        latt.calculate_distances(neighbour_dist)
        if size is not None:
            latt.build((size, 1))
        return latt

    @classmethod
    def square(cls, name="A", a=1.0, neighbour_dist=1, shape=None):
        latt = cls(np.eye(2) * a)
        latt.add_atom(name)
        latt.calculate_distances(neighbour_dist)
        if shape is not None:
            latt.build(shape)
        return latt

    @classmethod
    def hexagonal(cls, atom1="A", atom2="B", a=1.0, neighbour_dist=1, shape=None):
        latt = cls(a * np.array([[np.sqrt(3), np.sqrt(3) / 2], [0.0, 1.5]]))
        latt.add_atom(atom1)
        latt.add_atom(atom2, pos=[0.0, a])
        latt.calculate_distances(neighbour_dist)
        if shape is not None:
            latt.build(shape)
        return latt

    @classmethod
    def cubic(cls, name="A", a=1.0, neighbour_dist=1, shape=None):
        latt = cls(np.eye(3) * a)
        latt.add_atom(name)
        latt.calculate_distances(neighbour_dist)
        if shape is not None:
            latt.build(shape)
        return latt

    # ---- cell ----------------------------------------------------------------------------
    @property
    def n_base(self):
        return len(self.atoms)

    @property
    def n_dist(self):
        return len(self.distances)

    def add_atom(self, name="A", pos=None, neighbour_dist=0):
        """Add a basis atom at ``pos`` (default the cell origin) (lattice.py:469-496)."""
        pos = np.zeros(self.dim) if pos is None else np.asarray(pos, dtype=np.float64)
        if any(np.all(pos == p) for p in self.atom_positions):
            raise ValueError(f"Position {pos} allready occupied")
        self.atoms.append(name)
        self.atom_positions.append(pos)
        if neighbour_dist:
            self.calculate_distances(neighbour_dist)

    def get_position(self, n, alpha=0):
        return self.atom_positions[alpha] + self.vectors @ np.asarray(n, dtype=np.float64)

    def calculate_distances(self, n=1):
        """The ``n`` smallest neighbour distances of the structure (lattice.py:694-725)."""
        m = n + 1
        cells = vrange(self.dim * [np.arange(-m, m)])
        pts = np.array([self.get_position(c, a) for c in cells for a in range(self.n_base)])
        origin = np.array(self.atom_positions)
        d = np.sqrt(((pts[:, None, :] - origin[None, :, :]) ** 2).sum(axis=2))
        uniq = np.unique(np.round(d, DIST_DECIMALS))
        self.distances = [float(x) for x in uniq if x != 0.0][:n]

    # ---- finite lattice ------------------------------------------------------------------
    def build(self, shape):
        """Sites and neighbour shells of the open lattice of the given cell shape
        (lattice.py:880-904)."""
        if not self.atoms:
            self.add_atom()
        if not self.distances:
            self.calculate_distances(1)
        shape = np.atleast_1d(np.asarray(shape, dtype=np.int64))
        if shape.size < self.dim:
            shape = np.concatenate([shape, np.ones(self.dim - shape.size, dtype=np.int64)])
        cells = vrange([range(s) for s in shape])
        idx = np.array([[*c, a] for c in cells for a in range(self.n_base)], dtype=np.int64)
        pos = np.array([self.get_position(r[:-1], r[-1]) for r in idx])
        self.shape = shape
        self.indices = idx
        self.n_sites = len(idx)
        self._pos = pos
        d = np.round(np.sqrt(((pos[:, None, :] - pos[None, :, :]) ** 2).sum(axis=2)), DIST_DECIMALS)
        self.neighbours = [[list(np.nonzero(d[i] == dist)[0]) for dist in self.distances]
                           for i in range(self.n_sites)]
        return self.indices, self.shape

    def set_periodic_boundary(self, axis=0):
        """Join the two ends of each given axis (lattice.py:994-1033; see the module notes)."""
        if self.indices is None:
            raise ValueError("build the lattice before setting periodic boundaries")
        v = self.vectors
        orthogonal = np.allclose(v - np.diag(np.diag(v)), 0.0) and all(
            np.allclose(a, 0.0) for a in self.atom_positions)
        for ax in np.atleast_1d(axis):
            ax = int(ax)
            extent = int(self.shape[ax])
            if extent < 2:
                continue
            if not orthogonal:
                raise ValueError("periodic boundaries are defined for orthogonal one-atom cells only")
            first = np.nonzero(self.indices[:, ax] == 0)[0]
            for i in first:
                key = self.indices[i].copy()
                key[ax] = extent - 1
                j = int(np.nonzero((self.indices == key).all(axis=1))[0][0])
                for a, b in ((int(i), j), (j, int(i))):
                    if b not in self.neighbours[a][0]:
                        self.neighbours[a][0].append(b)
                        self.neighbours[a][0].sort()

    # ---- queries -------------------------------------------------------------------------
    def nearest_neighbours(self, i):
        return self.neighbours[i][0]

    def dist_neighbours(self, i, dist=1):
        return self.neighbours[i][dist - 1]

    def position(self, i):
        return self._pos[i]

    def positions(self):
        return np.array(self._pos)

    def adjacency(self):
        """Dense symmetric 0/1 nearest-neighbour matrix."""
        a = np.zeros((self.n_sites, self.n_sites))
        for i in range(self.n_sites):
            for j in self.neighbours[i][0]:
                a[i, j] = 1.0
        return a
