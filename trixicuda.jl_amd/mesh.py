"""Uniform TreeMesh on the host (periodic or not).

Mirrors what the reference obtains from Trixi's TreeMesh and
init_elements!/init_interfaces! before copying to the GPU
(src/solvers/cache.jl:130-158, containers_3d.jl:34-96):

* refining every cell of a cube `initial_refinement_level` times appends the
  2^d children of each cell in order, child c having offset bit d set iff
  ((c - 1) >> d) & 1, so the leaf (element) order is the Morton order with
  the coarsest level most significant;
* one interface per element and positive direction (+x, +y, +z in that
  order), left = the element, right = its neighbour (periodic wrap-around in
  periodic directions); in a non-periodic direction the faces on the domain
  boundary get no interface and are the mesh's boundary faces;
* inverse Jacobian 2/dx, node coordinates centre + dx/2 * LGL node.
"""
import numpy as np


class TreeMesh:
    def __init__(self, coordinates_min, coordinates_max, initial_refinement_level, n_cells_max=None,
                 periodicity=True):
        cmin = np.atleast_1d(np.asarray(coordinates_min, dtype=np.float64))
        cmax = np.atleast_1d(np.asarray(coordinates_max, dtype=np.float64))
        if cmin.shape != cmax.shape or cmin.ndim != 1 or not 1 <= cmin.size <= 3:
            raise ValueError("coordinates must be 1, 2 or 3 dimensional")
        lengths = cmax - cmin
        if not np.all(lengths == lengths[0]):
            raise ValueError("TreeMesh domains must be cubes")
        self.ndims = nd = cmin.size
        per = (bool(periodicity),) * nd if isinstance(periodicity, (bool, np.bool_)) else tuple(map(bool, periodicity))
        if len(per) != nd:
            raise ValueError("periodicity needs one entry per dimension")
        self.periodicity = per
        self.level = L = int(initial_refinement_level)
        self.coordinates_min, self.coordinates_max = cmin, cmax
        self.length = float(lengths[0])
        self.nelements = E = 1 << (nd * L)
        if n_cells_max is not None and E > n_cells_max:
            raise ValueError(f"n_cells_max={n_cells_max} too small for {E} leaf cells")
        e = np.arange(E, dtype=np.int64)
        n1 = 1 << L
        cells = np.zeros((E, nd), dtype=np.int64)
        for lvl in range(L):
            for d in range(nd):
                cells[:, d] |= ((e >> (lvl * nd + d)) & 1) << lvl
        self.cell_index = cells.astype(np.int32)

        def encode(k):
            code = np.zeros(len(k), dtype=np.int64)
            for lvl in range(L):
                for d in range(nd):
                    code |= ((k[:, d] >> lvl) & 1) << (lvl * nd + d)
            return code

        nbr = np.zeros((E, nd, 2), dtype=np.int32)
        ori = np.zeros((E, nd), dtype=np.int32)
        keep = np.ones((E, nd), dtype=bool)
        for d in range(nd):
            k = cells.copy()
            k[:, d] = (k[:, d] + 1) % n1
            nbr[:, d, 0] = e
            nbr[:, d, 1] = encode(k)
            ori[:, d] = d + 1
            if not per[d]:
                keep[:, d] = cells[:, d] + 1 < n1
        self.neighbor_ids = nbr[keep].reshape(-1, 2)
        self.orientations = ori[keep].reshape(-1)
        self.ninterfaces = int(keep.sum())
        # boundary faces (element, direction 0..2d-1 as -x, +x, -y, +y, -z, +z)
        bnd = []
        for d in range(nd):
            if not per[d]:
                for side, at in ((0, 0), (1, n1 - 1)):
                    for el in np.nonzero(cells[:, d] == at)[0]:
                        bnd.append((int(el), 2 * d + side))
        self.boundary_faces = np.array(sorted(bnd), dtype=np.int32).reshape(-1, 2)
        dx = self.length / n1
        self.inverse_jacobian = np.full(E, 1.0 / (dx / 2.0))
        # Cell centres by halving from the root centre (the tree's own
        # arithmetic, so the coordinates are exact binary fractions).
        centers = np.empty((E, nd))
        for d in range(nd):
            cc = np.full(E, 0.5 * (cmin[d] + cmax[d]))
            ln = self.length
            for lvl in range(L - 1, -1, -1):
                bit = (cells[:, d] >> lvl) & 1
                ln *= 0.5
                cc = cc + np.where(bit == 1, 0.5, -0.5) * ln
            centers[:, d] = cc
        self.centers = centers
        self.total_volume = self.length ** nd

    def node_coordinates(self, nodes):
        """Node coordinates [E, (N,)*nd, nd] in Trixi order (i fastest = last axis)."""
        nd, N = self.ndims, len(nodes)
        jac = (self.length / (1 << self.level)) / 2.0
        shape = (self.nelements,) + (N,) * nd + (nd,)
        x = np.empty(shape)
        nodes = np.asarray(nodes, dtype=np.float64)
        for d in range(nd):
            # node axes are (k, j, i): direction d varies along node axis nd-1-d
            shp = [1] * nd
            shp[nd - 1 - d] = N
            x[..., d] = self.centers[:, d].reshape((-1,) + (1,) * nd) + (jac * nodes).reshape(shp)[None]
        return x

    def __repr__(self):
        return (f"TreeMesh{self.ndims}D(level={self.level}, nelements={self.nelements}, "
                f"domain=[{self.coordinates_min}, {self.coordinates_max}], periodicity={self.periodicity})")
