"""TreeMesh on the host: uniform or with refinement patches, periodic or not.

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

With refinement patches the tree is built as Trixi's TreeMesh does it
(abstract_tree.jl refine!, refine_box!, rebalance!): refining a cell inserts
its 2^d children directly behind it (so the element order is the depth-first
order of the leaves); "box" patches refine the leaves whose centre lies
strictly inside the box; after each refinement the tree is rebalanced to 2:1
across faces.  Interfaces join same-level leaves (+ directions), mortars join
a leaf to the 2^(d-1) children of its same-level neighbour (L2 mortars, the
container of Trixi's init_mortars!: small elements in position order, first
tangential direction fastest, then the large element), and faces without any
neighbour are boundary faces.
"""
import numpy as np


class TreeMesh:
    def __init__(self, coordinates_min, coordinates_max, initial_refinement_level, n_cells_max=None,
                 periodicity=True, refinement_patches=()):
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
        self.nmortars = 0
        self.mortar_neighbor_ids = np.zeros((0, (1 << (nd - 1)) + 1), dtype=np.int32)
        self.mortar_orientations = np.zeros(0, dtype=np.int32)
        self.mortar_large_sides = np.zeros(0, dtype=np.int32)
        self.coordinates_min, self.coordinates_max = cmin, cmax
        self.length = float(lengths[0])
        if refinement_patches:
            if nd == 1:
                raise NotImplementedError("refinement patches are built for 2D and 3D meshes (L2 mortars)")
            self._build_tree(cmin, cmax, int(initial_refinement_level), per, refinement_patches, n_cells_max)
            return
        self.level = L = int(initial_refinement_level)
        self._tree = None
        self._n_cells_max = n_cells_max
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

    # -- general tree (refinement patches) -------------------------------------
    def _build_tree(self, cmin, cmax, level0, per, patches, n_cells_max):
        nd = self.ndims
        nch = 1 << nd
        length = float(cmax[0] - cmin[0])
        # cell: [level, integer coords at that level, children ids or None]
        cells = [[0, (0,) * nd, None]]
        order = [0]                       # storage order (Trixi's cell array)
        index = {(0, (0,) * nd): 0}       # (level, coords) -> cell id

        def same_level_neighbour(cid, d, side):
            lvl, k, _ = cells[cid]
            n1 = 1 << lvl
            kk = list(k)
            kk[d] += 1 if side else -1
            if not 0 <= kk[d] < n1:
                if not per[d]:
                    return None
                kk[d] %= n1
            return index.get((lvl, tuple(kk)))

        def refine_unbalanced(ids):
            pos = {c: i for i, c in enumerate(order)}
            new_order, done = [], []
            todo = set(ids)
            for c in order:
                new_order.append(c)
                if c in todo:
                    lvl, k, ch = cells[c]
                    assert ch is None, "only leaves are refined"
                    kids = []
                    for child in range(nch):
                        kc = tuple(2 * k[d] + ((child >> d) & 1) for d in range(nd))
                        cells.append([lvl + 1, kc, None])
                        cid = len(cells) - 1
                        index[(lvl + 1, kc)] = cid
                        kids.append(cid)
                    cells[c][2] = kids
                    new_order.extend(kids)
                    done.append(c)
            order[:] = new_order
            del pos
            return done

        def rebalance(refined):
            need = []
            for c in refined:
                lvl, k, _ = cells[c]
                if lvl == 0:
                    continue
                parent = index[(lvl - 1, tuple(x >> 1 for x in k))]
                for d in range(nd):
                    for side in (0, 1):
                        if same_level_neighbour(c, d, side) is not None:
                            continue
                        coarse = same_level_neighbour(parent, d, side)
                        if coarse is None:
                            continue  # domain boundary
                        if cells[coarse][2] is None:
                            need.append(coarse)
            return refine_unbalanced(sorted(set(need)))

        def refine(ids):
            refined = refine_unbalanced(ids)
            while refined:
                refined = rebalance(refined)

        def leaves():
            return [c for c in order if cells[c][2] is None]

        def centre(c):
            lvl, k, _ = cells[c]
            # halving from the root centre, as the tree computes child centres
            out = np.empty(nd)
            for d in range(nd):
                cc, ln = 0.5 * (cmin[d] + cmax[d]), length
                for b in range(lvl - 1, -1, -1):
                    ln *= 0.5
                    cc += (0.5 if (k[d] >> b) & 1 else -0.5) * ln
                out[d] = cc
            return out

        for _ in range(level0):
            refine([c for c in leaves() if cells[c][0] < level0])
        for patch in patches:
            ptype = patch["type"] if isinstance(patch, dict) else patch.type
            if ptype != "box":
                raise NotImplementedError("only box refinement patches are built")
            lo = np.asarray(patch["coordinates_min"] if isinstance(patch, dict) else patch.coordinates_min, float)
            hi = np.asarray(patch["coordinates_max"] if isinstance(patch, dict) else patch.coordinates_max, float)
            refine([c for c in leaves() if np.all(lo < centre(c)) and np.all(hi > centre(c))])

        elems = leaves()
        E = len(elems)
        self._tree = (cells, list(order), index, per)
        self._n_cells_max = n_cells_max
        if n_cells_max is not None and len(cells) > n_cells_max:
            raise ValueError(f"n_cells_max={n_cells_max} too small for {len(cells)} cells")
        c2e = {c: i for i, c in enumerate(elems)}
        self.nelements = E
        self.level = max(cells[c][0] for c in elems)
        self.levels = np.array([cells[c][0] for c in elems], dtype=np.int32)
        self.cell_index = np.array([cells[c][1] for c in elems], dtype=np.int32).reshape(E, nd)
        self.centers = np.array([centre(c) for c in elems]).reshape(E, nd)
        self.inverse_jacobian = np.array([1.0 / ((length / (1 << cells[c][0])) / 2.0) for c in elems])
        ifc, mort, bnd = [], [], []
        for e, c in enumerate(elems):
            for d in range(nd):
                for side in (0, 1):
                    nb = same_level_neighbour(c, d, side)
                    if nb is None:
                        lvl, k, _ = cells[c]
                        if lvl > 0:
                            parent = index[(lvl - 1, tuple(x >> 1 for x in k))]
                            if same_level_neighbour(parent, d, side) is not None:
                                continue  # small side of a mortar: owned by the large element
                        bnd.append((e, 2 * d + side))
                        continue
                    if cells[nb][2] is None:
                        if side == 1:
                            ifc.append((e, c2e[nb], d + 1))
                        continue
                    # mortar: the neighbour's children on the face towards this element
                    tdims = [q for q in range(nd) if q != d]
                    small = []
                    for pos in range(1 << (nd - 1)):
                        child = 0 if side else (1 << d)
                        for t, q in enumerate(tdims):
                            child |= ((pos >> t) & 1) << q
                        small.append(c2e[cells[nb][2][child]])
                    mort.append((small + [e], d + 1, 1 if side else 2))
        self.neighbor_ids = np.array([(a, b) for a, b, _ in ifc], dtype=np.int32).reshape(-1, 2)
        self.orientations = np.array([o for _, _, o in ifc], dtype=np.int32)
        self.ninterfaces = len(ifc)
        self.boundary_faces = np.array(sorted(bnd), dtype=np.int32).reshape(-1, 2)
        self.nmortars = len(mort)
        self.mortar_neighbor_ids = np.array([m[0] for m in mort], dtype=np.int32).reshape(-1, (1 << (nd - 1)) + 1)
        self.mortar_orientations = np.array([m[1] for m in mort], dtype=np.int32)
        self.mortar_large_sides = np.array([m[2] for m in mort], dtype=np.int32)
        self.total_volume = length ** nd

    def tree_arrays(self):
        """The cell tree as Trixi's TreeMesh stores it and its mesh file holds
        (abstract_tree.jl / serial_tree.jl; save_mesh_file of a TreeMesh):
        cells in storage order (depth first: refining a cell inserts its
        children directly behind it), 1-based cell ids, 0 for none.

        Returns a dict of parent_ids [n], child_ids [n, 2^d] (child c has
        offset bit d set iff (c >> d) & 1), neighbor_ids [n, 2d] (directions
        -x, +x, -y, +y, -z, +z: the same-level cell across the face, wrapping
        around periodic directions, so a periodic root is its own neighbour),
        levels [n], coordinates [n, d] (cell centres, halved from the root
        centre) and the scalars n_cells, capacity, n_leaf_cells,
        minimum_level, maximum_level (over the leaves), center_level_0,
        length_level_0, periodicity.  The leaves in storage order are the
        elements in order."""
        nd = self.ndims
        nch = 1 << nd
        per = self.periodicity
        if self._tree is not None:
            cells, order, index, _ = self._tree
        else:
            # a uniform tree: every cell refined down to self.level, depth first
            cells, order, index = [], [], {}

            def add(lvl, k):
                cid = len(cells)
                cells.append([lvl, k, None])
                index[(lvl, k)] = cid
                order.append(cid)
                if lvl < self.level:
                    kids = []
                    for child in range(nch):
                        kc = tuple(2 * k[d] + ((child >> d) & 1) for d in range(nd))
                        kids.append(add(lvl + 1, kc))
                    cells[cid][2] = kids
                return cid

            add(0, (0,) * nd)
        pos = {c: i for i, c in enumerate(order)}  # storage position, 0-based
        n = len(order)
        parent = np.zeros(n, dtype=np.int64)
        child_ids = np.zeros((n, nch), dtype=np.int64)
        nbr = np.zeros((n, 2 * nd), dtype=np.int64)
        levels = np.zeros(n, dtype=np.int64)
        coords = np.zeros((n, nd))
        centre0 = 0.5 * (self.coordinates_min + self.coordinates_max)
        for c in order:
            i = pos[c]
            lvl, k, kids = cells[c]
            levels[i] = lvl
            if kids is not None:
                for j, kc in enumerate(kids):
                    child_ids[i, j] = pos[kc] + 1
                    parent[pos[kc]] = i + 1
            n1 = 1 << lvl
            for d in range(nd):
                # centre by halving from the root centre (exact binary fractions)
                cc, ln = centre0[d], self.length
                for b in range(lvl - 1, -1, -1):
                    ln *= 0.5
                    cc += (0.5 if (k[d] >> b) & 1 else -0.5) * ln
                coords[i, d] = cc
                for side in (0, 1):
                    kk = list(k)
                    kk[d] += 1 if side else -1
                    if not 0 <= kk[d] < n1:
                        if not per[d]:
                            continue
                        kk[d] %= n1
                    nb = index.get((lvl, tuple(kk)))
                    if nb is not None:
                        nbr[i, 2 * d + side] = pos[nb] + 1
        leaf = np.all(child_ids == 0, axis=1)
        return dict(parent_ids=parent, child_ids=child_ids, neighbor_ids=nbr, levels=levels, coordinates=coords,
                    n_cells=n, capacity=int(self._n_cells_max) if self._n_cells_max is not None else n,
                    n_leaf_cells=int(leaf.sum()), minimum_level=int(levels[leaf].min()),
                    maximum_level=int(levels[leaf].max()), center_level_0=centre0.copy(),
                    length_level_0=self.length, periodicity=tuple(per))

    def node_coordinates(self, nodes):
        """Node coordinates [E, (N,)*nd, nd] in Trixi order (i fastest = last axis)."""
        nd, N = self.ndims, len(nodes)
        if hasattr(self, "levels"):
            jac = ((self.length / (1 << self.levels)) / 2.0).reshape((-1,) + (1,) * nd)
        else:
            jac = (self.length / (1 << self.level)) / 2.0
        shape = (self.nelements,) + (N,) * nd + (nd,)
        x = np.empty(shape)
        nodes = np.asarray(nodes, dtype=np.float64)
        for d in range(nd):
            # node axes are (k, j, i): direction d varies along node axis nd-1-d
            shp = [1] * nd
            shp[nd - 1 - d] = N
            x[..., d] = self.centers[:, d].reshape((-1,) + (1,) * nd) + jac * nodes.reshape(shp)[None]
        return x

    def __repr__(self):
        return (f"TreeMesh{self.ndims}D(level={self.level}, nelements={self.nelements}, "
                f"domain=[{self.coordinates_min}, {self.coordinates_max}], periodicity={self.periodicity})")
