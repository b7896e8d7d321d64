"""Structured hexahedral/tetrahedral meshes: Block and Mesh.

Host-side, vectorised restatement of the reference's structured mesher:
``Block`` (src/mesh/block.jl:101-212) and ``Mesh(blocks...)``
(src/mesh/genmesh.jl:69-123) → ``split_block!`` (src/mesh/structured.jl:
384-459 for HEX8, 461-552 for HEX20/TET10) → ``synchronize!``
(src/mesh/mesh.jl:335-391).  Node ids follow creation order (k outer, j, i
inner); coordinates are the HEX8 map of the block corners rounded to 8
decimals (src/node.jl:58-60), which is what fixes the DOF numbering.
"""
import numpy as np

HEX8, HEX20, TET10 = "HEX8", "HEX20", "TET10"
NNODES = {HEX8: 8, HEX20: 20, TET10: 10}

_HEX_SIGNS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                       [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)

# facet node lists (0-based) per element shape, as in the shapes' facet_idxs
# (src/shape/solids3d.jl: HEX8 :470, HEX20 :582, TET10 :121)
FACETS = {
    HEX8: np.array([[0, 4, 7, 3], [1, 2, 6, 5], [0, 1, 5, 4], [2, 3, 7, 6], [0, 3, 2, 1], [4, 5, 6, 7]]),
    HEX20: np.array([[0, 4, 7, 3, 16, 15, 19, 11], [1, 2, 6, 5, 9, 18, 13, 17], [0, 1, 5, 4, 8, 17, 12, 16],
                     [2, 3, 7, 6, 10, 19, 14, 18], [0, 3, 2, 1, 11, 10, 9, 8], [4, 5, 6, 7, 12, 13, 14, 15]]),
    TET10: np.array([[0, 3, 2, 7, 9, 6], [0, 1, 3, 4, 8, 7], [0, 2, 1, 6, 5, 4], [1, 2, 3, 5, 9, 8]]),
}
# edges of the facet shapes (QUAD4, QUAD8, TRI6 facet_idxs)
FACET_EDGES = {4: np.array([[0, 1], [1, 2], [2, 3], [3, 0]]),
               8: np.array([[0, 1, 4], [1, 2, 5], [2, 3, 6], [3, 0, 7]]),
               6: np.array([[0, 1, 3], [1, 2, 4], [2, 0, 5]])}
# six tetrahedra per hex cell, in the reference's order and node lists
# (src/mesh/structured.jl:537-542), as 1-based indices of the 27 cell points
_TET_SPLIT = [(2, 4, 1, 8, 25, 12, 9, 27, 20, 21), (2, 1, 5, 8, 9, 17, 23, 27, 21, 16),
              (2, 5, 6, 8, 23, 13, 18, 27, 16, 26), (2, 6, 7, 8, 18, 14, 22, 27, 26, 15),
              (2, 3, 4, 8, 10, 11, 25, 27, 24, 20), (2, 7, 3, 8, 22, 19, 10, 27, 15, 24)]
# grid offsets (di, dj, dk) of the 27 points of a 2x2x2 sub-grid cell, 1-based
# numbering p1..p27 of src/mesh/structured.jl:499-534
_P27 = {1: (0, 0, 0), 2: (2, 0, 0), 3: (2, 2, 0), 4: (0, 2, 0), 5: (0, 0, 2), 6: (2, 0, 2), 7: (2, 2, 2),
        8: (0, 2, 2), 9: (1, 0, 0), 10: (2, 1, 0), 11: (1, 2, 0), 12: (0, 1, 0), 13: (1, 0, 2), 14: (2, 1, 2),
        15: (1, 2, 2), 16: (0, 1, 2), 17: (0, 0, 1), 18: (2, 0, 1), 19: (2, 2, 1), 20: (0, 2, 1),
        21: (0, 1, 1), 22: (2, 1, 1), 23: (1, 0, 1), 24: (1, 2, 1), 25: (1, 1, 0), 26: (1, 1, 2), 27: (1, 1, 1)}


def round8(v):
    """round(x, digits=8) + 0.0: ties to even on the scaled value, -0 dropped."""
    return np.rint(v * 1e8) / 1e8 + 0.0


class Block:
    """Block(coords; nx, ny, nz, cellshape, tag) for 3D boxes
    (src/mesh/block.jl:115-211).  coords: 2×3 (opposite corners) or 8×3
    (HEX8-ordered corners)."""

    def __init__(self, coords, nx=1, ny=1, nz=1, cellshape=HEX8, tag=""):
        c = np.asarray(coords, dtype=float)
        if c.shape == (2, 3):
            (x1, y1, z1), (x2, y2, z2) = c
            lx, ly, lz = x2 - x1, y2 - y1, z2 - z1
            c = np.array([[x1, y1, z1], [x1 + lx, y1, z1], [x1 + lx, y1 + ly, z1], [x1, y1 + ly, z1],
                          [x1, y1, z1 + lz], [x1 + lx, y1, z1 + lz], [x1 + lx, y1 + ly, z1 + lz],
                          [x1, y1 + ly, z1 + lz]])
        if c.shape != (8, 3):
            raise ValueError("Block: expected 2x3 or 8x3 coordinates")
        if cellshape not in NNODES:
            raise ValueError(f"Block: unsupported cell shape {cellshape}")
        self.corners = c
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.cellshape = cellshape
        self.tag = tag


def _grid_points(bl, sub):
    """Natural coordinates of the block grid (sub = 1 or 2 points per cell
    edge) and their physical coordinates, k outer / j / i inner."""
    nx, ny, nz = bl.nx, bl.ny, bl.nz
    gi = np.arange(1, sub * nx + 2)
    gj = np.arange(1, sub * ny + 2)
    gk = np.arange(1, sub * nz + 2)
    r = -1.0 + 2.0 * ((1 / (sub * nx)) * (gi - 1))
    s = -1.0 + 2.0 * ((1 / (sub * ny)) * (gj - 1))
    t = -1.0 + 2.0 * ((1 / (sub * nz)) * (gk - 1))
    K, J, I = np.meshgrid(np.arange(len(gk)), np.arange(len(gj)), np.arange(len(gi)), indexing="ij")
    R, S, T = r[I.ravel()], s[J.ravel()], t[K.ravel()]
    X = np.zeros((R.size, 3))
    for a in range(8):
        Na = 0.125 * (1.0 + _HEX_SIGNS[a, 0] * R) * (1.0 + _HEX_SIGNS[a, 1] * S) * (1.0 + _HEX_SIGNS[a, 2] * T)
        for k in range(3):
            X[:, k] += Na * bl.corners[a, k]
    return round8(X), I.ravel(), J.ravel(), K.ravel(), (len(gi), len(gj), len(gk))


class Mesh:
    """Mesh(blocks...) for solid blocks of one cell shape.  With ``device``
    (a GPU index) a single block is generated there (``amr_block_mesh``):
    the same coordinates and connectivity, bit for bit."""

    def __init__(self, *blocks, device=None):
        blocks = [b for bl in blocks for b in (bl if isinstance(bl, (list, tuple)) else [bl])]
        if not blocks:
            raise ValueError("Mesh: no blocks")
        shapes = {b.cellshape for b in blocks}
        if len(shapes) != 1:
            raise ValueError("Mesh: all blocks must use one cell shape")
        self.cellshape = blocks[0].cellshape
        self._facets = None
        self._edges = None
        if device is not None:
            if len(blocks) != 1:
                raise ValueError("Mesh: device generation takes one block")
            self._device_block(blocks[0], int(device))
            return
        coords, conns, tags = [], [], []
        index = {}
        nnodes = 0
        for bl in blocks:
            sub = 1 if self.cellshape == HEX8 else 2
            X, I, J, K, (ni, nj, nk) = _grid_points(bl, sub)
            keep = np.ones(X.shape[0], dtype=bool)
            if self.cellshape == HEX20:
                ei, ej, ek = (I % 2 == 1), (J % 2 == 1), (K % 2 == 1)  # even 1-based index
                keep = ~((ei & ej) | (ej & ek) | (ek & ei))
            bnd = (I == 0) | (I == ni - 1) | (J == 0) | (J == nj - 1) | (K == 0) | (K == nk - 1)
            ids = -np.ones(X.shape[0], dtype=np.int64)
            if not index:
                # no earlier block: every kept point is new (vectorised)
                new_rows = np.nonzero(keep)[0]
                ids[new_rows] = nnodes + np.arange(new_rows.size)
                nnodes += new_rows.size
                if len(blocks) > 1:
                    for p in new_rows[bnd[new_rows]]:
                        index[tuple(X[p])] = ids[p]
            else:
                new_rows = []
                for p in np.nonzero(keep)[0]:
                    if bnd[p] and tuple(X[p]) in index:
                        ids[p] = index[tuple(X[p])]
                        continue
                    ids[p] = nnodes
                    nnodes += 1
                    new_rows.append(p)
                    if bnd[p]:
                        index[tuple(X[p])] = ids[p]
            coords.append(X[new_rows])
            P = ids.reshape(nk, nj, ni).transpose(2, 1, 0)  # P[i, j, k]
            conns.append(self._cells(P, bl, sub))
            tags += [bl.tag] * conns[-1].shape[0]
        self.coords = np.ascontiguousarray(np.vstack(coords))
        self.conn = np.ascontiguousarray(np.vstack(conns).astype(np.int32))
        self.tags = np.array(tags)
        self._facets = None
        self._edges = None

    def _device_block(self, bl, device):
        import ctypes
        from . import _lib
        et = {HEX8: _lib.AMR_HEX8, HEX20: _lib.AMR_HEX20, TET10: _lib.AMR_TET10}[self.cellshape]
        corners = np.ascontiguousarray(bl.corners, dtype=float)
        nn, ne = ctypes.c_int64(0), ctypes.c_int64(0)
        L = _lib.lib()
        _lib.check(L.amr_block_mesh(device, et, _lib.ptr(corners), bl.nx, bl.ny, bl.nz, ctypes.byref(nn),
                                    ctypes.byref(ne), None, None))
        self.coords = np.empty((nn.value, 3))
        self.conn = np.empty((ne.value, NNODES[self.cellshape]), dtype=np.int32)
        _lib.check(L.amr_block_mesh(device, et, _lib.ptr(corners), bl.nx, bl.ny, bl.nz, ctypes.byref(nn),
                                    ctypes.byref(ne), _lib.ptr(self.coords), _lib.ptr(self.conn)))
        self.tags = np.array([bl.tag] * ne.value)

    def _cells(self, P, bl, sub):
        nx, ny, nz = bl.nx, bl.ny, bl.nz
        if self.cellshape == HEX8:
            k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
            i, j, k = i.ravel(), j.ravel(), k.ravel()
            cols = [P[i + dx, j + dy, k + dz] for dx, dy, dz in
                    [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]]
            return np.stack(cols, axis=1)
        k, j, i = np.meshgrid(np.arange(0, 2 * nz, 2), np.arange(0, 2 * ny, 2), np.arange(0, 2 * nx, 2),
                              indexing="ij")
        i, j, k = i.ravel(), j.ravel(), k.ravel()
        pt = {q: P[i + d[0], j + d[1], k + d[2]] for q, d in _P27.items()
              if self.cellshape == TET10 or q <= 20}
        if self.cellshape == HEX20:
            return np.stack([pt[q] for q in range(1, 21)], axis=1)
        tets = [np.stack([pt[q] for q in split], axis=1) for split in _TET_SPLIT]
        return np.stack(tets, axis=1).reshape(-1, 10)

    @property
    def nnodes(self):
        return self.coords.shape[0]

    @property
    def ne(self):
        return self.conn.shape[0]

    def facets(self):
        """Outer facets (get_outer_facets, src/mesh/mesh.jl:69-85): facets seen
        once, first-appearance order.  Returns (owner element, node ids)."""
        if self._facets is None:
            self._facets = self._outer(None)
        return self._facets

    def facets_where(self, node_mask):
        """The outer facets whose nodes all satisfy ``node_mask``, in the order
        of ``facets()``.  A facet and its twin have the same nodes, so testing
        the mask before the seen-once test keeps exactly the outer facets
        among those that pass it, without sorting every facet of the mesh."""
        if self._facets is not None:
            owners, nodes = self._facets
            keep = node_mask[nodes].all(axis=1)
            return owners[keep], nodes[keep]
        return self._outer(node_mask)

    def _outer(self, node_mask):
        """Facets seen once among all (node_mask None) or among those whose
        nodes all pass node_mask.  In a conforming mesh two facets that share
        their three smallest corner ids are the same facet, so those three
        ids, packed in one int64, identify it (a 1-D sort instead of a sort
        of rows)."""
        F = FACETS[self.cellshape]
        nf, nfn = F.shape
        idx = None
        if node_mask is None:
            all_f = self.conn[:, F].reshape(-1, nfn)
        else:
            # only elements with at least a facet's worth of nodes in the mask
            # can own such a facet
            cm = node_mask[self.conn]
            el = np.nonzero(np.count_nonzero(cm, axis=1) >= nfn)[0]
            fm = cm[el][:, F].all(axis=2)
            e_i, f_i = np.nonzero(fm)
            idx = el[e_i] * nf + f_i   # ascending: el ascending, facets ascending within
            all_f = self.conn[el[e_i][:, None], F[f_i]]
        if self.nnodes < (1 << 21):
            c = np.sort(all_f[:, :3 if nfn == 6 else 4], axis=1)[:, :3].astype(np.int64)
            _, inv, cnt = np.unique((c[:, 0] << 42) | (c[:, 1] << 21) | c[:, 2], return_inverse=True,
                                    return_counts=True)
        else:
            _, inv, cnt = np.unique(np.sort(all_f, axis=1), axis=0, return_inverse=True, return_counts=True)
        outer = np.nonzero(cnt[inv.ravel()] == 1)[0]
        gidx = outer if idx is None else idx[outer]
        return gidx // nf, all_f[outer]

    def edges(self):
        """Unique edges of the outer facets, first-appearance order."""
        if self._edges is None:
            owners, fn = self.facets()
            E = FACET_EDGES[fn.shape[1]]
            all_e = fn[:, E].reshape(-1, E.shape[1])
            own = np.repeat(owners, E.shape[0])
            # an edge is identified by its two end nodes
            a, b = all_e[:, 0].astype(np.int64), all_e[:, 1].astype(np.int64)
            _, first = np.unique((np.minimum(a, b) << 32) | np.maximum(a, b), return_index=True)
            first = np.sort(first)
            self._edges = (own[first], all_e[first])
        return self._edges
