"""Python restatement of the host side of Amaru.jl's static mechanical path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Written with plain loops, for the small and medium cases the tests use.  The
element and material arithmetic is delegated to ``amaru_oracle.c``; the
linear solve is SuperLU (``scipy.sparse.linalg.splu``), the FP64 direct
oracle standing in for UMFPACK's ``lu(K11)`` of src/solver.jl:42-43.
"""
import ast
import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import HEX8, HEX20, NSTATE, TET10, lib, ptr, rule, shape_func

NNODES = {HEX8: 8, HEX20: 20, TET10: 10}
MAT_KIND = {"LinearElastic": 0, "VonMises": 1, "DruckerPrager": 2}
PRM_STRIDE = 6
EPS = 2.220446049250313e-16


# --------------------------------------------------------------- expressions

ARITH_TOL = 1e-6  # src/tools/expr.jl:96


def _jl_div(a, b):
    """Julia's div(a, b) on floats: (a - rem(a, b)) / b, truncated."""
    return float(math.trunc((a - math.fmod(a, b)) / b))


# op_dict (src/tools/expr.jl:98-125), one point at a time
_OPS = {
    "+": lambda *a: a[0] + a[1] if len(a) == 2 else +a[0],
    "-": lambda *a: a[0] - a[1] if len(a) == 2 else -a[0],
    "*": lambda a, b: a * b, "/": lambda a, b: a / b, "^": lambda a, b: a ** b,
    "div": _jl_div,
    ">": lambda a, b: a > b + ARITH_TOL, "<": lambda a, b: a < b - ARITH_TOL,
    ">=": lambda a, b: a > b - ARITH_TOL, "<=": lambda a, b: a < b + ARITH_TOL,
    "==": lambda a, b: abs(a - b) < ARITH_TOL, "!=": lambda a, b: abs(a - b) >= ARITH_TOL,
    "abs": abs, "sin": math.sin, "cos": math.cos, "tan": math.tan, "exp": math.exp, "log": math.log,
    "log10": math.log10, "sqrt": math.sqrt, "max": max, "min": min,
}
_PYOP = {"Add": "+", "Sub": "-", "Mult": "*", "Div": "/", "Pow": "^", "Gt": ">", "Lt": "<", "GtE": ">=",
         "LtE": "<=", "Eq": "==", "NotEq": "!=", "USub": "-", "UAdd": "+"}


def _expr(src):
    """Parse a reference-style condition/value expression (Julia syntax:
    && || comparisons and arithmetic on x, y, z, t) into a tree that
    ``evaluate`` reduces as reduce! does (src/tools/expr.jl:141-176)."""
    if not isinstance(src, str):
        return src
    s = src.replace("&&", " and ").replace("||", " or ").replace("^", "**")
    if "!" in s.replace("!=", ""):
        raise ValueError("operation ! not allowed in this context")
    try:
        return ("tree", ast.parse(s, mode="eval").body)
    except SyntaxError as e:
        raise ValueError(f"cannot parse {src!r}: {e}") from None


def _reduce(node, env):
    if isinstance(node, ast.Constant):
        return node.value
    if isinstance(node, ast.Name):
        if node.id == "pi":
            return math.pi
        if node.id not in env:
            raise ValueError(f"variable {node.id} not defined for this context")
        return env[node.id]
    if isinstance(node, ast.BoolOp):  # && / || short-circuit
        want = isinstance(node.op, ast.Or)
        for v in node.values:
            if bool(_reduce(v, env)) == want:
                return want
        return not want
    if isinstance(node, ast.Compare):  # :comparison chains
        vals = [_reduce(node.left, env)] + [_reduce(c, env) for c in node.comparators]
        for i, op in enumerate(node.ops):
            if not _OPS[_PYOP[type(op).__name__]](vals[i], vals[i + 1]):
                return False
        return True
    if isinstance(node, (ast.BinOp, ast.UnaryOp)):
        name = _PYOP.get(type(node.op).__name__)
        if name is None:
            raise ValueError(f"operation {type(node.op).__name__} not allowed in this context")
        args = [node.left, node.right] if isinstance(node, ast.BinOp) else [node.operand]
        return _OPS[name](*[_reduce(a, env) for a in args])
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Name):
        if node.func.id not in _OPS:
            raise ValueError(f"operation {node.func.id} not allowed in this context")
        args = [_reduce(a, env) for a in node.args]
        try:
            return _OPS[node.func.id](*args)
        except TypeError:  # Julia: no method matching the argument count
            raise ValueError(f"no method {node.func.id} for {len(args)} arguments") from None
    raise ValueError(f"operation {ast.dump(node)} not allowed in this context")


def evaluate(val, x=0.0, y=0.0, z=0.0, t=0.0):
    """evaluate (src/tools/expr.jl:178-181): numbers pass through,
    expressions are reduced with the point coordinates and time bound."""
    if isinstance(val, (int, float)):
        return float(val)
    if callable(val):
        return float(val(x, y, z, t))
    tree = _expr(val) if isinstance(val, str) else val
    return _reduce(tree[1], {"x": x, "y": y, "z": z, "t": t})


# -------------------------------------------------------------------- mesh

def _round8(v):
    """Node(x,y,z) rounding: round(x, digits=8) + 0.0 (src/node.jl:58-60)."""
    r = round(v * 1e8) / 1e8  # Python round: half to even, like Julia's RoundNearest
    return r + 0.0


def _box_corners(lo, hi):
    """box_coords (src/mesh/block.jl:3-30): HEX8-ordered corners."""
    x1, y1, z1 = lo
    lx, ly, lz = hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]
    return np.array([
        [x1, y1, z1], [x1 + lx, y1, z1], [x1 + lx, y1 + ly, z1], [x1, y1 + ly, z1],
        [x1, y1, z1 + lz], [x1 + lx, y1, z1 + lz], [x1 + lx, y1 + ly, z1 + lz], [x1, y1 + ly, z1 + lz]],
        dtype=float)


def _point(corners, r, s, t):
    N = shape_func(HEX8, np.array([r, s, t]))
    c = [0.0, 0.0, 0.0]
    for a in range(8):
        for k in range(3):
            c[k] += N[a] * corners[a, k]
    return tuple(_round8(v) for v in c)


def block_mesh(lo, hi, nx, ny, nz, etype):
    """split_block! for HEX8 (src/mesh/structured.jl:384-428), HEX20 and
    TET10 (:461-542) with uniform spacing, plus synchronize! numbering
    (src/mesh/mesh.jl:348-350: ids in creation order).
    Returns (coords[nnodes,3], conn[ne,nen]) with 0-based ids."""
    corners = _box_corners(lo, hi)
    nodes = []
    index = {}

    def new_node(c, boundary):
        if boundary:
            if c in index:
                return index[c]
            index[c] = len(nodes)
        nodes.append(c)
        return len(nodes) - 1

    conn = []
    if etype == HEX8:
        P = {}
        for k in range(1, nz + 2):
            for j in range(1, ny + 2):
                for i in range(1, nx + 2):
                    r = -1.0 + 2.0 * ((1 / nx) * (i - 1))
                    s = -1.0 + 2.0 * ((1 / ny) * (j - 1))
                    t = -1.0 + 2.0 * ((1 / nz) * (k - 1))
                    c = _point(corners, r, s, t)
                    bnd = i in (1, nx + 1) or j in (1, ny + 1) or k in (1, nz + 1)
                    P[i, j, k] = new_node(c, bnd)
        for k in range(1, nz + 1):
            for j in range(1, ny + 1):
                for i in range(1, nx + 1):
                    conn.append([P[i, j, k], P[i + 1, j, k], P[i + 1, j + 1, k], P[i, j + 1, k],
                                 P[i, j, k + 1], P[i + 1, j, k + 1], P[i + 1, j + 1, k + 1], P[i, j + 1, k + 1]])
    else:
        P = {}
        for k in range(1, 2 * nz + 2):
            for j in range(1, 2 * ny + 2):
                for i in range(1, 2 * nx + 2):
                    if etype == HEX20:
                        ei, ej, ek = i % 2 == 0, j % 2 == 0, k % 2 == 0
                        if (ei and ej) or (ej and ek) or (ek and ei):
                            continue
                    r = -1.0 + 2.0 * ((1 / (2 * nx)) * (i - 1))
                    s = -1.0 + 2.0 * ((1 / (2 * ny)) * (j - 1))
                    t = -1.0 + 2.0 * ((1 / (2 * nz)) * (k - 1))
                    c = _point(corners, r, s, t)
                    bnd = i in (1, 2 * nx + 1) or j in (1, 2 * ny + 1) or k in (1, 2 * nz + 1)
                    P[i, j, k] = new_node(c, bnd)
        for k in range(1, 2 * nz, 2):
            for j in range(1, 2 * ny, 2):
                for i in range(1, 2 * nx, 2):
                    p = [None] * 28
                    p[1], p[2], p[3], p[4] = P[i, j, k], P[i + 2, j, k], P[i + 2, j + 2, k], P[i, j + 2, k]
                    p[5], p[6], p[7], p[8] = (P[i, j, k + 2], P[i + 2, j, k + 2], P[i + 2, j + 2, k + 2],
                                              P[i, j + 2, k + 2])
                    p[9], p[10], p[11], p[12] = P[i + 1, j, k], P[i + 2, j + 1, k], P[i + 1, j + 2, k], P[i, j + 1, k]
                    p[13], p[14], p[15], p[16] = (P[i + 1, j, k + 2], P[i + 2, j + 1, k + 2], P[i + 1, j + 2, k + 2],
                                                  P[i, j + 1, k + 2])
                    p[17], p[18], p[19], p[20] = (P[i, j, k + 1], P[i + 2, j, k + 1], P[i + 2, j + 2, k + 1],
                                                  P[i, j + 2, k + 1])
                    if etype == HEX20:
                        conn.append(p[1:21])
                    else:
                        p[21], p[22], p[23] = P[i, j + 1, k + 1], P[i + 2, j + 1, k + 1], P[i + 1, j, k + 1]
                        p[24], p[25], p[26] = P[i + 1, j + 2, k + 1], P[i + 1, j + 1, k], P[i + 1, j + 1, k + 2]
                        p[27] = P[i + 1, j + 1, k + 1]
                        # six tets per hex cell, src/mesh/structured.jl:537-542
                        for q in ((2, 4, 1, 8, 25, 12, 9, 27, 20, 21), (2, 1, 5, 8, 9, 17, 23, 27, 21, 16),
                                  (2, 5, 6, 8, 23, 13, 18, 27, 16, 26), (2, 6, 7, 8, 18, 14, 22, 27, 26, 15),
                                  (2, 3, 4, 8, 10, 11, 25, 27, 24, 20), (2, 7, 3, 8, 22, 19, 10, 27, 15, 24)):
                            conn.append([p[v] for v in q])
    return np.array(nodes, dtype=float), np.array(conn, dtype=np.int32)


# Facet node lists per element shape (1-based in the reference:
# src/shape/solids3d.jl facet_idxs_HEX8 / _HEX20 / _TET10), and the facet
# shape used for surface integrals.
FACETS = {
    HEX8: [[0, 4, 7, 3], [1, 2, 6, 5], [0, 1, 5, 4], [2, 3, 7, 6], [0, 3, 2, 1], [4, 5, 6, 7]],
    HEX20: [[0, 4, 7, 3, 16, 15, 19, 11], [1, 2, 6, 5, 9, 18, 13, 17], [0, 1, 5, 4, 8, 17, 12, 16],
            [2, 3, 7, 6, 10, 19, 14, 18], [0, 3, 2, 1, 11, 10, 9, 8], [4, 5, 6, 7, 12, 13, 14, 15]],
    TET10: [[0, 3, 2, 7, 9, 6], [0, 1, 3, 4, 8, 7], [0, 2, 1, 6, 5, 4], [1, 2, 3, 5, 9, 8]],
}
# edge node lists of the facet shapes (QUAD4/QUAD8/TRI6 facet_idxs)
FACET_EDGES = {4: [[0, 1], [1, 2], [2, 3], [3, 0]],
               8: [[0, 1, 4], [1, 2, 5], [2, 3, 6], [3, 0, 7]],
               6: [[0, 1, 3], [1, 2, 4], [2, 0, 5]]}


def outer_facets(conn, etype):
    """get_outer_facets (src/mesh/mesh.jl:69-85): facets seen once, in order
    of first appearance.  Returns list of (owner element, node ids)."""
    order = {}
    for e, el in enumerate(conn):
        for f in FACETS[etype]:
            nodes = [int(el[i]) for i in f]
            key = frozenset(nodes)
            if key in order:
                del order[key]
            else:
                order[key] = (e, nodes)
    return list(order.values())


def facet_edges(facets):
    """Unique edges of a facet list, first appearance order."""
    seen = {}
    for owner, nodes in facets:
        for ed in FACET_EDGES[len(nodes)]:
            en = [nodes[i] for i in ed]
            key = frozenset(en)
            if key not in seen:
                seen[key] = (owner, en)
    return list(seen.values())


# ------------------------------------------------------- facet integration

def _facet_shape(npts, R):
    """QUAD4 / QUAD8 / TRI6 / LIN2 / LIN3 shape functions and derivatives
    (src/shape/solids2d.jl:141-151, 317-427; src/shape/lines.jl)."""
    r, s = R[0], R[1]
    if npts == 4:
        nat = [(-1, -1), (1, -1), (1, 1), (-1, 1)]
        N = [0.25 * (1 + a * r) * (1 + b * s) for a, b in nat]
        D = [[0.25 * a * (1 + b * s), 0.25 * b * (1 + a * r)] for a, b in nat]
    elif npts == 8:
        nat = [(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0)]
        N, D = [], []
        for a, b in nat:
            if a != 0 and b != 0:
                N.append(0.25 * (1 + a * r) * (1 + b * s) * (a * r + b * s - 1))
                D.append([0.25 * a * (1 + b * s) * (2 * a * r + b * s), 0.25 * b * (1 + a * r) * (a * r + 2 * b * s)])
            elif a == 0:
                N.append(0.5 * (1 - r * r) * (1 + b * s))
                D.append([-r * (1 + b * s), 0.5 * b * (1 - r * r)])
            else:
                N.append(0.5 * (1 + a * r) * (1 - s * s))
                D.append([0.5 * a * (1 - s * s), -s * (1 + a * r)])
    elif npts == 6:
        L0 = 1 - r - s
        N = [L0 * (2 * L0 - 1), r * (2 * r - 1), s * (2 * s - 1), 4 * r * L0, 4 * r * s, 4 * s * L0]
        D = [[-(4 * L0 - 1), -(4 * L0 - 1)], [4 * r - 1, 0.0], [0.0, 4 * s - 1],
             [4 * (L0 - r), -4 * r], [4 * s, 4 * r], [-4 * s, 4 * (L0 - s)]]
    elif npts == 2:
        N = [0.5 * (1 - r), 0.5 * (1 + r)]
        D = [[-0.5], [0.5]]
    elif npts == 3:
        N = [0.5 * r * (r - 1), 0.5 * r * (r + 1), 1 - r * r]
        D = [[r - 0.5], [r + 0.5], [-2 * r]]
    else:
        raise ValueError(npts)
    return np.array(N), np.array(D)


G2 = 0.577350269189626
_FACET_RULES = {
    4: ([(-G2, -G2), (G2, -G2), (-G2, G2), (G2, G2)], [1.0] * 4),          # QUAD_IP4
    8: ([(-G2, -G2), (G2, -G2), (-G2, G2), (G2, G2)], [1.0] * 4),          # QUAD_IP4
    6: ([(1 / 6, 1 / 6), (2 / 3, 1 / 6), (1 / 6, 2 / 3)], [1 / 6] * 3),    # TRI_IP3
    2: ([(-0.577350269189625764509149, 0.0), (0.577350269189625764509149, 0.0)], [1.0, 1.0]),  # LIN_IP2
    3: ([(-0.774596669241483377035835, 0.0), (0.0, 0.0), (0.774596669241483377035835, 0.0)],
        [0.555555555555555555555556, 0.888888888888888888888889, 0.555555555555555555555556]),  # LIN_IP3
}


def boundary_forces(C, key, val, t=0.0):
    """mech_boundary_forces (src/mech/elem/distributed.jl:75-150) for a facet
    or an edge with nodal coordinates C (n×3).  Returns F (n×3)."""
    n = C.shape[0]
    F = np.zeros((n, 3))
    pts, ws = _FACET_RULES[n]
    edge = key in ("qx", "qy", "qz", "qn")
    for (r, s), w in zip(pts, ws):
        N, D = _facet_shape(n, (r, s))
        J = C.T @ D
        X = C.T @ N
        vip = evaluate(val, X[0], X[1], X[2], t)
        Q = np.zeros(3)
        if key in ("tx", "qx"):
            Q[0] = vip
        elif key in ("ty", "qy"):
            Q[1] = vip
        elif key in ("tz", "qz"):
            Q[2] = vip
        elif key == "tn":
            nrm = np.cross(J[:, 0], J[:, 1])
            Q = vip * nrm / np.linalg.norm(nrm)
        else:
            raise ValueError(f"boundary key {key} not supported")
        if edge:
            jn = np.linalg.norm(J[:, 0])
        else:
            jn = np.linalg.norm(np.cross(J[:, 0], J[:, 1]))
        F += (jn * w) * np.outer(N, Q)
    return F


def body_forces(etype, nips, C, key, val):
    """mech_solid_body_forces (src/mech/elem/distributed.jl:155-210)."""
    R, w = rule(etype, 0)
    F = np.zeros((C.shape[0], 3))
    from . import shape_deriv
    for q in range(len(w)):
        N = shape_func(etype, R[q])
        D = shape_deriv(etype, R[q])
        J = C.T @ D
        X = C.T @ N
        vip = evaluate(val, X[0], X[1], X[2])
        Q = np.zeros(3)
        Q["xyz".index(key[1])] = vip
        F += (np.linalg.det(J) * w[q]) * np.outer(N, Q)
    return F


# ------------------------------------------------------------------- model

class BC:
    """NodeBC / SurfaceBC / EdgeBC / BodyC (src/bc.jl:10-153)."""

    def __init__(self, kind, **conds):
        self.kind = kind
        self.conds = dict(conds)


def NodeBC(**c):
    return BC("node", **c)


def SurfaceBC(**c):
    return BC("surface", **c)


def EdgeBC(**c):
    return BC("edge", **c)


def BodyC(**c):
    return BC("body", **c)


class Model:
    """FEModel (src/fe-model.jl:66-255) for one solid element type."""

    def __init__(self, coords, conn, etype, mats, emat=None, nips=0):
        self.X = np.ascontiguousarray(coords, dtype=float)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32)
        self.etype = etype
        self.nen = NNODES[etype]
        self.ne = self.conn.shape[0]
        self.nnodes = self.X.shape[0]
        R, w = rule(etype, nips)
        self.nips = len(w)
        self.kinds = np.array([MAT_KIND[m[0]] for m in mats], dtype=np.int32)
        self.prms = np.zeros((len(mats), PRM_STRIDE))
        for i, (name, p) in enumerate(mats):
            self.prms[i, :len(p)] = p
        self.emat = np.zeros(self.ne, dtype=np.int32) if emat is None else np.ascontiguousarray(emat, np.int32)
        self.state = np.zeros((self.ne * self.nips, NSTATE))
        self.facets = outer_facets(self.conn, etype)
        self.edges = facet_edges(self.facets)
        # nodal values (dof.vals): u and f per (node, dir)
        self.uvals = np.zeros((self.nnodes, 3))
        self.fvals = np.zeros((self.nnodes, 3))

    def node_filter(self, expr):
        code = _expr(expr)
        return [i for i in range(self.nnodes) if evaluate(code, *self.X[i])]

    def _facet_filter(self, items, expr):
        code = _expr(expr)
        return [it for it in items if all(evaluate(code, *self.X[n]) for n in it[1])]

    def configure_dofs(self, bcs):
        """configure_dofs! (src/bc.jl:198-233): DOFs of active nodes in node
        order (ux,uy,uz), unknowns first (stable), eq ids = positions."""
        active = sorted(set(int(v) for v in self.conn.ravel()))
        presc = np.zeros((self.nnodes, 3), dtype=bool)
        self.bc_sel = []
        for flt, bc in bcs:
            if bc.kind == "node":
                sel = self.node_filter(flt)
                nodes = sel
            elif bc.kind == "surface":
                sel = self._facet_filter(self.facets, flt)
                nodes = [n for f in sel for n in f[1]]
            elif bc.kind == "edge":
                sel = self._facet_filter(self.edges, flt)
                nodes = [n for f in sel for n in f[1]]
            else:
                code = _expr(flt)
                sel = [e for e in range(self.ne) if all(evaluate(code, *self.X[n]) for n in self.conn[e])]
                nodes = [int(n) for e in sel for n in self.conn[e]]
            for key in bc.conds:
                if key in ("ux", "uy", "uz"):
                    for n in nodes:
                        presc[n, "xyz".index(key[1])] = True
            self.bc_sel.append(sel)
        dofs = [(n, d) for n in active for d in range(3)]
        ud = [x for x in dofs if not presc[x]]
        pd = [x for x in dofs if presc[x]]
        self.eq = -np.ones((self.nnodes, 3), dtype=np.int32)
        for i, (n, d) in enumerate(ud + pd):
            self.eq[n, d] = i
        self.ndofs = len(dofs)
        self.nu = len(ud)
        self.eqmap = np.ascontiguousarray(self.eq[self.conn].reshape(self.ne, 3 * self.nen))
        return self.ndofs, self.nu

    def bc_vals(self, bcs, t=0.0):
        """get_bc_vals / compute_bc_vals! (src/bc.jl:43-58, 116-136, 175-194,
        237-249)."""
        U = np.zeros(self.ndofs)
        F = np.zeros(self.ndofs)
        for (flt, bc), sel in zip(bcs, self.bc_sel):
            if bc.kind == "node":
                for n in sel:
                    x, y, z = self.X[n]
                    for key, cond in bc.conds.items():
                        d = "xyz".index(key[1])
                        if key[0] == "u":
                            U[self.eq[n, d]] = evaluate(cond, x, y, z, t)
                        elif key[0] == "f":
                            F[self.eq[n, d]] += evaluate(cond, x, y, z, t)
            elif bc.kind in ("surface", "edge"):
                for owner, nodes in sel:
                    for key, val in bc.conds.items():
                        if key in ("ux", "uy", "uz"):
                            for n in nodes:
                                U[self.eq[n, "xyz".index(key[1])]] = evaluate(val, *self.X[n], t)
                        else:
                            Fd = boundary_forces(self.X[nodes], key, val, t)
                            for a, n in enumerate(nodes):
                                for d in range(3):
                                    F[self.eq[n, d]] += Fd[a, d]
            else:
                for e in sel:
                    for key, val in bc.conds.items():
                        nodes = self.conn[e]
                        if key in ("ux", "uy", "uz"):
                            for n in nodes:
                                U[self.eq[n, "xyz".index(key[1])]] = evaluate(val, *self.X[n], t)
                        else:
                            Fd = body_forces(self.etype, self.nips, self.X[nodes], key, val)
                            for a, n in enumerate(nodes):
                                for d in range(3):
                                    F[self.eq[n, d]] += Fd[a, d]
        return U, F

    # ---------------------------------------------------------- hot path
    def elem_stiffness(self, state=None, nthreads=1):
        """Ke for all elements (C restatement of elem_stiffness)."""
        st = self.state if state is None else state
        nd = 3 * self.nen
        Ke = np.zeros((self.ne, nd, nd))
        bad = lib().ora_elem_stiffness(self.etype, self.nips, self.ne, ptr(self.X), ptr(self.conn),
                                       ptr(self.emat), ptr(self.kinds), ptr(self.prms), PRM_STRIDE,
                                       ptr(np.ascontiguousarray(st)), ptr(Ke), nthreads)
        if bad:
            raise RuntimeError(f"Negative Jacobian determinant in cell {bad}")
        return Ke

    def mount_K(self, state=None, nthreads=1):
        """mount_K (src/mech/mech-solver.jl:78-110) → scipy CSR (ndofs²)."""
        Ke = self.elem_stiffness(state, nthreads)
        return csr_from_Ke(Ke, self.eqmap, self.ndofs)

    def elem_mass(self, rho):
        """Me for all elements (C restatement of elem_mass,
        src/mech/elem/mech-solid.jl:169-205); rho per element."""
        nd = 3 * self.nen
        Me = np.zeros((self.ne, nd, nd))
        rho = np.ascontiguousarray(np.broadcast_to(np.asarray(rho, dtype=float), (self.ne,)))
        bad = lib().ora_elem_mass(self.etype, self.nips, self.ne, ptr(self.X), ptr(self.conn), ptr(rho), ptr(Me))
        if bad:
            raise RuntimeError(f"Negative Jacobian determinant in cell {bad}")
        return Me

    def mount_M(self, rho):
        """mount_M (src/mech/dyn-solver.jl:72-104): the mount_K triplet rule
        over the element mass matrices → scipy CSR (ndofs²)."""
        return csr_from_Ke(self.elem_mass(rho), self.eqmap, self.ndofs)

    def update_state(self, dUt, state_in):
        """update_state!(elems, ΔUt, t) (src/mech/mech-solver.jl:124-144).
        Returns (ΔFin, new_state, status_message_or_None)."""
        nd = 3 * self.nen
        st = np.array(state_in, copy=True)
        dF = np.zeros((self.ne, nd))
        status = np.zeros(self.ne, dtype=np.int32)
        lib().ora_update_elems(self.etype, self.nips, self.ne, ptr(self.X), ptr(self.conn), ptr(self.eqmap),
                               ptr(self.emat), ptr(self.kinds), ptr(self.prms), PRM_STRIDE,
                               ptr(np.ascontiguousarray(dUt, dtype=float)), ptr(st), ptr(dF), ptr(status), 1)
        dFin = np.zeros(self.ndofs)
        lib().ora_gather_forces(self.ne, nd, ptr(dF), ptr(self.eqmap), self.ndofs, ptr(dFin))
        if status.any():
            e = int(np.nonzero(status)[0][0])
            msg = "Negative Jacobian determinant" if status[e] == 2 else "material update failed"
            return dFin, st, f"{msg} in cell {e + 1}"
        if np.isnan(dFin).any():
            return dFin, st, "solve_system!: NaN values in internal forces vector"
        return dFin, st, None

    def ip_stress(self, e):
        """σxx..σxy per IP of element e (stress_strain_dict,
        src/tools/tensors.jl:162-215): shear = Mandel / sqrt(2)."""
        s = self.state[e * self.nips:(e + 1) * self.nips, :6].copy()
        s[:, 3:] /= math.sqrt(2.0)
        return s


def csr_from_Ke(Ke, eqmap, ndofs):
    """Apply the mount_K triplet rule to a stack of element matrices."""
    ne, nd, _ = Ke.shape
    Ke = np.ascontiguousarray(Ke)
    ntrip = lib().ora_count_triplets(ne, nd, ptr(Ke))
    rowptr = np.zeros(ndofs + 1, dtype=np.int64)
    col = np.zeros(max(ntrip, 1), dtype=np.int32)
    val = np.zeros(max(ntrip, 1))
    nnz = lib().ora_mount_K(ne, nd, ptr(Ke), ptr(np.ascontiguousarray(eqmap)), ndofs, ptr(rowptr), ptr(col),
                            ptr(val))
    if nnz < 0:
        raise MemoryError("mount_K")
    return sp.csr_matrix((val[:nnz], col[:nnz], rowptr), shape=(ndofs, ndofs))


def solve_system(K, U, F, nu):
    """solve_system! (src/solver.jl:5-79) with SuperLU in place of UMFPACK.
    Mutates U[:nu] and F[nu:]; returns None or a failure message."""
    K = K.tocsr()
    U2 = U[nu:]
    F2 = K[nu:, nu:] @ U2
    U1 = np.zeros(nu)
    if nu > 0:
        RHS = F[:nu] - K[:nu, nu:] @ U2
        try:
            lu = spla.splu(K[:nu, :nu].tocsc())
            U1 = lu.solve(RHS)
            F2 = F2 + K[nu:, :nu] @ U1
        except RuntimeError as err:
            msg = ""
            if np.isnan(K[:nu, :nu].data).any():
                msg = "\nsolve_system!: NaN values in coefficients matrix"
            return f"{msg}\nsolve_system!: {err}"
    if nu > 0 and np.max(np.abs(U1)) > 1e8:
        return "\nsolve_system!: Possible syngular matrix"
    U[:nu] = U1
    F[nu:] = F2
    return None


# ------------------------------------------------------------ stage solver

# (p1, q11, a1, a2) of the predictor-corrector schemes (src/mech/mech-solver.jl:280-288)
SCHEMES = {"FE": (1.0, 1.0, None, None), "ME": (1.0, 1.0, 0.5, 0.5), "BE": (1.0, 1.0, 0.0, 1.0),
           "Ralston": (2 / 3, 2 / 3, 1 / 4, 3 / 4)}


def nr_step(model, state_bk, state, dUa, dUi, R, nu, scheme="FE"):
    """The solve and state update of one Newton-Raphson iteration of
    mech_stage_solver! (src/mech/mech-solver.jl:325-349): the predictor
    solve_system!(K, p1 ΔUi, q11 R) with K = mount_K at ``state`` and
    update_state!(ΔUa + ΔUitr) from ``state_bk``; for :ME, :BE and :Ralston
    the corrector with K = a1 K + a2 K2 (K2 = mount_K at the predicted
    state), solve_system!(K, ΔUi, R) and update_state!(ΔUa + ΔUi).
    ``dUi`` holds the iteration's essential values.  Returns (ΔUi, ΔFin,
    state, message or None)."""
    p1, q11, a1, a2 = SCHEMES[scheme]
    K = model.mount_K(state)
    dUitr = p1 * dUi
    Rtr = q11 * R
    msg = solve_system(K, dUitr, Rtr, nu)
    if msg is not None:
        return dUi, None, state, msg
    dFin, state, msg = model.update_state(dUa + dUitr, state_bk)
    if msg is not None or scheme == "FE":
        return dUitr, dFin, state, msg
    K2 = model.mount_K(state)
    K = (a1 * K + a2 * K2).tocsr()
    dUi = dUi.copy()
    msg = solve_system(K, dUi, R.copy(), nu)
    if msg is not None:
        return dUi, None, state, msg
    dFin, state, msg = model.update_state(dUa + dUi, state_bk)
    return dUi, dFin, state, msg


def solve_stage(model, bcs, nincs=1, nouts=0, tol=0.01, rtol=0.01, rspan=0.01, dT0=0.01, dTmin=1e-7,
                dTmax=0.1, maxits=5, autoinc=False, on_increment=None, log=None, scheme="FE"):
    """mech_stage_solver! (src/mech/mech-solver.jl:186-492) with the
    predictor-corrector schemes :FE, :ME, :BE, :Ralston (:319-349),
    including the stage_iterator! nincs/nouts adjustment (src/solver.jl:128-142).
    Returns (success, message)."""
    p1, q11, a1, a2 = SCHEMES[scheme]
    if nouts > 0 and not autoinc:
        if nouts > nincs:
            nincs = nouts
    ndofs, nu = model.configure_dofs(bcs)
    Uex, Fex = model.bc_vals(bcs)
    state_bk = model.state.copy()
    state = model.state.copy()
    dofs_u = np.zeros(ndofs)
    # eq id → (node, dir)
    inv = np.zeros((ndofs, 2), dtype=np.int64)
    for n in range(model.nnodes):
        for d in range(3):
            if model.eq[n, d] >= 0:
                inv[model.eq[n, d]] = (n, d)

    dTbk = 0.0
    dTcheck = 1.0 / nouts if nouts > 0 else 1.0
    Tcheck = dTcheck
    T = 0.0
    dT = 1.0 / nincs
    if autoinc:
        dT = min(dT, dTmax, dTcheck, dT0)
    U = np.zeros(ndofs)
    F = np.zeros(ndofs)
    R = np.zeros(ndofs)
    dFin = np.zeros(ndofs)
    dUa = np.zeros(ndofs)
    dUi = np.zeros(ndofs)
    Rc = np.zeros(ndofs)
    still_linear = True
    inc = 0
    status_msg = None
    while T < 1.0 - dTmin:
        inc += 1
        dUex, dFex = dT * Uex, dT * Fex
        dTcr = min(rspan, 1 - T)
        acr = min(dT / dTcr, 1.0)
        if T < 1 - rspan:
            dFex = dFex + acr * Rc
        R[:] = dFex
        dUa[:] = 0.0
        dUi = dUex.copy()
        nits = 0
        err = 0.0
        res = 0.0
        res1 = 0.0
        converged = False
        syserror = False
        sysmsg = None
        for it in range(1, maxits + 1):
            nits += 1
            if it > 1:
                dUi[:] = 0.0
            lastres = res
            dUi, dFin, state, sysmsg = nr_step(model, state_bk, state, dUa, dUi, R, nu, scheme)
            if sysmsg is not None:
                syserror = True
                break
            dUa += dUi
            R = dFex - dFin
            R[nu:] = 0.0
            res = np.max(np.abs(R)) if ndofs else 0.0
            err = np.max(np.abs(dUi)) / np.max(np.abs(dUa)) if ndofs else 0.0
            if log is not None:
                log.append((inc, it, res))
            if it == 1:
                res1 = res
            if it > 1:
                still_linear = False
            if res < tol:
                converged = True
                break
            if err < rtol:
                converged = True
                break
            if math.isnan(res):
                break
            if it > 1 and res > lastres:
                break
        if syserror:
            converged = False
        if converged:
            U += dUa
            F += dFin
            Rc = (1.0 - acr) * Rc + R
            state_bk = state.copy()
            for i in range(ndofs):
                n, d = inv[i]
                model.uvals[n, d] += dUa[i]
                model.fvals[n, d] += dFin[i]
            T += dT
            checkpoint = T > Tcheck - dTmin
            if checkpoint:
                Tcheck += dTcheck
            model.state = state_bk.copy()
            if on_increment is not None:
                on_increment(model, T)
            if autoinc:
                if dTbk > 0.0:
                    dT = min(dTbk, Tcheck - T)
                    dTbk = 0.0
                else:
                    q = 1 + math.tanh(math.log10(tol / (res1 + EPS)))
                    q = max(q, 1.1)
                    dTtr = min(q * dT, dTmax, 1 - T)
                    if T + dTtr > Tcheck - dTmin:
                        dTbk = dT
                        dT = Tcheck - T
                    else:
                        dT = dTtr
                        dTbk = 0.0
        else:
            inc -= 1
            state = state_bk.copy()
            if autoinc:
                q = 1 + math.tanh(math.log10(tol / (res1 + EPS)))
                q = min(max(q, 0.2), 0.9)
                if syserror:
                    q = 0.7
                dT = q * dT
                dT = float(f"{dT:.3g}")
                if dT < dTmin:
                    status_msg = "Solver did not converge."
                    break
            else:
                status_msg = "Solver did not converge. Try `autoinc=true`. "
                break
    model.state = state_bk.copy()
    return status_msg is None, status_msg
