"""FE model, materials and boundary conditions (host side).

Mirrors the reference's model API: ``FEModel(mesh, matbinds, ctx)``
(src/fe-model.jl:66-255), the material constructors (LinearElastic
src/mech/mat/linear-elastic.jl:5-20, VonMises von-mises.jl:4-26,
DruckerPrager drucker-prager.jl:5-34), the boundary conditions NodeBC /
SurfaceBC / EdgeBC / BodyC (src/bc.jl:10-153), ``configure_dofs!``
(src/bc.jl:198-233) and ``get_bc_vals`` (src/bc.jl:237-249) with the
distributed loads of src/mech/elem/distributed.jl:75-210.
"""
import ctypes

import numpy as np

from . import _lib
from .expr import compile_expr
from .mesh import HEX8, HEX20, NNODES, TET10

ETYPE = {HEX8: _lib.AMR_HEX8, HEX20: _lib.AMR_HEX20, TET10: _lib.AMR_TET10}
NSTATE = 14


# ------------------------------------------------------------- materials

class Material:
    kind = None
    names = ()

    def params(self):
        p = np.zeros(6)
        v = [getattr(self, n) for n in self.names]
        p[:len(v)] = v
        return p


class LinearElastic(Material):
    """LinearElastic(E, nu) (src/mech/mat/linear-elastic.jl:5-20)."""
    kind = _lib.AMR_MAT_LE
    names = ("E", "nu")

    def __init__(self, E, nu=0.0):
        if not E > 0.0:
            raise ValueError("LinearElastic: E>0 required")
        if not 0.0 <= nu < 0.5:
            raise ValueError("LinearElastic: 0<=nu<0.5 required")
        self.E, self.nu = float(E), float(nu)


class VonMises(Material):
    """VonMises(E, nu, fy, H) with linear hardening (src/mech/mat/von-mises.jl:4-26)."""
    kind = _lib.AMR_MAT_VM
    names = ("E", "nu", "fy", "H")

    def __init__(self, E, fy, nu=0.0, H=0.0, rho=0.0):
        if not (E > 0.0 and 0.0 <= nu < 0.5 and fy >= 0.0 and H >= 0.0):
            raise ValueError("VonMises: invalid parameters")
        self.E, self.nu, self.fy, self.H = float(E), float(nu), float(fy), float(H)


class DruckerPrager(Material):
    """DruckerPrager(E, nu, alpha, kappa, H) (src/mech/mat/drucker-prager.jl:5-34)."""
    kind = _lib.AMR_MAT_DP
    names = ("E", "nu", "alpha", "kappa", "H")

    def __init__(self, E, nu=0.0, alpha=0.0, kappa=0.0, H=0.0, rho=0.0):
        if not (E > 0.0 and 0.0 <= nu < 0.5 and alpha >= 0.0 and kappa > 0.0 and H >= 0.0):
            raise ValueError("DruckerPrager: invalid parameters")
        self.E, self.nu, self.alpha, self.kappa, self.H = float(E), float(nu), float(alpha), float(kappa), float(H)


class MechSolid:
    """The isoparametric 3D bulk element (src/mech/elem/mech-solid.jl).  In a
    material binding it may appear as the class or as an instance carrying the
    element properties of MechSolidProps (src/mech/elem/mech-solid.jl:13-27):
    ``MechSolid(rho=7850.0)``."""

    def __init__(self, rho=0.0, gamma=0.0):
        if rho < 0:
            raise ValueError("MechSolid: rho must be >= 0")
        self.rho = float(rho)
        self.gamma = float(gamma)


# ------------------------------------------------------- boundary conditions

class BC:
    kind = None

    def __init__(self, **conds):
        if not conds:
            raise ValueError(f"{type(self).__name__} must have at least one condition")
        self.conds = dict(conds)


class NodeBC(BC):
    kind = "node"


class SurfaceBC(BC):
    kind = "surface"


FaceBC = SurfaceBC


class EdgeBC(BC):
    kind = "edge"


class BodyC(BC):
    kind = "body"


ESSENTIAL = ("ux", "uy", "uz")
DIR = {"x": 0, "y": 1, "z": 2, "n": -1}


# ------------------------------------------------ facet shapes and rules

def _facet_tables(npts):
    """Shape functions / derivatives of QUAD4, QUAD8, TRI6, LIN2, LIN3 at the
    default rule (QUAD_IP4, TRI_IP3, LIN_IP2, LIN_IP3; src/shape/quadrature.jl)."""
    g2 = 0.577350269189626
    if npts in (4, 8):
        pts = [(-g2, -g2), (g2, -g2), (-g2, g2), (g2, g2)]
        w = [1.0] * 4
    elif npts == 6:
        pts = [(1 / 6, 1 / 6), (2 / 3, 1 / 6), (1 / 6, 2 / 3)]
        w = [1 / 6] * 3
    elif npts == 2:
        a = 0.577350269189625764509149
        pts, w = [(-a, 0.0), (a, 0.0)], [1.0, 1.0]
    else:
        a = 0.774596669241483377035835
        pts = [(-a, 0.0), (0.0, 0.0), (a, 0.0)]
        w = [0.555555555555555555555556, 0.888888888888888888888889, 0.555555555555555555555556]
    Ns, Ds = [], []
    for r, s in pts:
        if npts == 4:
            nat = [(-1, -1), (1, -1), (1, 1), (-1, 1)]
            N = [0.25 * (1 + a * r) * (1 + b * s) for a, b in nat]
            D = [[0.25 * a * (1 + b * s), 0.25 * b * (1 + a * r)] for a, b in nat]
        elif npts == 8:
            nat = [(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0)]
            N, D = [], []
            for a, b in nat:
                if a and b:
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
            D = [[-(4 * L0 - 1), -(4 * L0 - 1)], [4 * r - 1, 0.0], [0.0, 4 * s - 1], [4 * (L0 - r), -4 * r],
                 [4 * s, 4 * r], [-4 * s, 4 * (L0 - s)]]
        elif npts == 2:
            N = [0.5 * (1 - r), 0.5 * (1 + r)]
            D = [[-0.5], [0.5]]
        else:
            N = [0.5 * r * (r - 1), 0.5 * r * (r + 1), 1 - r * r]
            D = [[r - 0.5], [r + 0.5], [-2 * r]]
        Ns.append(N)
        Ds.append(D)
    return np.array(Ns), np.array(Ds), np.array(w)


def distributed_forces(C, key, val, t=0.0):
    """mech_boundary_forces (src/mech/elem/distributed.jl:75-150) for many
    facets/edges at once: C is (nf, npts, 3); returns (nf, npts, 3)."""
    nf, npts, _ = C.shape
    N, D, w = _facet_tables(npts)
    f = compile_expr(val)
    edge = key[0] == "q"
    F = np.zeros((nf, npts, 3))
    for q in range(len(w)):
        J = np.einsum("fak,ai->fki", C, D[q])        # (nf, 3, ndim_facet)
        X = np.einsum("fak,a->fk", C, N[q])
        vip = f(X[:, 0], X[:, 1], X[:, 2], t)
        Q = np.zeros((nf, 3))
        d = DIR[key[1]]
        if d >= 0:
            Q[:, d] = vip
        else:
            n = np.cross(J[:, :, 0], J[:, :, 1])
            Q = vip[:, None] * n / np.linalg.norm(n, axis=1)[:, None]
        if edge:
            jn = np.linalg.norm(J[:, :, 0], axis=1)
        else:
            jn = np.linalg.norm(np.cross(J[:, :, 0], J[:, :, 1]), axis=1)
        F += (jn * w[q])[:, None, None] * N[q][None, :, None] * Q[:, None, :]
    return F


def cell_loads(device, C, N, D, w, d, val, t=0.0):
    """amr_cell_loads: force vectors (ncells, npts, 3) of a distributed load
    over cells with node coordinates C (ncells, npts, 3), computed on the
    device; N, D, w the cell shape's tables, d a direction (0..2) or -1 for
    the facet normal, val a number or expression over x, y, z, t."""
    from .expr import compile_program
    ops, consts = compile_program(val)
    C = np.ascontiguousarray(C, dtype=float)
    nc, npts, _ = C.shape
    N = np.ascontiguousarray(N, dtype=float)
    D = np.ascontiguousarray(D, dtype=float)
    w = np.ascontiguousarray(w, dtype=float)
    nq = len(w)
    ndim = D.shape[-1] if D.ndim == 3 else 1
    D = np.ascontiguousarray(D.reshape(nq, npts, ndim))
    F = np.empty_like(C)
    _lib.check(_lib.lib().amr_cell_loads(int(device), nc, npts, ndim, nq, _lib.ptr(C), _lib.ptr(N),
                                         _lib.ptr(D), _lib.ptr(w), int(d),
                                         _lib.ptr(ops), _lib.ptr(consts) if consts.size else None, len(ops),
                                         len(consts), float(t), _lib.ptr(F)))
    return F


def distributed_forces_device(device, C, key, val, t=0.0):
    """distributed_forces on the device: the same (nf, npts, 3) result."""
    npts = C.shape[1]
    N, D, w = _facet_tables(npts)
    return cell_loads(device, C, N, D, w, DIR[key[1]], val, t)


# ------------------------------------------------------------------ model

class FEModel:
    """FEModel(mesh, matbinds) for one solid element shape.

    matbinds: list of ``(filter, MechSolid, material)`` or
    ``(filter, material)``; filter is "all", a mesh tag, or an expression
    over node coordinates that every node of an element must satisfy.
    nips selects the IP rule (0 = shape default; 27 is
    ``changequadrature!(elems, 27)`` for HEX20, src/element.jl:185-188).
    """

    def __init__(self, mesh, matbinds, nips=0):
        self.mesh = mesh
        self.shape = mesh.cellshape
        self.etype = ETYPE[self.shape]
        self.nen = NNODES[self.shape]
        self.coords = mesh.coords
        self.conn = mesh.conn
        self.nnodes, self.ne = mesh.nnodes, mesh.ne
        self.nips = self._rule(nips)
        mats, emat = [], -np.ones(self.ne, dtype=np.int32)
        self.rho = np.zeros(self.ne)   # MechSolidProps.rho per element
        for bind in matbinds:
            flt, mat = bind[0], bind[-1]
            if not isinstance(mat, Material):
                raise TypeError("FEModel: material binding must end with a material instance")
            sel = self.elem_filter(flt)
            mats.append(mat)
            emat[sel] = len(mats) - 1
            props = bind[1] if len(bind) == 3 else None
            self.rho[sel] = props.rho if isinstance(props, MechSolid) else 0.0
        if (emat < 0).any():
            raise ValueError(f"FEModel: missing material definition for {int((emat < 0).sum())} elements")
        self.mats = mats
        self.emat = emat
        self.state = np.zeros((self.ne * self.nips, NSTATE))
        self.uvals = np.zeros((self.nnodes, 3))   # dof.vals[ux,uy,uz]
        self.fvals = np.zeros((self.nnodes, 3))   # dof.vals[fx,fy,fz]
        self.eq = None

    # The IP states (one row of NSTATE values per IP).  After an analysis
    # stage they are the device model's committed states, copied to the host
    # when first read (or when that device model is closed or starts another
    # stage), so that analyses whose states nobody reads pay no copy.
    @property
    def state(self):
        dev = self.__dict__.get("_state_dev")
        if dev is not None:
            self._state_dev = None
            self._state = dev.state_get(committed=True)
        return self._state

    @state.setter
    def state(self, st):
        self._state_dev = None
        self._state = st

    def state_on_device(self, dev):
        """The states are dev's committed ones (read when first needed)."""
        self._state_dev = dev

    def _rule(self, nips):
        if self.shape in (HEX8, HEX20):
            if nips in (0, 8):
                return 8
            if nips == 27:
                return 27
        elif nips in (0, 4):
            return 4
        raise ValueError(f"{self.shape} cannot use {nips} integration points")

    def changequadrature(self, nips):
        """changequadrature!(elems, n): new rule, states re-initialised."""
        self.nips = self._rule(nips)
        self.state = np.zeros((self.ne * self.nips, NSTATE))

    # -------------------------------------------------------- filters
    def _node_mask(self, expr):
        X = self.coords
        return np.asarray(compile_expr(expr)(X[:, 0], X[:, 1], X[:, 2]), dtype=bool)

    def node_filter(self, expr):
        return np.nonzero(self._node_mask(expr))[0]

    def elem_filter(self, flt):
        if flt in ("all", ":bulks", ":solids", "bulks", "solids"):
            return np.arange(self.ne)
        if isinstance(flt, str) and flt in set(self.mesh.tags.tolist()):
            return np.nonzero(self.mesh.tags == flt)[0]
        m = self._node_mask(flt)
        return np.nonzero(m[self.conn].all(axis=1))[0]

    def _sel(self, bc, flt):
        if bc.kind == "node":
            n = self.node_filter(flt)
            return n, n
        if bc.kind == "body":
            e = self.elem_filter(flt)
            return e, self.conn[e].ravel()
        m = self._node_mask(flt)
        if bc.kind == "surface":
            owners, nodes = self.mesh.facets_where(m)
            return (owners, nodes), nodes.ravel()
        owners, nodes = self.mesh.edges()
        keep = m[nodes].all(axis=1)
        return (owners[keep], nodes[keep]), nodes[keep].ravel()

    # ----------------------------------------------------------- dofs
    def configure_dofs(self, bcs, device=None):
        """configure_dofs! (src/bc.jl:198-233): DOFs of the active nodes in node
        order (ux,uy,uz), stable partition unknowns-first, eq ids = position.
        With ``device`` (a GPU index) the numbering is computed there
        (``amr_dof_map``) from the prescribed flags selected here."""
        # the same boundary conditions again (a stage re-configures the DOFs it
        # was set up with): the numbering cannot have changed
        dkey = tuple((str(flt), bc.kind, tuple(sorted((k, repr(v)) for k, v in bc.conds.items())))
                     for flt, bc in bcs)
        if self.eq is not None and getattr(self, "_dof_key", None) == dkey:
            return self.ndofs, self.nu
        active = np.nonzero(np.bincount(self.conn.ravel(), minlength=self.nnodes))[0]
        presc = np.zeros((self.nnodes, 3), dtype=bool)
        self._bcsel = []
        for flt, bc in bcs:
            sel, nodes = self._sel(bc, flt)
            for key in bc.conds:
                if key in ESSENTIAL:
                    presc[nodes, DIR[key[1]]] = True
            self._bcsel.append(sel)
        if device is not None:
            act = np.zeros(self.nnodes, dtype=np.uint8)
            act[active] = 1
            pre = np.ascontiguousarray(presc.astype(np.uint8).ravel())
            eq = np.empty(self.nnodes * 3, dtype=np.int32)
            nd, nu = ctypes.c_int64(0), ctypes.c_int64(0)
            _lib.check(_lib.lib().amr_dof_map(int(device), self.nnodes, _lib.ptr(act), _lib.ptr(pre), _lib.ptr(eq),
                                              ctypes.byref(nd), ctypes.byref(nu)))
            self.eq = eq.reshape(self.nnodes, 3)
            self.ndofs, self.nu = int(nd.value), int(nu.value)
            self._dof_key = dkey
            return self.ndofs, self.nu
        dof = (active[:, None] * 3 + np.arange(3)[None, :]).ravel()
        p = presc.ravel()[dof]
        order = np.concatenate([dof[~p], dof[p]])
        eq = -np.ones(self.nnodes * 3, dtype=np.int32)
        eq[order] = np.arange(order.size, dtype=np.int32)
        self.eq = eq.reshape(self.nnodes, 3)
        self.ndofs = int(order.size)
        self.nu = int((~p).sum())
        self._dof_key = dkey
        return self.ndofs, self.nu

    def bc_vals(self, bcs, t=0.0, device=None):
        """get_bc_vals (src/bc.jl:237-249): essential values U and natural
        loads F, in eq order.  With ``device`` (a GPU index) the distributed
        loads (tractions, line loads, body forces) are integrated there."""
        U = np.zeros(self.ndofs)
        F = np.zeros(self.ndofs)
        X = self.coords
        for (flt, bc), sel in zip(bcs, self._bcsel):
            for key, val in bc.conds.items():
                f = compile_expr(val)
                if bc.kind == "node":
                    d = DIR[key[1]]
                    v = f(X[sel, 0], X[sel, 1], X[sel, 2], t)
                    if key in ESSENTIAL:
                        U[self.eq[sel, d]] = v
                    elif key[0] == "f":
                        np.add.at(F, self.eq[sel, d], v)
                    continue
                if bc.kind == "body":
                    nodes = self.conn[sel]
                else:
                    nodes = sel[1]
                if key in ESSENTIAL:
                    nn = nodes.ravel()
                    U[self.eq[nn, DIR[key[1]]]] = f(X[nn, 0], X[nn, 1], X[nn, 2], t)
                    continue
                if bc.kind == "body":
                    Fd = self._body_forces(sel, key, f, val, device)
                elif device is not None:
                    Fd = distributed_forces_device(device, X[nodes], key, val, t)
                else:
                    Fd = distributed_forces(X[nodes], key, val, t)
                np.add.at(F, self.eq[nodes].reshape(-1), Fd.reshape(-1))
        return U, F

    def _body_forces(self, elems, key, f, val=None, device=None):
        """mech_solid_body_forces (src/mech/elem/distributed.jl:155-210) with
        the shape's default rule (values at t = 0)."""
        from .shapes import rule_tables
        N, D, w = rule_tables(self.shape, 0)
        C = self.coords[self.conn[elems]]
        d = DIR[key[1]]
        if device is not None:
            return cell_loads(device, C, N, D, w, d, val, 0.0)
        F = np.zeros(C.shape)
        for q in range(len(w)):
            J = np.einsum("eak,ai->eki", C, D[q])
            X = np.einsum("eak,a->ek", C, N[q])
            vip = f(X[:, 0], X[:, 1], X[:, 2])
            coef = np.linalg.det(J) * w[q]
            F[:, :, d] += (coef * vip)[:, None] * N[q][None, :]
        return F

    # -------------------------------------------------------- outputs
    def ip_stress(self):
        """sigma_xx..sigma_xy per IP (stress_strain_dict,
        src/tools/tensors.jl:162-215: shear = Mandel / sqrt(2))."""
        s = self.state[:, :6].copy()
        s[:, 3:] /= np.sqrt(2.0)
        return s

    def node_data(self):
        """get_data(model.nodes) (src/node.jl:270-281): ux uy uz fx fy fz."""
        u, f = self.uvals, self.fvals
        return {"ux": u[:, 0], "uy": u[:, 1], "uz": u[:, 2], "fx": f[:, 0], "fy": f[:, 1], "fz": f[:, 2]}


OUTPUT_FIELDS = ("σxx", "σyy", "σzz", "σyz", "σxz", "σxy", "σvm", "σ1", "σ2", "σ3",
                 "εxx", "εyy", "εzz", "εyz", "εxz", "εxy")
_EXTRA_FIELDS = {"LinearElastic": (), "VonMises": ((16, "ep"),),
                 "DruckerPrager": ((16, "epa"), (17, "j1"), (18, "srj2d"))}


def output_fields(model):
    """(column, name) of the output table for the model's material
    (stress_strain_dict, src/tools/tensors.jl:190-213, plus the material's
    own entries)."""
    kinds = {type(m).__name__ for m in model.mats}
    plastic = kinds - {"LinearElastic"}
    if len(plastic) > 1:
        raise ValueError("output fields: von Mises and Drucker-Prager in one model")
    # a model that mixes kinds: the stress_strain_dict fields of every
    # element, then the plastic kind's extras (recovered on the subpatches of
    # the elements that carry them)
    return list(enumerate(OUTPUT_FIELDS)) + list(_EXTRA_FIELDS[plastic.pop() if plastic else "LinearElastic"])


class DeviceModel:
    """An FEModel resident on one GPU behind the C ABI."""

    def __init__(self, model, device=0):
        import ctypes
        L = _lib.lib()
        if model.eq is None:
            model.configure_dofs([])
        self.model = model
        kinds = np.array([m.kind for m in model.mats], dtype=np.int32)
        prms = np.ascontiguousarray(np.stack([m.params() for m in model.mats]))
        self._h = ctypes.c_void_p()
        self._keep = (np.ascontiguousarray(model.coords), np.ascontiguousarray(model.conn, dtype=np.int32),
                      kinds, prms, np.ascontiguousarray(model.emat, dtype=np.int32),
                      np.ascontiguousarray(model.eq.ravel(), dtype=np.int32))
        X, C, K, Pm, E, Q = self._keep
        _lib.check(L.amr_model_create(ctypes.byref(self._h), device, model.etype, model.nips, model.nnodes,
                                      _lib.ptr(X), model.ne, _lib.ptr(C), len(model.mats), _lib.ptr(K),
                                      _lib.ptr(Pm), _lib.ptr(E), _lib.ptr(Q), model.ndofs, model.nu))
        self.device = device
        self._eq = Q.copy()
        if model.state.any():
            self.state_set(model.state)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            if self.model.__dict__.get("_state_dev") is self:
                self.model.state  # noqa: B018  (copy the committed states first)
            _lib.lib().amr_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        out = np.zeros(8, dtype=np.int64)
        _lib.check(_lib.lib().amr_model_info(self._h, _lib.ptr(out)))
        return dict(zip(["nnodes", "ne", "nips", "nblk", "nnz_max", "ncontrib", "ndofs", "nu"], out.tolist()))

    def set_dofs(self):
        """Upload the model's DOF numbering if it changed since the last upload."""
        m = self.model
        eq = np.ascontiguousarray(m.eq.ravel(), dtype=np.int32)
        if np.array_equal(eq, self._eq):
            return
        _lib.check(_lib.lib().amr_model_set_dofs(self._h, _lib.ptr(eq), m.ndofs, m.nu))
        self._eq = eq

    def state_set(self, st):
        st = np.ascontiguousarray(st, dtype=float)
        _lib.check(_lib.lib().amr_state_set(self._h, _lib.ptr(st)))

    def state_get(self, committed=False):
        st = np.empty((self.model.ne * self.model.nips, NSTATE))
        _lib.check(_lib.lib().amr_state_get(self._h, _lib.ptr(st), 1 if committed else 0))
        return st

    def commit(self):
        _lib.check(_lib.lib().amr_state_commit(self._h))

    def restore(self):
        _lib.check(_lib.lib().amr_state_restore(self._h))

    def assemble(self):
        _lib.check(_lib.lib().amr_assemble(self._h))

    def Ke(self):
        nd = 3 * self.model.nen
        out = np.zeros((self.model.ne, nd, nd))
        _lib.check(_lib.lib().amr_get_Ke(self._h, _lib.ptr(out)))
        return out

    def K_csc(self):
        import ctypes
        import scipy.sparse as sp
        L = _lib.lib()
        nnz = ctypes.c_int64(0)
        _lib.check(L.amr_get_K_csc(self._h, ctypes.byref(nnz), None, None, None))
        n = self.model.ndofs
        colptr = np.zeros(n + 1, dtype=np.int64)
        rowval = np.zeros(max(nnz.value, 1), dtype=np.int32)
        nzval = np.zeros(max(nnz.value, 1))
        _lib.check(L.amr_get_K_csc(self._h, ctypes.byref(nnz), _lib.ptr(colptr), _lib.ptr(rowval), _lib.ptr(nzval)))
        return sp.csc_matrix((nzval[:nnz.value], rowval[:nnz.value], colptr), shape=(n, n))

    def solve_system(self, U, F, precision="fp64", rtol=1e-12, maxiter=200000, check_every=32, replace_ratio=0.0,
                     drift_limit=0.0, precond="jacobi"):
        """solve_system!(K, U, F, nu): mutates U[:nu] and F[nu:]."""
        import ctypes
        o = self.opts(precision, rtol, maxiter=maxiter, check_every=check_every, replace_ratio=replace_ratio,
                      drift_limit=drift_limit, precond=precond)
        info = _lib.SolveInfo()
        assert U.flags["C_CONTIGUOUS"] and F.flags["C_CONTIGUOUS"] and U.dtype == F.dtype == np.float64
        rc = _lib.lib().amr_solve_system(self._h, _lib.ptr(U), _lib.ptr(F), ctypes.byref(o), ctypes.byref(info))
        _lib.check(rc)
        return {"iterations": info.iterations, "replacements": info.refinements, "relres": info.relres,
                "converged": bool(info.converged), "fallback": bool(info.fallback)}

    # ------------------------------------------- device-resident stage driver
    @staticmethod
    def opts(precision="fp64", rtol=1e-10, maxiter=200000, check_every=32, replace_ratio=0.0, drift_limit=0.0,
             precond="jacobi"):
        prec = {"fp64": _lib.AMR_PREC_FP64, "mixed": _lib.AMR_PREC_MIXED, "mixed64acc": _lib.AMR_PREC_MIXED_F64ACC}
        pc = {"jacobi": _lib.AMR_PRECOND_JACOBI, "block_jacobi": _lib.AMR_PRECOND_BLOCK_JACOBI,
              "two_level": _lib.AMR_PRECOND_TWO_LEVEL}
        return _lib.SolveOpts(prec[precision], pc[precond], int(maxiter), int(check_every), float(rtol), 0.0,
                              float(replace_ratio), float(drift_limit))

    def stage_begin(self, Uex, Fex):
        Uex = np.ascontiguousarray(Uex, dtype=float)
        Fex = np.ascontiguousarray(Fex, dtype=float)
        _lib.check(_lib.lib().amr_stage_begin(self._h, _lib.ptr(Uex), _lib.ptr(Fex)))

    def increment_begin(self, dT, acr, add_residual):
        _lib.check(_lib.lib().amr_increment_begin(self._h, float(dT), float(acr), 1 if add_residual else 0))

    def nr_iteration(self, it, opts):
        """One NR iteration on the device; returns the amr_nr_info fields.
        A failed iteration is reported in 'status'/'message', not raised."""
        import ctypes
        info = _lib.NRInfo()
        rc = _lib.lib().amr_nr_iteration(self._h, int(it), ctypes.byref(opts), ctypes.byref(info))
        return self._nr_result(rc, info)

    def nr_prepare(self, it):
        """First half of nr_iteration: mount_K and the solve's right-hand side, queued without waiting."""
        _lib.check(_lib.lib().amr_nr_prepare(self._h, int(it)))

    def nr_finish(self, solve_info):
        """Second half of nr_iteration, given the linear solve's outcome (a SolveInfo, or None
        for a model without unknowns); returns what nr_iteration returns."""
        import ctypes
        info = _lib.NRInfo()
        si = ctypes.byref(solve_info) if solve_info is not None else None
        rc = _lib.lib().amr_nr_finish(self._h, si, ctypes.byref(info))
        return self._nr_result(rc, info)

    def nr_finish_begin(self, solve_info):
        """First half of nr_finish: queue the update and the residual read-back."""
        import ctypes
        si = ctypes.byref(solve_info) if solve_info is not None else None
        _lib.check(_lib.lib().amr_nr_finish_begin(self._h, si))

    @staticmethod
    def nr_prepare_many(devs, its):
        """nr_prepare(it) of every model, their element kernels batched (one
        launch per kernel family for all of them)."""
        import ctypes
        n = len(devs)
        if n == 0:
            return
        hs = (ctypes.c_void_p * n)(*[d._h.value for d in devs])
        iv = (ctypes.c_int * n)(*[int(i) for i in its])
        _lib.check(_lib.lib().amr_nr_prepare_many(hs, iv, n))

    @staticmethod
    def nr_finish_begin_many(devs, solve_infos):
        """nr_finish_begin of every model (a SolveInfo each), their update
        kernels batched."""
        import ctypes
        n = len(devs)
        if n == 0:
            return
        hs = (ctypes.c_void_p * n)(*[d._h.value for d in devs])
        si = (_lib.SolveInfo * n)(*solve_infos)
        _lib.check(_lib.lib().amr_nr_finish_begin_many(hs, si, n))

    def nr_finish_end(self):
        """Second half of nr_finish: wait and return what nr_iteration returns."""
        import ctypes
        info = _lib.NRInfo()
        rc = _lib.lib().amr_nr_finish_end(self._h, ctypes.byref(info))
        return self._nr_result(rc, info)

    def _nr_result(self, rc, info):
        out = {"res": info.res, "err": info.err, "relres": info.relres, "pcg_iterations": info.pcg_iterations,
               "replacements": info.replacements, "fallback": bool(info.fallback), "status": rc, "message": None}
        if rc != _lib.AMR_OK:
            out["message"] = _lib.lib().amr_last_error().decode(errors="replace")
            if rc in (-1, -2, -7):   # argument, runtime or memory errors are not solver failures
                raise _lib.AmaruError(rc, out["message"])
        return out

    def set_coarse(self, target_nodes=256, max_aggregates=170):
        _lib.check(_lib.lib().amr_set_coarse(self._h, int(target_nodes), int(max_aggregates)))

    COARSE_REFRESH = {"dof_map": 0, "increment": 1, "assembly": 2}

    def set_coarse_refresh(self, policy="dof_map"):
        """When the two-level coarse operator is formed again: "dof_map" (the
        default), "increment" (non-elastic models) or "assembly"."""
        _lib.check(_lib.lib().amr_set_coarse_refresh(self._h, self.COARSE_REFRESH[policy]))

    def coarse_info(self):
        out = np.zeros(2, dtype=np.int64)
        _lib.check(_lib.lib().amr_get_coarse(self._h, _lib.ptr(out)))
        return {"aggregates": int(out[0]), "coarse_dofs": int(out[1])}

    def coarse_space(self):
        """(aggregate of each node, Kc^-1) of the two-level preconditioner."""
        info = self.coarse_info()
        agg = np.zeros(self.model.nnodes, dtype=np.int32)
        kinv = np.zeros((info["coarse_dofs"], info["coarse_dofs"]))
        _lib.check(_lib.lib().amr_get_coarse_space(self._h, _lib.ptr(agg), _lib.ptr(kinv)))
        return agg, kinv

    def device_bytes(self):
        import ctypes
        b = ctypes.c_int64(0)
        _lib.check(_lib.lib().amr_model_device_bytes(self._h, ctypes.byref(b)))
        return b.value

    def increment_commit(self):
        _lib.check(_lib.lib().amr_increment_commit(self._h))

    def increment_reject(self):
        _lib.check(_lib.lib().amr_increment_reject(self._h))

    def stage_get(self, last=False):
        U = np.zeros(self.model.ndofs)
        F = np.zeros(self.model.ndofs)
        _lib.check(_lib.lib().amr_stage_get(self._h, _lib.ptr(U), _lib.ptr(F), 1 if last else 0))
        return U, F

    def spmv_lanes(self):
        out = np.zeros(3, dtype=np.int32)
        _lib.check(_lib.lib().amr_get_spmv_lanes(self._h, _lib.ptr(out)))
        return {"fp64": int(out[0]), "fp32": int(out[1]), "fp32_f64acc": int(out[2])}

    def set_spmv_lanes(self, lanes_fp64=0, lanes_fp32=0):
        _lib.check(_lib.lib().amr_set_spmv_lanes(self._h, int(lanes_fp64), int(lanes_fp32)))

    def tune_spmv_lanes(self):
        """Time every lanes choice on the assembled matrix; keep the fastest."""
        _lib.check(_lib.lib().amr_tune_spmv_lanes(self._h))

    def timing(self, enable):
        import ctypes
        ms = ctypes.c_double(0.0)
        n = ctypes.c_int64(0)
        _lib.check(_lib.lib().amr_timing(self._h, 1 if enable else 0, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def assemble_M(self, rho=None):
        """mount_M (src/mech/dyn-solver.jl:72-104) on the device: element
        mass matrices (elem_mass, src/mech/elem/mech-solid.jl:169-205) and the
        eps-filtered assembly; rho per element (default: the model's
        MechSolid densities)."""
        r = self.model.rho if rho is None else rho
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(r, dtype=float), (self.model.ne,)))
        _lib.check(_lib.lib().amr_assemble_M(self._h, _lib.ptr(r)))

    def M_csc(self):
        """The assembled M as scipy CSC over the current DOF numbering."""
        import scipy.sparse as sp
        nnz = ctypes.c_int64(0)
        L = _lib.lib()
        _lib.check(L.amr_get_M_csc(self._h, ctypes.byref(nnz), None, None, None))
        colptr = np.zeros(self.model.ndofs + 1, dtype=np.int64)
        rowval = np.zeros(max(nnz.value, 1), dtype=np.int32)
        nzval = np.zeros(max(nnz.value, 1))
        _lib.check(L.amr_get_M_csc(self._h, ctypes.byref(nnz), _lib.ptr(colptr), _lib.ptr(rowval), _lib.ptr(nzval)))
        return sp.csc_matrix((nzval[:nnz.value], rowval[:nnz.value], colptr), shape=(self.model.ndofs,) * 2)

    def spmv_col_bytes(self):
        """Bytes per block of the column indices the PCG SpMV reads (2 or 4)."""
        b = ctypes.c_int(0)
        _lib.check(_lib.lib().amr_spmv_col_bytes(self._h, ctypes.byref(b)))
        return b.value

    def set_ke_path(self, path):
        """Element stiffness of the NR iterations: "exact" (bit for bit the
        oracle's) or "mfma" (HEX20 on the FP64 matrix cores, equal to rounding)."""
        _lib.check(_lib.lib().amr_set_ke_path(self._h, {"exact": 0, "mfma": 1}[path]))

    def elem_tiles(self):
        """(ach, epw): the element stiffness and update tilings in use."""
        a, e = ctypes.c_int(0), ctypes.c_int(0)
        _lib.check(_lib.lib().amr_get_elem_tiles(self._h, ctypes.byref(a), ctypes.byref(e)))
        return a.value, e.value

    def set_elem_tiles(self, ach=0, epw=0):
        _lib.check(_lib.lib().amr_set_elem_tiles(self._h, int(ach), int(epw)))

    def tune_elem_tiles(self):
        """Times the tilings and keeps the fastest (for this element shape in
        the process too); returns (ach, epw)."""
        a, e = ctypes.c_int(0), ctypes.c_int(0)
        _lib.check(_lib.lib().amr_tune_elem_tiles(self._h, ctypes.byref(a), ctypes.byref(e)))
        return a.value, e.value

    SCHEMES = {"FE": 0, "ME": 1, "BE": 2, "Ralston": 3}

    def set_scheme(self, scheme):
        """Predictor-corrector scheme of nr_iteration (mech_stage_solver!'s
        scheme: FE, ME, BE or Ralston)."""
        _lib.check(_lib.lib().amr_set_scheme(self._h, self.SCHEMES[str(scheme).lstrip(":")]))

    def set_spmv_col_bytes(self, nbytes):
        """4: the SpMV reads the 32-bit columns; 2: the 16-bit offsets
        (when every block's fits).  Results do not depend on it."""
        _lib.check(_lib.lib().amr_set_spmv_col_bytes(self._h, int(nbytes)))

    def combine_KM(self, a, b):
        """The device K becomes a K + b M (the dynamic solver's pseudo-stiffness)."""
        _lib.check(_lib.lib().amr_combine_KM(self._h, float(a), float(b)))

    def matvec(self, which, x):
        """K x ("K": the device K), M x ("M") or K x with K as assembled
        before the last combine_KM ("K0"); x and the result in eq order."""
        x = np.ascontiguousarray(x, dtype=float)
        assert x.shape == (self.model.ndofs,)
        y = np.zeros(self.model.ndofs)
        _lib.check(_lib.lib().amr_matvec(self._h, {"K": 0, "M": 1, "K0": 2}[which], _lib.ptr(x), _lib.ptr(y)))
        return y

    # -------------------------------------------------------- outputs
    def ip_values(self, committed=False):
        """ip_state_vals of every IP (nip_total x 19; ``OUTPUT_FIELDS``)."""
        out = np.empty((self.model.ne * self.model.nips, _lib.AMR_NFIELDS))
        _lib.check(_lib.lib().amr_ip_values(self._h, int(committed), _lib.ptr(out)))
        return out

    def nodal_values(self, committed=False):
        """nodal_patch_recovery (src/fe-model.jl:506-692) of the IP values:
        (nnodes x 19 values, patches reaching each node)."""
        out = np.empty((self.model.nnodes, _lib.AMR_NFIELDS))
        reps = np.empty(self.model.nnodes, dtype=np.int32)
        _lib.check(_lib.lib().amr_nodal_recovery(self._h, int(committed), _lib.ptr(out), _lib.ptr(reps)))
        return out, reps

    def node_fields(self, committed=False):
        """The recovered nodal fields by the reference's names for the
        model's material (σxx … εxy, plus ep or epa, j1, srj2d)."""
        vals, _ = self.nodal_values(committed)
        return {name: vals[:, i] for i, name in output_fields(self.model)}

    def update_state(self, dUt):
        """update_state!(elems, ΔUt, t) → ΔFin (raises on failure)."""
        dUt = np.ascontiguousarray(dUt, dtype=float)
        out = np.zeros(self.model.ndofs)
        _lib.check(_lib.lib().amr_update_state(self._h, _lib.ptr(dUt), _lib.ptr(out)))
        return out
