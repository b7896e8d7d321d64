"""Python restatement of Amaru.jl's output post-processing.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

- ``ip_values``: ``ip_state_vals`` of the three solid materials
  (src/mech/mat/linear-elastic.jl:140-142, von-mises.jl:159-168,
  drucker-prager.jl:152-163) over ``stress_strain_dict``
  (src/tools/tensors.jl:162-215) for the 3D stress model.
- ``nodal_patch_recovery``: src/fe-model.jl:506-692, one thread (patch
  contributions summed in patch order), with ``pinv`` as Julia's: SVD with
  singular values below ``eps*min(m, n)*sigma_max`` dropped.

Values are returned in a fixed table of 19 columns (``FIELDS``); a material
uses the first 16 and its own extras (``MAT_FIELDS``).
"""
import numpy as np

from . import HEX8, HEX20, TET10, rule, shape_func
from .fem import outer_facets

SR2 = 2.0 ** 0.5
EPS = 2.220446049250313e-16
FIELDS = ("σxx", "σyy", "σzz", "σyz", "σxz", "σxy", "σvm", "σ1", "σ2", "σ3",
          "εxx", "εyy", "εzz", "εyz", "εxz", "εxy", "εp", "j1", "srj2d")
# extras per material kind (0 LE, 1 VM, 2 DP): (column, reference name)
MAT_FIELDS = {0: (), 1: ((16, "ep"),), 2: ((16, "epa"), (17, "j1"), (18, "srj2d"))}
NCORNERS = {HEX8: 8, HEX20: 8, TET10: 4}


def J2(s):
    """tensors.jl:24-27 (Mandel input, shear / sqrt 2)."""
    t11, t22, t33 = s[..., 0], s[..., 1], s[..., 2]
    t23, t13, t12 = s[..., 3] / SR2, s[..., 4] / SR2, s[..., 5] / SR2
    return 1 / 6 * ((t11 - t22) ** 2 + (t22 - t33) ** 2 + (t33 - t11) ** 2) + t23 ** 2 + t13 ** 2 + t12 ** 2


def ip_values(state, kind):
    """(nip, 19) values of ``ip_state_vals`` per IP; ``kind`` is the material
    kind of every IP.  Eigenvalues: tensors.jl:80-93, descending."""
    s, e = state[:, 0:6], state[:, 6:12]
    n = len(state)
    V = np.zeros((n, len(FIELDS)))
    V[:, 0:3] = s[:, 0:3]
    V[:, 3:6] = s[:, 3:6] / SR2
    j2 = J2(s)
    V[:, 6] = np.sqrt(3 * j2)
    T = np.zeros((n, 3, 3))
    T[:, 0, 0], T[:, 1, 1], T[:, 2, 2] = s[:, 0], s[:, 1], s[:, 2]
    T[:, 1, 2] = T[:, 2, 1] = s[:, 3] / SR2
    T[:, 0, 2] = T[:, 2, 0] = s[:, 4] / SR2
    T[:, 0, 1] = T[:, 1, 0] = s[:, 5] / SR2
    V[:, 7:10] = np.linalg.eigvalsh(T)[:, ::-1]
    V[:, 10:13] = e[:, 0:3]
    V[:, 13:16] = e[:, 3:6] / SR2
    vm, dp = kind == 1, kind == 2
    V[vm, 16] = state[vm, 12]
    V[dp, 16] = state[dp, 12]
    V[dp, 17] = s[dp, 0] + s[dp, 1] + s[dp, 2]
    V[dp, 18] = np.sqrt(j2[dp])
    return V


def ip_coords(X, conn, etype, nips=0):
    """Global IP coordinates ``C'*N`` per element and IP (element.jl:160-163)."""
    R, _ = rule(etype, nips)
    N = np.array([shape_func(etype, r) for r in R])      # (nips, nn)
    C = X[conn]                                            # (ne, nn, 3)
    return np.einsum("qa,eak->eqk", N, C).reshape(-1, 3)


def reg_terms(P, nterms):
    """fe-model.jl:498-503 for ndim 3."""
    x, y, z = P[:, 0], P[:, 1], P[:, 2]
    one = np.ones_like(x)
    if nterms == 7:
        return np.stack([one, x, y, z, x * y, y * z, x * z], axis=1)
    if nterms == 4:
        return np.stack([one, x, y, z], axis=1)
    return one[:, None]


def patches(conn, etype, nnodes):
    """The patch of every node (fe-model.jl:522-578): lists of elements, in
    element order; empty where a node has none."""
    at_bound = np.zeros(nnodes, bool)
    for _, nodes in outer_facets(conn, etype):
        at_bound[nodes] = True
    nc = NCORNERS[etype]
    inner = [[] for _ in range(nnodes)]
    bry = [[] for _ in range(nnodes)]
    for e, el in enumerate(conn):
        for n in el[:nc]:
            (bry if at_bound[n] else inner)[n].append(e)
    has = np.zeros(nnodes, bool)
    for p in inner:
        for e in p:
            has[conn[e]] = True
    orphans = [n for n in range(nnodes) if not has[n] and at_bound[n]]
    if orphans:
        for nmin in (3, 2, 1):
            for n in orphans:
                if len(bry[n]) >= nmin:
                    inner[n] = bry[n]
            for n in orphans:
                for e in inner[n]:
                    has[conn[e]] = True
            orphans = [n for n in orphans if not has[n]]
            if not orphans:
                break
    return inner


def nodal_patch_recovery(X, conn, etype, W, nips=0, has=None):
    """(nnodes, nf) recovered nodal values of the IP values ``W`` (ne*nips,
    nf), and the number of patches that reached each node.

    ``has`` (ne, nf) marks the fields each element carries (default: all).
    A field is fitted on the subpatch of the patch's elements that carry it
    and averaged over the subpatches that reach a node (fe-model.jl:614-680:
    ``subpatch``, ``V_reps[node, field]``)."""
    nnodes = len(X)
    R, _ = rule(etype, nips)
    nq = len(R)
    nf = W.shape[1]
    XI = ip_coords(X, conn, etype, nips)
    if has is None:
        has = np.ones((len(conn), nf), dtype=bool)
    V = np.zeros((nnodes, nf))
    reps = np.zeros((nnodes, nf), dtype=np.int64)
    for patch in patches(conn, etype, nnodes):
        if not patch:
            continue
        last = None
        for f in range(nf):
            sub = [e for e in patch if has[e, f]]
            if not sub:
                continue
            if sub != last:
                last = sub
                ips = np.concatenate([np.arange(e * nq, (e + 1) * nq) for e in sub])
                nodes = list(dict.fromkeys(int(n) for e in sub for n in conn[e]))
                m = len(ips)
                nterms = 7 if m >= 7 else 4 if m >= 4 else 1
                M = reg_terms(XI[ips], nterms)
                invM = np.linalg.pinv(M, rcond=EPS * min(M.shape))
                N = reg_terms(X[nodes], nterms)
            V[nodes, f] += N @ (invM @ W[ips, f])
            reps[nodes, f] += 1
    with np.errstate(invalid="ignore", divide="ignore"):
        out = V / reps
    out[np.isnan(out)] = 0.0
    return out, reps[:, 0]
