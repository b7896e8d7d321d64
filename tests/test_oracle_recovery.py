"""CPU: pins of the oracle's output post-processing (oracle/recovery.py).

No reference test covers ip_state_vals or nodal_patch_recovery; the pins are
what the mathematics fixes:
- principal stresses: the invariants of the 3x3 tensor (trace, second
  invariant, determinant) and textbook states (uniaxial, pure shear,
  hydrostatic);
- von Mises: sqrt(3 J2) = sqrt(3/2 s:s) with s the deviator;
- patch recovery: exact reproduction of fields in the regression space
  {1, x, y, z, xy, yz, xz}; on a single Hex8 the fit of xyz is xyz minus the
  centred product (the one 2x2x2-grid monomial outside the space); the number
  of patches reaching each node, counted by hand on small blocks.
"""
import numpy as np
import pytest

from oracle import HEX8, HEX20, TET10, fem, recovery as R

SR2 = np.sqrt(2.0)


def mandel(T):
    """Mandel vector (xx yy zz yz xz xy, shear * sqrt 2) of symmetric 3x3 tensors."""
    return np.stack([T[:, 0, 0], T[:, 1, 1], T[:, 2, 2], SR2 * T[:, 1, 2], SR2 * T[:, 0, 2], SR2 * T[:, 0, 1]], 1)


def states(T, E=None, epa=None):
    n = len(T)
    st = np.zeros((n, 14))
    st[:, 0:6] = mandel(T)
    if E is not None:
        st[:, 6:12] = mandel(E)
    if epa is not None:
        st[:, 12] = epa
    return st


def test_ip_values_textbook_states():
    s, tau, p = 3.0e6, 2.0e6, -5.0e5
    T = np.zeros((3, 3, 3))
    T[0, 0, 0] = s                       # uniaxial
    T[1, 0, 1] = T[1, 1, 0] = tau        # pure shear
    T[2] = p * np.eye(3)                 # hydrostatic
    V = R.ip_values(states(T), np.full(3, 2))
    assert np.allclose(V[0, 6], s) and np.allclose(V[0, 7:10], [s, 0, 0])
    assert np.allclose(V[1, 5], tau) and np.allclose(V[1, 6], np.sqrt(3) * tau)
    assert np.allclose(V[1, 7:10], [tau, 0, -tau])
    assert abs(V[2, 6]) < 1e-9 and np.allclose(V[2, 7:10], p) and np.allclose(V[2, 17], 3 * p)
    assert abs(V[2, 18]) < 1e-9


def test_ip_values_invariants_and_columns():
    rng = np.random.default_rng(3)
    n = 200
    A = rng.standard_normal((n, 3, 3)) * 1e6
    T = A + A.transpose(0, 2, 1)
    A = rng.standard_normal((n, 3, 3)) * 1e-3
    Eps = A + A.transpose(0, 2, 1)
    epa = rng.uniform(0, 0.01, n)
    kinds = np.arange(n) % 3
    V = R.ip_values(states(T, Eps, epa), kinds)
    # stress and strain components, shear without the Mandel factor
    for c, (i, j) in enumerate([(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]):
        assert np.allclose(V[:, c], T[:, i, j], rtol=1e-14, atol=0)
        assert np.allclose(V[:, 10 + c], Eps[:, i, j], rtol=1e-14, atol=0)
    # principal stresses: descending, and the three invariants of T
    L = V[:, 7:10]
    assert np.all(L[:, 0] >= L[:, 1]) and np.all(L[:, 1] >= L[:, 2])
    scale = np.abs(T).max()
    tr = np.trace(T, axis1=1, axis2=2)
    i2 = 0.5 * (tr ** 2 - np.einsum("nij,nji->n", T, T))
    assert np.allclose(L.sum(1), tr, atol=1e-9 * scale)
    assert np.allclose(L[:, 0] * L[:, 1] + L[:, 1] * L[:, 2] + L[:, 0] * L[:, 2], i2, atol=1e-9 * scale ** 2)
    assert np.allclose(L.prod(1), np.linalg.det(T), atol=1e-9 * scale ** 3)
    # von Mises from the deviator
    S = T - tr[:, None, None] / 3 * np.eye(3)
    svm = np.sqrt(1.5 * np.einsum("nij,nij->n", S, S))
    assert np.allclose(V[:, 6], svm, rtol=1e-12)
    # material extras: LE none, VM ep, DP epa, j1, srj2d
    le, vm, dp = kinds == 0, kinds == 1, kinds == 2
    assert np.all(V[le, 16:] == 0) and np.all(V[vm, 17:] == 0)
    assert np.array_equal(V[vm, 16], epa[vm]) and np.array_equal(V[dp, 16], epa[dp])
    assert np.allclose(V[dp, 17], tr[dp]) and np.allclose(V[dp, 18], svm[dp] / np.sqrt(3), rtol=1e-12)


def poly(P):
    x, y, z = P[:, 0], P[:, 1], P[:, 2]
    return 1.5 + 2 * x - 3 * y + 0.5 * z + x * y - 0.7 * y * z + 2 * x * z


@pytest.mark.parametrize("etype,n,nips", [(HEX8, (3, 2, 2), 0), (HEX20, (2, 2, 2), 27), (HEX20, (2, 2, 1), 8),
                                          (TET10, (2, 2, 2), 0), (TET10, (3, 1, 1), 0)])
def test_recovery_reproduces_the_regression_space(etype, n, nips):
    X, conn = fem.block_mesh([0, 0, 0], [2.0, 1.2, 0.9], *n, etype)
    XI = R.ip_coords(X, conn, etype, nips)
    W = np.stack([poly(XI), np.ones(len(XI)), XI[:, 2]], 1)
    V, reps = R.nodal_patch_recovery(X, conn, etype, W, nips)
    assert np.all(reps > 0)
    assert np.allclose(V[:, 0], poly(X), rtol=0, atol=1e-12 * np.abs(poly(X)).max())
    assert np.allclose(V[:, 1], 1.0, atol=1e-12)
    assert np.allclose(V[:, 2], X[:, 2], atol=1e-12)


def test_recovery_single_hex8_xyz():
    X, conn = fem.block_mesh([0, 0, 0], [1.0, 1.0, 1.0], 1, 1, 1, HEX8)
    XI = R.ip_coords(X, conn, HEX8)
    g = 0.577350269189626
    assert np.allclose(np.sort(np.unique(np.round(XI[:, 0], 12))), [0.5 - g / 2, 0.5 + g / 2])
    V, reps = R.nodal_patch_recovery(X, conn, HEX8, XI.prod(1)[:, None])
    assert np.all(reps == 8)            # all 8 corners are orphans, each takes its 1-element boundary patch
    c = X - 0.5
    assert np.allclose(V[:, 0], X.prod(1) - c.prod(1), atol=1e-14)


def test_recovery_patch_counts():
    # 3x3x3 Hex8: the 8 inner corner nodes own the patches; each covers the
    # 3x3x3 nodes around it, so a node with grid indices (i, j, k) is reached
    # by c(i) c(j) c(k) patches, c = (1, 2, 2, 1)
    X, conn = fem.block_mesh([0, 0, 0], [1.0, 1.0, 1.0], 3, 3, 3, HEX8)
    _, reps = R.nodal_patch_recovery(X, conn, HEX8, np.zeros((len(conn) * 8, 1)))
    idx = np.rint(X * 3).astype(int)
    c = np.array([1, 2, 2, 1])
    assert np.array_equal(reps, c[idx[:, 0]] * c[idx[:, 1]] * c[idx[:, 2]])
    # 3x1x1 Hex8: every node is on the boundary; no patch has 3 elements, so
    # the corners at x = 1/3 and 2/3 take their 2-element patches, which reach
    # the end layers 4 times and the inner layers 8 times
    X, conn = fem.block_mesh([0, 0, 0], [1.0, 1.0, 1.0], 3, 1, 1, HEX8)
    P = R.patches(conn, HEX8, len(X))
    assert sorted(len(p) for p in P if p) == [2] * 8
    _, reps = R.nodal_patch_recovery(X, conn, HEX8, np.zeros((len(conn) * 8, 1)))
    layer = np.rint(X[:, 0] * 3).astype(int)
    assert np.array_equal(reps, np.array([4, 8, 8, 4])[layer])


@pytest.mark.parametrize("etype,n,nips", [(HEX8, (4, 3, 2), 0), (TET10, (3, 2, 2), 0), (HEX20, (3, 2, 2), 8)])
def test_subpatches_of_fields_some_elements_carry(etype, n, nips):
    """A field carried only by the elements with x-centre > 1 (fe-model.jl:
    614-680): fitted on the subpatch of those elements, it reproduces a
    linear field exactly at every node that only full-rank subpatches reach
    (a rank-deficient one takes pinv's minimum-norm fit), and is 0 at the
    nodes no subpatch reaches; a field every element carries is unchanged by
    the split.  Fitting the whole patch (zeros included) would break both."""
    X, conn = fem.block_mesh([0, 0, 0], [2.0, 1.2, 0.9], *n, etype)
    XI = R.ip_coords(X, conn, etype, nips)
    nq = len(XI) // len(conn)
    right = X[conn].mean(1)[:, 0] > 1.0
    lin = lambda P: 0.5 + 2 * P[:, 0] - P[:, 1] + 3 * P[:, 2]  # noqa: E731
    carry = np.repeat(right, nq)
    W = np.stack([poly(XI), np.where(carry, lin(XI), 0.0)], 1)
    has = np.stack([np.ones(len(conn), bool), right], 1)
    V, reps = R.nodal_patch_recovery(X, conn, etype, W, nips, has=has)
    V0, reps0 = R.nodal_patch_recovery(X, conn, etype, W[:, :1], nips)
    assert np.array_equal(reps, reps0) and np.array_equal(V[:, 0], V0[:, 0])
    # the nodes of right elements in patches that hold one; those reached by
    # a rank-deficient subpatch
    reached = np.zeros(len(X), bool)
    deficient = np.zeros(len(X), bool)
    for patch in R.patches(conn, etype, len(X)):
        sub = [e for e in patch if right[e]]
        if not sub:
            continue
        ips = np.concatenate([np.arange(e * nq, (e + 1) * nq) for e in sub])
        nt = 7 if len(ips) >= 7 else 4 if len(ips) >= 4 else 1
        nodes = np.unique(conn[sub])
        reached[nodes] = True
        if np.linalg.matrix_rank(R.reg_terms(XI[ips], nt)) < nt:
            deficient[nodes] = True
    full = reached & ~deficient
    assert full.sum() >= 10 and not reached.all()
    assert np.allclose(V[full, 1], lin(X[full]), rtol=0, atol=1e-11 * np.abs(lin(X)).max())
    assert np.all(V[~reached, 1] == 0.0)
