"""Pins for the oracle's element stiffness (CPU).

The reference has no test of K itself; these checks fix it by mathematics:
  * unit-cube HEX8 K against a closed form built from 1-D integrals in Lamé
    form (independent of the Mandel B / D route the oracle takes),
  * six rigid-body null vectors and symmetry for every element type,
  * the patch test (a linear displacement field leaves interior nodes in
    equilibrium) on distorted meshes.
"""
import numpy as np
import pytest

from oracle import HEX8, HEX20, TET10, fem

HEX20_EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
TET10_EDGES = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def hex8_cube_closed_form(E, nu):
    lam, mu = lame(E, nu)
    M1 = np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]])
    S1 = np.array([[1.0, -1.0], [-1.0, 1.0]])
    sg = [-1.0, 1.0]
    C1 = np.array([[sg[i] * 0.5 for j in range(2)] for i in range(2)])  # int phi_i' phi_j
    nat = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]
    idx = [tuple((c + 1) // 2 for c in n) for n in nat]

    def G(a, p, b, q):
        v = 1.0
        for ax in range(3):
            ia, ib = idx[a][ax], idx[b][ax]
            if ax == p and ax == q:
                v *= S1[ia, ib]
            elif ax == p:
                v *= C1[ia, ib]
            elif ax == q:
                v *= C1[ib, ia]
            else:
                v *= M1[ia, ib]
        return v

    K = np.zeros((24, 24))
    for a in range(8):
        for b in range(8):
            for i in range(3):
                for j in range(3):
                    v = lam * G(a, i, b, j) + mu * G(a, j, b, i)
                    if i == j:
                        v += mu * sum(G(a, p, b, p) for p in range(3))
                    K[3 * a + i, 3 * b + j] = v
    return K


def test_hex8_unit_cube_closed_form():
    X, C = fem.block_mesh((0, 0, 0), (1, 1, 1), 1, 1, 1, HEX8)
    m = fem.Model(X, C, HEX8, [("LinearElastic", (200e9, 0.3))])
    Ke = m.elem_stiffness()[0]
    # block_mesh orders the cube's nodes in HEX8 order for a single cell
    Kc = hex8_cube_closed_form(200e9, 0.3)
    assert np.max(np.abs(Ke - Kc)) / np.max(np.abs(Kc)) < 1e-14


def distorted(et, nips, seed=0):
    rng = np.random.default_rng(seed)
    X, C = fem.block_mesh((0, 0, 0), (1, 1, 1), 2, 2, 2, et)
    # move only cell-grid vertices that are interior, then put edge nodes
    # back on the straight edge midpoints
    corner = [c for c in range(X.shape[0]) if np.all(np.isclose(X[c] * 2, np.round(X[c] * 2)))]
    for c in corner:
        if np.all((X[c] > 0) & (X[c] < 1)):
            X[c] += rng.uniform(-0.12, 0.12, 3)
    if et == HEX20:
        for el in C:
            for k, (a, b) in enumerate(HEX20_EDGES):
                X[el[8 + k]] = 0.5 * (X[el[a]] + X[el[b]])
    if et == TET10:
        for el in C:
            for k, (a, b) in enumerate(TET10_EDGES):
                X[el[4 + k]] = 0.5 * (X[el[a]] + X[el[b]])
    return fem.Model(X, C, et, [("LinearElastic", (30e9, 0.2))], nips=nips)


def rigid_modes(X):
    n = X.shape[0]
    modes = []
    for d in range(3):
        u = np.zeros((n, 3))
        u[:, d] = 1.0
        modes.append(u.ravel())
    for w in np.eye(3):
        modes.append(np.cross(w, X).ravel())
    return modes


@pytest.mark.parametrize("et,nips", [(HEX8, 8), (HEX20, 8), (HEX20, 27), (TET10, 4)])
def test_rigid_body_and_symmetry(et, nips):
    m = distorted(et, nips, seed=1)
    Ke = m.elem_stiffness()
    scale = np.abs(Ke).max()
    for e in range(m.ne):
        Xe = m.X[m.conn[e]]
        for u in rigid_modes(Xe):
            assert np.abs(Ke[e] @ u).max() < 1e-12 * scale * max(1.0, np.abs(u).max())
        assert np.abs(Ke[e] - Ke[e].T).max() < 1e-13 * scale


@pytest.mark.parametrize("et,nips", [(HEX8, 8), (HEX20, 27), (TET10, 4)])
def test_patch_test(et, nips):
    m = distorted(et, nips, seed=2)
    m.configure_dofs([])
    K = m.mount_K()
    A = np.array([[1e-3, 2e-4, -3e-4], [5e-4, -2e-3, 1e-4], [-1e-4, 3e-4, 7e-4]])
    u_nodes = m.X @ A.T + np.array([1e-3, -2e-3, 5e-4])
    u = np.zeros(m.ndofs)
    for n in range(m.nnodes):
        for d in range(3):
            u[m.eq[n, d]] = u_nodes[n, d]
    f = K @ u
    interior = [n for n in range(m.nnodes) if np.all((m.X[n] > 1e-9) & (m.X[n] < 1 - 1e-9))]
    assert interior
    fi = np.array([f[m.eq[n, d]] for n in interior for d in range(3)])
    assert np.abs(fi).max() < 1e-10 * np.abs(f).max()
    # constant strain reproduced at every IP: stresses equal De * eps
    dFin, st, msg = m.update_state(u, m.state)
    assert msg is None
    eps = 0.5 * (A + A.T)
    epsM = np.array([eps[0, 0], eps[1, 1], eps[2, 2], np.sqrt(2) * eps[1, 2], np.sqrt(2) * eps[0, 2],
                     np.sqrt(2) * eps[0, 1]])
    lam, mu = lame(30e9, 0.2)
    sig = 2 * mu * epsM + lam * np.trace(eps) * np.array([1, 1, 1, 0, 0, 0])
    assert np.abs(st[:, :6] - sig).max() < 1e-9 * np.abs(sig).max()
    # for a linear material the internal forces equal K u
    assert np.abs(dFin - f).max() < 1e-9 * np.abs(f).max()
