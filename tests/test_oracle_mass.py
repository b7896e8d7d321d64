"""CPU: pins of the oracle's element mass matrices and mount_M
(oracle/amaru_oracle.c ora_elem_mass, oracle/fem.py Model.mount_M).

No reference test covers elem_mass; the pins are closed forms:
- the unit-cube HEX8 consistent mass, the tensor product of the 1-D linear
  element's h/6 [2 1; 1 2] (2x2x2 Gauss is exact for it);
- total mass: sum_ab N_a N_b = 1, so the entries of one direction sum to
  rho times the volume (exactly integrated for undistorted elements, and for
  HEX8 with moved interior nodes, whose total volume is the box's);
- N'N couples equal directions only: mount_M keeps no cross-direction entry.
"""
import numpy as np
import pytest

from oracle import HEX8, HEX20, TET10, fem


def model(et, n, nips=0, hi=(2.0, 1.0, 0.5), distort=0.0, seed=0):
    X, conn = fem.block_mesh([0, 0, 0], list(hi), *n, et)
    if distort:
        inner = np.all((X > 1e-9) & (X < np.array(hi) - 1e-9), axis=1)
        X = X.copy()
        X[inner] += np.random.default_rng(seed).uniform(-distort, distort, (inner.sum(), 3))
    m = fem.Model(X, conn, et, [("LinearElastic", (1e9, 0.3))], nips=nips)
    m.configure_dofs([("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0))])
    return m


def test_unit_cube_hex8_closed_form():
    m = model(HEX8, (1, 1, 1), hi=(1.0, 1.0, 1.0))
    rho = 2500.0
    Me = m.elem_mass(rho)[0]
    nodes = m.X[m.conn[0]]
    for a in range(8):
        for b in range(8):
            w = np.prod(np.where(nodes[a] == nodes[b], 2.0, 1.0)) / 216.0
            blk = Me[3 * a:3 * a + 3, 3 * b:3 * b + 3]
            assert np.allclose(blk, rho * w * np.eye(3), rtol=1e-14, atol=0)


@pytest.mark.parametrize("et,n,nips,distort", [(HEX8, (3, 2, 2), 0, 0.05), (HEX20, (2, 2, 1), 27, 0.0),
                                               (HEX20, (2, 2, 1), 8, 0.0), (TET10, (2, 2, 2), 0, 0.0)])
def test_total_mass_and_structure(et, n, nips, distort):
    m = model(et, n, nips, distort=distort)
    rho = np.linspace(1000.0, 3000.0, m.ne)
    Me = m.elem_mass(rho)
    assert np.array_equal(Me, Me.transpose(0, 2, 1))
    V = 2.0 * 1.0 * 0.5
    if distort == 0.0:
        vol = np.full(m.ne, V / m.ne)
    else:   # per-element volumes from the exact HEX8 rule: use rho = 1
        vol = m.elem_mass(1.0)[:, 0::3, 0::3].sum(axis=(1, 2))
        assert np.isclose(vol.sum(), V, rtol=1e-13)
    for d in range(3):
        assert np.allclose(Me[:, d::3, d::3].sum(axis=(1, 2)), rho * vol, rtol=1e-13)
    M = m.mount_M(rho).tocoo()
    comp = np.full(m.ndofs, -1)
    for n_ in range(m.nnodes):
        for d in range(3):
            if m.eq[n_, d] >= 0:
                comp[m.eq[n_, d]] = d
    assert np.all(comp[M.row] == comp[M.col])
    assert np.isclose(M.sum(), 3 * (rho * vol).sum(), rtol=1e-13)
