"""Pins for the oracle's mount_K rule and solve_system! (CPU).

mount_K (src/mech/mech-solver.jl:78-110) drops |Ke_ij| < eps() and lets
sparse() sum duplicates; the oracle sums them in insertion (element) order.
Those two rules are checked on hand-made element matrices whose result
depends on them; solve_system! is checked by its residual, by equilibrium
and against a dense solve.
"""
import numpy as np
import scipy.sparse as sp

from oracle import HEX8, fem


def test_eps_filter_and_insertion_order():
    nd = 2
    eqmap = np.array([[0, 1], [0, 1], [0, 1]], dtype=np.int32)
    Ke = np.zeros((3, nd, nd))
    Ke[0] = [[1e16, 1e-17], [0.0, 2.0]]
    Ke[1] = [[1.0, 3e-16], [-0.0, 1e-300]]
    Ke[2] = [[-1e16, 0.0], [0.0, 5.0]]
    K = fem.csr_from_Ke(Ke, eqmap, 2)
    # (1e16 + 1) + (-1e16) == 0 in insertion order (another order would give 1)
    assert K[0, 0] == 0.0
    # 1e-17 dropped, 3e-16 kept; (1,0) only had zeros -> not in the pattern
    assert K[0, 1] == 3e-16
    dense_pattern = np.zeros((2, 2), dtype=bool)
    dense_pattern[K.nonzero()] = True
    assert K.nnz == 3 and (1, 0) not in set(zip(*K.nonzero()))
    # an entry summing to zero keeps its slot (numerical zero, like sparse())
    assert K.indptr[1] - K.indptr[0] == 2
    assert K[1, 1] == 7.0


def cantilever(nx=20, ny=4, nz=4):
    X, C = fem.block_mesh((0, 0, 0), (1.0, 0.1, 0.1), nx, ny, nz, HEX8)
    m = fem.Model(X, C, HEX8, [("LinearElastic", (200e9, 0.3))])
    bcs = [("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", fem.SurfaceBC(tz=-1e5))]
    return m, bcs


def test_hex8_pattern_is_topological_cfg1_sizes():
    m, bcs = cantilever()
    ndofs, nu = m.configure_dofs(bcs)
    assert (ndofs, nu, m.nnodes, m.ne) == (1575, 1500, 525, 320)
    K = m.mount_K()
    # topological pattern: every (dof, dof) pair of nodes sharing an element
    pairs = set()
    for el in m.conn:
        for a in el:
            for b in el:
                pairs.add((int(a), int(b)))
    assert K.nnz == 9 * len(pairs) == 92781
    assert K[:nu, :nu].nnz == 88218


def test_solve_system_residual_equilibrium_and_dense():
    m, bcs = cantilever(10, 2, 2)
    ndofs, nu = m.configure_dofs(bcs)
    U, F = m.bc_vals(bcs)
    assert abs(F.sum() + 1e5 * 0.01) < 1e-9  # -1 kN resultant
    K = m.mount_K()
    U1, F1 = U.copy(), F.copy()
    assert fem.solve_system(K, U1, F1, nu) is None
    Kd = K.toarray()
    ud = np.linalg.solve(Kd[:nu, :nu], F[:nu])
    assert np.abs(U1[:nu] - ud).max() < 1e-10 * np.abs(ud).max()  # two direct solvers, kappa ~ 1e6
    r = Kd[:nu, :nu] @ U1[:nu] - F[:nu]
    # backward error relative to |K| |U| (the scale a stable direct solve is bound by)
    assert np.abs(r).max() < 1e-10 * np.abs(Kd[:nu, :nu]).sum(axis=1).max() * np.abs(U1).max()
    # reactions: F2 = K21 U1 + K22 U2 balance the applied load
    assert abs(F1[nu:].sum() + F[:nu].sum()) < 1e-10 * 1e3


def test_solve_system_singular_reports_failure():
    # an exactly singular K11 (no surviving entries) makes the factorisation
    # raise; solve_system! turns that into a failure status (src/solver.jl:58-65)
    Ke = np.zeros((1, 2, 2))
    K = fem.csr_from_Ke(Ke, np.array([[0, 1]], dtype=np.int32), 2)
    U, F = np.zeros(2), np.ones(2)
    msg = fem.solve_system(K, U, F, 2)
    assert msg is not None and "solve_system!" in msg
