"""CPU: pins of the oracle's dynamic stage solver (oracle/dynamic.py).

No reference test covers the dynamic solver; the pins are properties the
mathematics fixes for a linear elastic body under a load applied at t = 0:
- the average-acceleration (trapezoidal) Newmark rule conserves the total
  energy  V'MV/2 + U'KU/2 - F'U  of an undamped linear system, which is 0
  at rest, at every step and for any step size;
- the initial acceleration solves M A = F;
- one Newton-Raphson iteration per increment (linear problem);
- with mass-proportional damping the energy decreases monotonically.
"""
import numpy as np
import scipy.sparse.linalg as spla

from oracle import HEX8, fem
from oracle.dynamic import dyn_stage_solver


def cantilever():
    X, conn = fem.block_mesh([0, 0, 0], [1.0, 0.1, 0.1], 8, 2, 2, HEX8)
    m = fem.Model(X, conn, HEX8, [("LinearElastic", (200e9, 0.3))])
    bcs = [("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", fem.SurfaceBC(tz=-1e6))]
    return m, bcs


def energy(m, bcs, rho):
    ndofs, nu = m.configure_dofs(bcs)
    K = m.mount_K().tocsc()[:nu, :nu]
    M = m.mount_M(rho).tocsc()[:nu, :nu]
    _, F = m.bc_vals(bcs)
    eq = m.eq
    u = np.zeros(ndofs)
    v = np.zeros(ndofs)
    has = eq >= 0
    u[eq[has]] = m.uvals[has]
    v[eq[has]] = m.vvals[has]
    u, v = u[:nu], v[:nu]
    return 0.5 * v @ (M @ v) + 0.5 * u @ (K @ u) - F[:nu] @ u, 0.5 * u @ (K @ u)


def test_energy_conserved_undamped():
    for nincs in (1, 3, 7):
        m, bcs = cantilever()
        log = []
        ok, msg = dyn_stage_solver(m, bcs, 7850.0, nincs=nincs, tspan=2e-3, log=log)
        assert ok, msg
        assert [it for it, _ in log] == [1] * nincs
        E, strain = energy(m, bcs, 7850.0)
        assert strain > 0
        assert abs(E) <= 1e-8 * strain


def test_initial_acceleration_and_damping():
    m, bcs = cantilever()
    ndofs, nu = m.configure_dofs(bcs)
    M = m.mount_M(7850.0).tocsc()[:nu, :nu]
    _, F = m.bc_vals(bcs)
    a0 = spla.spsolve(M, F[:nu])
    ok, _ = dyn_stage_solver(m, bcs, 7850.0, nincs=1, tspan=1e-12)
    assert ok
    a = np.zeros(ndofs)
    has = m.eq >= 0
    a[m.eq[has]] = m.avals[has]
    # after a vanishing step the acceleration is the initial one
    assert np.linalg.norm(a[:nu] - a0) <= 1e-6 * np.linalg.norm(a0)
    Es = []
    for n in (2, 4, 8):
        m, bcs = cantilever()
        ok, _ = dyn_stage_solver(m, bcs, 7850.0, nincs=n, tspan=n * 1e-3, alpha=500.0)
        assert ok
        Es.append(energy(m, bcs, 7850.0)[0])
    assert Es[0] < 0 and Es[1] < Es[0] and Es[2] < Es[1]
