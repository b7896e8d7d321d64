"""CPU: host logic of the dynamic stage driver (amaru_amd.dynamic).

A stand-in device backed by the oracle's mount_K / mount_M / update_state
and SuperLU runs the product driver; it must take the oracle driver's
Newton-Raphson path and reach its displacements, velocities and
accelerations (the same arithmetic but for the damping products, which the
driver forms as alpha (M x) + beta (K x) and the oracle as (alpha M + beta K) x).
"""
import numpy as np
import pytest

from oracle import HEX8 as O8, TET10 as O10, fem
from oracle.dynamic import dyn_stage_solver

OSHAPE = {"HEX8": O8, "TET10": O10}


def oracle_mat(mat):
    if type(mat).__name__ == "VonMises":
        return ("VonMises", (mat.E, mat.nu, mat.fy, mat.H))
    return ("LinearElastic", (mat.E, mat.nu))


class OracleDev:
    def __init__(self, model, device=0):
        self.model = model
        self.om = fem.Model(model.coords, model.conn, OSHAPE[model.shape], [oracle_mat(model.mats[0])])
        self.om.eq = model.eq
        self.om.eqmap = np.ascontiguousarray(model.eq[model.conn].reshape(model.ne, -1))
        self.om.ndofs = model.ndofs
        self.st_com = np.zeros_like(self.om.state)
        self.st_cur = self.st_com.copy()

    def set_dofs(self):
        pass

    def assemble(self):
        self.K = self.om.mount_K(self.st_cur).tocsc()

    def assemble_M(self):
        self.M = self.om.mount_M(self.model.rho).tocsc()

    def combine_KM(self, a, b):
        self.K0 = self.K
        self.K = (a * self.K + b * self.M).tocsc()

    def matvec(self, which, x):
        return {"K": self.K, "M": self.M, "K0": self.K0}[which] @ x

    def solve_system(self, U, F, **kw):
        assert fem.solve_system(self.K, U, F, self.model.nu) is None

    def update_state(self, dUt):
        dFin, self.st_cur, msg = self.om.update_state(dUt, self.st_com)
        assert msg is None
        return dFin

    def commit(self):
        self.st_com = self.st_cur.copy()

    def restore(self):
        self.st_cur = self.st_com.copy()


@pytest.mark.parametrize("shape,alpha,beta", [("HEX8", 0.0, 0.0), ("TET10", 40.0, 2e-5)])
def test_driver_follows_the_oracle(amaru, monkeypatch, shape, alpha, beta):
    from amaru_amd import dynamic as D
    monkeypatch.setattr(D, "DeviceModel", OracleDev)
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1.0, 0.1, 0.1]], nx=8, ny=2, nz=2, cellshape=shape))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid(rho=7850.0), amaru.LinearElastic(E=200e9, nu=0.3))])
    ana = D.DynamicAnalysis(model)
    bcs = [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", amaru.SurfaceBC(tz="-1e6*(1 + y)"))]
    D.addstage(ana, bcs, nincs=6, tspan=3e-3)
    ok, msg = D.dyn_solve(ana, alpha=alpha, beta=beta)
    assert ok, msg
    om = fem.Model(mesh.coords, mesh.conn, OSHAPE[shape], [("LinearElastic", (200e9, 0.3))])
    obcs = [("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", fem.SurfaceBC(tz="-1e6*(1 + y)"))]
    log = []
    ok, msg = dyn_stage_solver(om, obcs, 7850.0, nincs=6, tspan=3e-3, alpha=alpha, beta=beta, log=log)
    assert ok, msg
    assert [it for _, it, _ in ana.log] == [it for it, _ in log]
    for name in ("uvals", "vvals", "avals"):
        a, b = getattr(model, name), getattr(om, name)
        assert np.linalg.norm(a - b) <= 1e-10 * np.linalg.norm(b), name


def test_auto_increment_follows_the_oracle(amaru, monkeypatch):
    """A von Mises cantilever the fixed increments cannot carry: the driver's
    cuts and retries follow the oracle's."""
    from amaru_amd import dynamic as D
    monkeypatch.setattr(D, "DeviceModel", OracleDev)
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1.0, 0.1, 0.1]], nx=8, ny=2, nz=2, cellshape="HEX8"))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid(rho=7850.0), amaru.VonMises(E=200e9, nu=0.3, fy=250e6))])
    ana = D.DynamicAnalysis(model)
    bcs = [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", amaru.SurfaceBC(tz=-1.2e7))]
    D.addstage(ana, bcs, nincs=8, tspan=4e-3)
    ok, msg = D.dyn_solve(ana, alpha=20.0, maxits=10, autoinc=True)
    assert ok, msg
    om = fem.Model(mesh.coords, mesh.conn, O8, [("VonMises", (200e9, 0.3, 250e6, 0.0))])
    log = []
    ok, msg = dyn_stage_solver(om, [("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", fem.SurfaceBC(tz=-1.2e7))],
                               7850.0, nincs=8, tspan=4e-3, alpha=20.0, maxits=10, autoinc=True, log=log)
    assert ok, msg
    assert [it for _, it, _ in ana.log] == [it for it, _ in log]
    assert np.linalg.norm(model.uvals - om.uvals) <= 1e-10 * np.linalg.norm(om.uvals)
    # without auto-increment the same problem fails, as the reference's does
    model2 = amaru.FEModel(mesh, [("all", amaru.MechSolid(rho=7850.0), amaru.VonMises(E=200e9, nu=0.3, fy=250e6))])
    ana2 = D.DynamicAnalysis(model2)
    D.addstage(ana2, bcs, nincs=8, tspan=4e-3)
    ok, msg = D.dyn_solve(ana2, alpha=20.0, maxits=10)
    assert not ok and "did not converge" in msg


def test_failed_update_cuts_the_step(amaru, monkeypatch):
    """update_state! failing inside an increment (here: NaN internal forces
    injected at the third update, i.e. the first iteration of the second
    increment) is not an error of the stage: the iteration counts as an
    unbounded residue, the increment is retried from the committed states
    with dT cut by 0.2, and the stage finishes; the oracle driver with the
    same injection takes the same path to the same displacements."""
    from amaru_amd import dynamic as D
    from amaru_amd._lib import AmaruError
    from oracle import dynamic as OD

    calls = {"dev": 0, "ora": 0}

    class FailingDev(OracleDev):
        def update_state(self, dUt):
            calls["dev"] += 1
            if calls["dev"] == 3:
                raise AmaruError(-6, "solve_system!: NaN values in internal forces vector")
            return super().update_state(dUt)

    monkeypatch.setattr(D, "DeviceModel", FailingDev)
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1.0, 0.1, 0.1]], nx=6, ny=2, nz=2, cellshape="HEX8"))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid(rho=7850.0), amaru.LinearElastic(E=200e9, nu=0.3))])
    ana = D.DynamicAnalysis(model)
    bcs = [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", amaru.SurfaceBC(tz=-1e6))]
    D.addstage(ana, bcs, nincs=4, tspan=2e-3)
    ok, msg = D.dyn_solve(ana, autoinc=True)
    assert ok, msg
    # increment 1 (it 1, 2), increment 2 fails at once, then the cut step
    assert ana.log[2][1] == 1 and ana.log[2][2] == float("inf")
    assert all(np.isfinite(r) for _, _, r in ana.log[3:])
    # without auto-increment the failure ends the stage as non-converged
    calls["dev"] = 0
    model2 = amaru.FEModel(mesh, [("all", amaru.MechSolid(rho=7850.0), amaru.LinearElastic(E=200e9, nu=0.3))])
    ana2 = D.DynamicAnalysis(model2)
    D.addstage(ana2, bcs, nincs=4, tspan=2e-3)
    ok, msg = D.dyn_solve(ana2)
    assert not ok and "did not converge" in msg

    om = fem.Model(mesh.coords, mesh.conn, O8, [("LinearElastic", (200e9, 0.3))])
    real = om.update_state

    def failing(dUt, st):
        calls["ora"] += 1
        if calls["ora"] == 3:
            dFin, st2, _ = real(dUt, st)
            return dFin * np.nan, st2, "solve_system!: NaN values in internal forces vector"
        return real(dUt, st)

    om.update_state = failing
    log = []
    ok, msg = OD.dyn_stage_solver(om, [("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", fem.SurfaceBC(tz=-1e6))],
                                  7850.0, nincs=4, tspan=2e-3, autoinc=True, log=log)
    assert ok, msg
    assert [it for _, it, _ in ana.log] == [it for it, _ in log]
    assert np.linalg.norm(model.uvals - om.uvals) <= 1e-10 * np.linalg.norm(om.uvals)
