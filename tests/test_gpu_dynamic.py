"""GPU: the dynamic stage solver on the device (amaru_amd.dynamic) against the
oracle's restatement of dyn_stage_solver! (oracle/dynamic.py).

Bar: the same Newton-Raphson iterations per increment; displacements,
velocities and accelerations within 1e-8 relative of the oracle's (PCG to
1e-12 against SuperLU), with and without Rayleigh damping.
"""
import numpy as np
import pytest

from oracle import HEX8 as O8, HEX20 as O20, TET10 as O10, fem
from oracle.dynamic import dyn_stage_solver

pytestmark = pytest.mark.gpu

OSHAPE = {"HEX8": O8, "HEX20": O20, "TET10": O10}


@pytest.mark.parametrize("shape,nips,alpha,beta", [("HEX8", 0, 0.0, 0.0), ("TET10", 0, 40.0, 2e-5),
                                                   ("HEX20", 27, 0.0, 1e-5)])
def test_dynamic_stage_matches_oracle(amaru, shape, nips, alpha, beta):
    from amaru_amd import dynamic as D
    n = (8, 2, 2) if shape != "HEX20" else (4, 1, 1)
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1.0, 0.1, 0.1]], nx=n[0], ny=n[1], nz=n[2], cellshape=shape))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid(rho=7850.0), amaru.LinearElastic(E=200e9, nu=0.3))],
                          nips=nips)
    ana = D.DynamicAnalysis(model)
    bcs = [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", amaru.SurfaceBC(tz="-1e6*(1 + y)"))]
    D.addstage(ana, bcs, nincs=6, tspan=3e-3)
    ok, msg = D.dyn_solve(ana, alpha=alpha, beta=beta)
    assert ok, msg
    om = fem.Model(mesh.coords, mesh.conn, OSHAPE[shape], [("LinearElastic", (200e9, 0.3))], nips=nips)
    obcs = [("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", fem.SurfaceBC(tz="-1e6*(1 + y)"))]
    log = []
    ok, msg = dyn_stage_solver(om, obcs, 7850.0, nincs=6, tspan=3e-3, alpha=alpha, beta=beta, log=log)
    assert ok, msg
    assert [it for _, it, _ in ana.log] == [it for it, _ in log]
    for name in ("uvals", "vvals", "avals"):
        a, b = getattr(model, name), getattr(om, name)
        assert np.linalg.norm(a - b) <= 1e-8 * np.linalg.norm(b), name
    assert np.linalg.norm(model.uvals) > 0
    ana.dev.close()


def test_unsupported_forms_are_refused(amaru):
    from amaru_amd import dynamic as D
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1.0, 0.1, 0.1]], nx=2, ny=1, nz=1, cellshape="HEX8"))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid(rho=7850.0), amaru.LinearElastic(E=200e9, nu=0.3))])
    ana = D.DynamicAnalysis(model)
    D.addstage(ana, [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0))], nincs=2, tspan=1e-3)
    with pytest.raises(NotImplementedError):
        D.dyn_solve(ana, sism=True)


def test_plastic_auto_increment_matches_oracle(amaru):
    """Von Mises cantilever under a step load that the fixed increments cannot
    carry: auto-increment cuts failed increments; the device takes the
    oracle's path (every NR iteration, every cut) and ends in its state."""
    from amaru_amd import dynamic as D
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1.0, 0.1, 0.1]], nx=8, ny=2, nz=2, cellshape="HEX8"))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid(rho=7850.0), amaru.VonMises(E=200e9, nu=0.3, fy=250e6))])
    ana = D.DynamicAnalysis(model)
    D.addstage(ana, [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", amaru.SurfaceBC(tz=-1.2e7))], nincs=8,
               tspan=4e-3)
    ok, msg = D.dyn_solve(ana, alpha=20.0, maxits=10, autoinc=True)
    assert ok, msg
    om = fem.Model(mesh.coords, mesh.conn, O8, [("VonMises", (200e9, 0.3, 250e6, 0.0))])
    log = []
    ok, msg = dyn_stage_solver(om, [("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", fem.SurfaceBC(tz=-1.2e7))],
                               7850.0, nincs=8, tspan=4e-3, alpha=20.0, maxits=10, autoinc=True, log=log)
    assert ok, msg
    assert [it for _, it, _ in ana.log] == [it for it, _ in log]
    assert len(log) > 8 * 3   # cuts happened
    st = ana.dev.state_get(committed=True)
    assert np.array_equal(st[:, 12] > 0, om.state[:, 12] > 0) and (st[:, 12] > 0).any()
    for name in ("uvals", "vvals", "avals"):
        a, b = getattr(model, name), getattr(om, name)
        assert np.linalg.norm(a - b) <= 1e-8 * np.linalg.norm(b), name
    ana.dev.close()
