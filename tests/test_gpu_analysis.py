"""GPU end to end: MechAnalysis / addstage / solve on the device against the
reference's own known answers and against the oracle's stage solver."""
import numpy as np
import pytest

from oracle import HEX8 as O8, HEX20 as O20, TET10 as O10, fem

pytestmark = pytest.mark.gpu


def base_hex8(am):
    return [("x==0 && y==0 && z==0", am.NodeBC(ux=0, uy=0)), ("x==1 && y==0 && z==0", am.NodeBC(uy=0)),
            ("x==0 && y==1 && z==0", am.NodeBC(ux=0)), ("z==0", am.NodeBC(uz=0))]


@pytest.mark.parametrize("case", range(4))
def test_elastic_hex8_reference(amaru, pins, case):
    p = pins["elastic_hex8"]
    loads = [("z==1", amaru.NodeBC(fz=1)), ("y==1 && z==1", amaru.EdgeBC(qy=2)),
             ("x==1", amaru.SurfaceBC(tx="3*z")), ("x>=0", amaru.BodyC(wz=-1))]
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 1]], nx=1, ny=1, nz=1, cellshape="HEX8", tag="solid"))
    model = amaru.FEModel(mesh, [("solid", amaru.MechSolid, amaru.LinearElastic(E=p["E"], nu=p["nu"]))])
    ana = amaru.MechAnalysis(model)
    amaru.addstage(ana, base_hex8(amaru) + [loads[case]], nouts=1)
    ok, msg = amaru.solve(ana)
    assert ok, msg
    assert np.allclose(model.uvals[:, 2], p["uz"][case], atol=p["atol"])


@pytest.mark.parametrize("shape", ["TET10", "HEX8", "HEX20"])
def test_elastic_elems_3d_reference(amaru, pins, shape):
    p = pins["elastic_elems_3d"]
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 1]], nx=2, ny=2, nz=2, cellshape=shape))
    model = amaru.FEModel(mesh, [("all", amaru.LinearElastic(E=p["E"], nu=p["nu"]))])
    ana = amaru.MechAnalysis(model)
    amaru.addstage(ana, [("z==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("x==0 || x==1", amaru.SurfaceBC(ux=0)),
                         ("z==1", amaru.SurfaceBC(tz=p["tz"]))])
    ok, msg = amaru.solve(ana)
    assert ok, msg
    top = np.nonzero(mesh.coords[:, 2] == 1)[0][0]
    assert np.allclose(model.uvals[top, 1:], p["uy_uz"], atol=p["atol"])


def test_vm_3d_reference_and_oracle(amaru, pins):
    p = pins["vm_3d"]
    th = 0.05
    mesh = amaru.Mesh(amaru.Block([[0, 0, -0.05], [0.05, 1.0, 0.05]], nx=1, ny=50, nz=2, cellshape="HEX20"))
    model = amaru.FEModel(mesh, [(":bulks", amaru.MechSolid, amaru.VonMises(E=p["E"], nu=p["nu"], fy=p["fy"], H=p["H"]))])
    ana = amaru.MechAnalysis(model)
    bcs = [("y==0", amaru.NodeBC(uy=0)), ("y==0 && z==0", amaru.NodeBC(uz=0)),
           (f"x=={th}/2 && y==0 && z==0", amaru.NodeBC(ux=0)), (f"x=={th}/2 && y==1 && z==0", amaru.NodeBC(uz=p["uz"]))]
    X = mesh.coords
    node = np.nonzero((X[:, 0] == th / 2) & (X[:, 1] == 1) & (X[:, 2] == 0))[0][0]
    rec = []
    ana.loggers.append(lambda m, T: rec.append(m.fvals[node, 2]))
    amaru.addstage(ana, bcs, nincs=20, nouts=1)
    ok, msg = amaru.solve(ana, autoinc=True)
    assert ok, msg
    assert rec[-1] == pytest.approx(p["fz_end"], abs=p["atol"])
    # the oracle's stage solver on the same problem
    om = fem.Model(mesh.coords, mesh.conn, O20, [("VonMises", (p["E"], p["nu"], p["fy"], p["H"]))])
    obcs = [("y==0", fem.NodeBC(uy=0)), ("y==0 && z==0", fem.NodeBC(uz=0)),
            (f"x=={th}/2 && y==0 && z==0", fem.NodeBC(ux=0)), (f"x=={th}/2 && y==1 && z==0", fem.NodeBC(uz=p["uz"]))]
    orec = []
    ok, msg = fem.solve_stage(om, obcs, nincs=20, nouts=1, autoinc=True,
                              on_increment=lambda mm, T: orec.append(mm.fvals[node, 2]))
    assert ok
    assert len(rec) == len(orec)
    assert rec[-1] == pytest.approx(orec[-1], rel=1e-6)
    assert np.allclose(model.uvals, om.uvals, rtol=0, atol=1e-8 * np.abs(om.uvals).max())


def test_dp_two_stages_reference(amaru, pins):
    p = pins["dp"]
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 0.5]], nx=2, ny=2, nz=2, cellshape="HEX8", tag="solids"))
    model = amaru.FEModel(mesh, [("solids", amaru.MechSolid,
                                  amaru.DruckerPrager(E=p["E"], nu=p["nu"], alpha=p["alpha"], kappa=p["kappa"]))])
    ana = amaru.MechAnalysis(model)
    bcs = [("z==0.0", amaru.NodeBC(ux=0, uy=0, uz=0)), ("z==0.5", amaru.NodeBC(uz=p["uz1"])),
           ("x==0 || x==1.0", amaru.NodeBC(ux=0, uy=0)), ("y==0 || y==1.0", amaru.NodeBC(ux=0, uy=0))]
    amaru.addstage(ana, bcs, nincs=10)
    bcs2 = list(bcs)
    bcs2[1] = ("z==0.5", amaru.NodeBC(uz=p["uz2"]))
    amaru.addstage(ana, bcs2, nincs=10)
    ok, msg = amaru.solve(ana, tol=1e-2, autoinc=True)
    assert ok, msg
    assert model.state[:, 12].max() > 0
