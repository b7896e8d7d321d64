"""The oracle against the reference's own known-answer tests (CPU).

Values come from tests/golden/reference_pins.json, which re-types them from
the reference's test files (cited there).
"""
import numpy as np
import pytest

from oracle import HEX8, HEX20, TET10, fem

SHAPES = {"HEX8": HEX8, "HEX20": HEX20, "TET10": TET10}


@pytest.mark.parametrize("name", ["HEX8", "HEX20", "TET10"])
def test_structured_node_counts(pins, name):
    X, C = fem.block_mesh((0, 0, 0), (1, 1, 1), 10, 10, 10, SHAPES[name])
    assert X.shape[0] == pins["structured_node_counts"][name]


HEX8_BASE = [("x==0 && y==0 && z==0", fem.NodeBC(ux=0, uy=0)), ("x==1 && y==0 && z==0", fem.NodeBC(uy=0)),
             ("x==0 && y==1 && z==0", fem.NodeBC(ux=0)), ("z==0", fem.NodeBC(uz=0))]
HEX8_CASES = [("z==1", fem.NodeBC(fz=1)), ("y==1 && z==1", fem.EdgeBC(qy=2)),
              ("x==1", fem.SurfaceBC(tx="3*z")), ("x>=0", fem.BodyC(wz=-1))]


@pytest.mark.parametrize("case", range(4))
def test_elastic_hex8(pins, case):
    p = pins["elastic_hex8"]
    X, C = fem.block_mesh((0, 0, 0), (1, 1, 1), 1, 1, 1, HEX8)
    m = fem.Model(X, C, HEX8, [("LinearElastic", (p["E"], p["nu"]))])
    ok, msg = fem.solve_stage(m, HEX8_BASE + [HEX8_CASES[case]], nouts=1)
    assert ok, msg
    assert np.allclose(m.uvals[:, 2], p["uz"][case], atol=p["atol"])
    # reactions balance the applied load (sum of fz over all nodes = 0 within rounding)
    assert abs(m.fvals[:, 2].sum()) < 1e-12


@pytest.mark.parametrize("name", ["TET10", "HEX8", "HEX20"])
def test_elastic_elems_3d(pins, name):
    p = pins["elastic_elems_3d"]
    X, C = fem.block_mesh((0, 0, 0), (1, 1, 1), 2, 2, 2, SHAPES[name])
    m = fem.Model(X, C, SHAPES[name], [("LinearElastic", (p["E"], p["nu"]))])
    bcs = [("z==0", fem.SurfaceBC(ux=0, uy=0, uz=0)), ("x==0 || x==1", fem.SurfaceBC(ux=0)),
           ("z==1", fem.SurfaceBC(tz=p["tz"]))]
    ok, msg = fem.solve_stage(m, bcs)
    assert ok, msg
    top = [i for i in range(m.nnodes) if X[i, 2] == 1][0]
    assert np.allclose(m.uvals[top, 1:], p["uy_uz"], atol=p["atol"])


def test_vm_3d(pins):
    p = pins["vm_3d"]
    th = 0.05
    X, C = fem.block_mesh((0, 0, -0.05), (0.05, 1.0, 0.05), 1, 50, 2, HEX20)
    m = fem.Model(X, C, HEX20, [("VonMises", (p["E"], p["nu"], p["fy"], p["H"]))])
    bcs = [("y==0", fem.NodeBC(uy=0)), ("y==0 && z==0", fem.NodeBC(uz=0)),
           (f"x=={th}/2 && y==0 && z==0", fem.NodeBC(ux=0)),
           (f"x=={th}/2 && y==1 && z==0", fem.NodeBC(uz=p["uz"]))]
    node = [i for i in range(m.nnodes) if X[i, 0] == th / 2 and X[i, 1] == 1 and X[i, 2] == 0][0]
    rec = []
    ok, msg = fem.solve_stage(m, bcs, nincs=20, nouts=1, autoinc=True,
                              on_increment=lambda mm, T: rec.append(mm.fvals[node, 2]))
    assert ok, msg
    assert rec[-1] == pytest.approx(p["fz_end"], abs=p["atol"])


def test_dp_two_stages(pins):
    p = pins["dp"]
    X, C = fem.block_mesh((0, 0, 0), (1, 1, 0.5), 2, 2, 2, HEX8)
    m = fem.Model(X, C, HEX8, [("DruckerPrager", (p["E"], p["nu"], p["alpha"], p["kappa"], 0.0))])
    bcs = [("z==0.0", fem.NodeBC(ux=0, uy=0, uz=0)), ("z==0.5", fem.NodeBC(uz=p["uz1"])),
           ("x==0 || x==1.0", fem.NodeBC(ux=0, uy=0)), ("y==0 || y==1.0", fem.NodeBC(ux=0, uy=0))]
    ok, msg = fem.solve_stage(m, bcs, nincs=10, tol=1e-2, autoinc=True)
    assert ok, msg
    bcs[1] = ("z==0.5", fem.NodeBC(uz=p["uz2"]))
    ok, msg = fem.solve_stage(m, bcs, nincs=10, tol=1e-2, autoinc=True)
    assert ok, msg
    assert m.state[:, 12].max() > 0  # plastic flow happened
