"""Host side of the product against the oracle (CPU only).

The structured mesher, DOF numbering and boundary-condition vectors of
``amaru.jl_amd`` must give the oracle's node coordinates and connectivity
bit for bit, the same equation ids, and the same load vectors (up to
summation order).
"""
import numpy as np
import pytest

from oracle import HEX8 as O8, HEX20 as O20, TET10 as O10, fem

CASES = [("HEX8", O8, (0, 0, 0), (1.0, 0.1, 0.1), (20, 4, 4)),
         ("HEX20", O20, (0, 0, -0.05), (0.05, 1.0, 0.05), (1, 50, 2)),
         ("TET10", O10, (0, 0, 0), (1, 1, 1), (3, 2, 2)),
         ("HEX20", O20, (0, 0, 0), (1, 1, 1), (3, 3, 3)),
         ("HEX8", O8, (-1.5, 0.25, 2.0), (1.0, 1.0, 3.0), (7, 3, 5))]


@pytest.mark.parametrize("shape,oshape,lo,hi,n", CASES)
def test_mesh_matches_oracle_bitwise(amaru, shape, oshape, lo, hi, n):
    m = amaru.Mesh(amaru.Block([lo, hi], nx=n[0], ny=n[1], nz=n[2], cellshape=shape))
    X, C = fem.block_mesh(lo, hi, *n, oshape)
    assert m.coords.shape == X.shape
    assert np.array_equal(m.coords, X)
    assert np.array_equal(m.conn, C)


@pytest.mark.parametrize("shape,count", [("HEX8", 1331), ("HEX20", 4961), ("TET10", 9261)])
def test_node_counts_reference(amaru, pins, shape, count):
    m = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 1]], nx=10, ny=10, nz=10, cellshape=shape))
    assert m.nnodes == count == pins["structured_node_counts"][shape]


def test_outer_facets_match_oracle(amaru):
    m = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 1]], nx=2, ny=3, nz=2, cellshape="TET10"))
    owners, nodes = m.facets()
    ref = fem.outer_facets(m.conn, O10)
    assert [(int(o), list(map(int, f))) for o, f in zip(owners, nodes)] == [(o, f) for o, f in ref]


def _bcs(amaru, kind):
    if kind == "cantilever":
        return ([("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", amaru.SurfaceBC(tz=-1e5))],
                [("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", fem.SurfaceBC(tz=-1e5))])
    return ([("z==0", amaru.NodeBC(ux=0, uy=0, uz=0)), ("x==0 || x==1", amaru.SurfaceBC(ux=0)),
             ("z==1", amaru.SurfaceBC(tn="-10*x")), ("y==1 && z==1", amaru.EdgeBC(qy=2)),
             ("x>=0", amaru.BodyC(wz=-1)), ("x==1 && y==1 && z==1", amaru.NodeBC(fx=3.0))],
            [("z==0", fem.NodeBC(ux=0, uy=0, uz=0)), ("x==0 || x==1", fem.SurfaceBC(ux=0)),
             ("z==1", fem.SurfaceBC(tn="-10*x")), ("y==1 && z==1", fem.EdgeBC(qy=2)),
             ("x>=0", fem.BodyC(wz=-1)), ("x==1 && y==1 && z==1", fem.NodeBC(fx=3.0))])


@pytest.mark.parametrize("shape,oshape", [("HEX8", O8), ("HEX20", O20), ("TET10", O10)])
@pytest.mark.parametrize("kind", ["cantilever", "mixed"])
def test_dofs_and_bc_vectors_match_oracle(amaru, shape, oshape, kind):
    lo, hi = (0, 0, 0), (1, 1, 1)
    mesh = amaru.Mesh(amaru.Block([lo, hi], nx=3, ny=2, nz=2, cellshape=shape))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid, amaru.LinearElastic(E=1e3, nu=0.25))])
    bcs, obcs = _bcs(amaru, kind)
    ndofs, nu = model.configure_dofs(bcs)
    om = fem.Model(mesh.coords, mesh.conn, oshape, [("LinearElastic", (1e3, 0.25))])
    ondofs, onu = om.configure_dofs(obcs)
    assert (ndofs, nu) == (ondofs, onu)
    assert np.array_equal(model.eq, om.eq)
    U, F = model.bc_vals(bcs)
    Uo, Fo = om.bc_vals(obcs)
    assert np.array_equal(U, Uo)
    assert np.allclose(F, Fo, rtol=1e-13, atol=1e-13 * np.abs(Fo).max())
    assert abs(F.sum() - Fo.sum()) < 1e-12 * max(1.0, np.abs(Fo).sum())
