"""GPU: the sparsity pattern and gather maps made on the device.

The block rows, the per-block contribution lists (ascending element) and
the per-node force lists are built on the device when a model is created;
a node whose elements bring more contributions than one workgroup sorts
(2048) sends the model to the host builder.  Both must give K (pattern and
values) equal to the oracle's mount_K bit for bit (src/mech/mech-solver.jl:
86-102), on the edges of each path:

* a fan of HEX8 cubes sharing one corner node: 255 cubes (2040
  contributions, the largest sort on the device) and 257 cubes (2056, the
  host builder);
* one HEX8 cube whose nodes are numbered 0-3 and 39996-39999 in a 40000-node
  model (the rest unused): column offsets beyond 16 bits, so the SpMV reads
  32-bit columns, and rows without blocks;
* TET10 and HEX20 blocks with the node numbering shuffled.
"""
import numpy as np
import pytest

from oracle import HEX8 as O8, HEX20 as O20, TET10 as O10, fem

pytestmark = pytest.mark.gpu

E, NU = 200e9, 0.3


def unit_cube(amaru):
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 1]], nx=1, ny=1, nz=1, cellshape="HEX8"))
    return mesh, mesh.coords[mesh.conn[0]] - mesh.coords[mesh.conn[0][0]]


def check_K(amaru, mesh, oshape):
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid, amaru.LinearElastic(E=E, nu=NU))])
    model.configure_dofs([])
    om = fem.Model(mesh.coords, mesh.conn, oshape, [("LinearElastic", (E, NU))])
    om.configure_dofs([])
    dev = amaru.DeviceModel(model)
    dev.assemble()
    A = dev.K_csc()
    B = om.mount_K().tocsc()
    assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
    assert np.array_equal(A.data, B.data)
    return dev


@pytest.mark.parametrize("ncubes", [255, 257])
def test_fan_of_cubes(amaru, ncubes):
    mesh, C8 = unit_cube(amaru)
    X = np.zeros((1 + 7 * ncubes, 3))
    conn = np.zeros((ncubes, 8), dtype=np.int64)
    for e in range(ncubes):
        ids = 1 + 7 * e + np.arange(7)
        X[ids] = C8[1:] * (1.0 + 0.01 * e)
        conn[e] = np.concatenate([[0], ids])
    mesh.coords, mesh.conn = X, conn.astype(mesh.conn.dtype)
    dev = check_K(amaru, mesh, O8)
    dev.close()


def test_far_columns_and_empty_rows(amaru):
    mesh, C8 = unit_cube(amaru)
    n = 40000
    ids = np.array([0, 1, 2, 3, n - 4, n - 3, n - 2, n - 1])
    X = np.zeros((n, 3))
    X[ids] = C8
    mesh.coords, mesh.conn = X, ids[None, :].astype(mesh.conn.dtype)
    dev = check_K(amaru, mesh, O8)
    assert dev.spmv_col_bytes() == 4
    dev.close()


@pytest.mark.parametrize("shape", ["TET10", "HEX20"])
def test_shuffled_numbering(amaru, shape):
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1.0, 0.6, 0.5]], nx=4, ny=3, nz=2, cellshape=shape))
    perm = np.random.default_rng(5).permutation(mesh.nnodes)
    X = np.empty_like(mesh.coords)
    X[perm] = mesh.coords
    mesh.coords, mesh.conn = X, perm[mesh.conn].astype(mesh.conn.dtype)
    dev = check_K(amaru, mesh, O10 if shape == "TET10" else O20)
    assert dev.spmv_col_bytes() == 2
    dev.close()
