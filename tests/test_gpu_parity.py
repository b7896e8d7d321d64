"""GPU parity: the HIP path through the C ABI against the oracle.

Bars: element matrices, K (pattern and values), IP states and internal
forces bit-exact (the kernels follow the oracle's operation order);
displacements within 1e-8 relative of the SuperLU direct solve.
"""
import numpy as np
import pytest

from oracle import HEX8 as O8, HEX20 as O20, TET10 as O10, fem

pytestmark = pytest.mark.gpu

OSHAPE = {"HEX8": O8, "HEX20": O20, "TET10": O10}
MATS = {"LE": ("LinearElastic", (200e9, 0.3)), "VM": ("VonMises", (200e9, 0.3, 250e6, 2e9)),
        "DP": ("DruckerPrager", (50e6, 0.3, 0.23094, 24000.0, 0.0))}


def product_mat(am, name):
    kind, p = MATS[name]
    if kind == "LinearElastic":
        return am.LinearElastic(E=p[0], nu=p[1])
    if kind == "VonMises":
        return am.VonMises(E=p[0], nu=p[1], fy=p[2], H=p[3])
    return am.DruckerPrager(E=p[0], nu=p[1], alpha=p[2], kappa=p[3], H=p[4])


def distorted_mesh(am, shape, n=(3, 2, 2), seed=0, amp=0.08):
    mesh = am.Mesh(am.Block([[0, 0, 0], [1.0, 0.6, 0.5]], nx=n[0], ny=n[1], nz=n[2], cellshape=shape))
    rng = np.random.default_rng(seed)
    X = mesh.coords
    inner = np.all((X > 1e-9) & (X < np.array([1.0, 0.6, 0.5]) - 1e-9), axis=1)
    X[inner] += rng.uniform(-amp, amp, (inner.sum(), 3)) * 0.2
    return mesh


def pair(am, shape, mat="LE", nips=0, n=(3, 2, 2), seed=0):
    mesh = distorted_mesh(am, shape, n, seed)
    model = am.FEModel(mesh, [("all", am.MechSolid, product_mat(am, mat))], nips=nips)
    om = fem.Model(mesh.coords, mesh.conn, OSHAPE[shape], [MATS[mat]], nips=nips)
    return mesh, model, om


CASES = [("HEX8", 0), ("HEX20", 8), ("HEX20", 27), ("TET10", 0)]


@pytest.mark.parametrize("shape,nips", CASES)
def test_Ke_and_K_bit_exact_elastic(amaru, shape, nips):
    mesh, model, om = pair(amaru, shape, "LE", nips)
    bcs = [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0))]
    model.configure_dofs(bcs)
    om.configure_dofs([("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0))])
    dev = amaru.DeviceModel(model)
    dev.assemble()
    Kg = dev.Ke()
    Ko = om.elem_stiffness()
    assert np.array_equal(Kg, Ko), f"max |dKe| = {np.abs(Kg - Ko).max()}"
    A = dev.K_csc()
    B = om.mount_K().tocsc()
    assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
    assert np.array_equal(A.data, B.data)
    dev.close()


@pytest.mark.parametrize("shape,oshape", [("HEX20", O20), ("TET10", O10)])
def test_eps_dropped_pattern_matches(amaru, shape, oshape):
    """E = 1 makes round-off-level Ke entries fall on either side of eps();
    the exported pattern must drop exactly the oracle's entries."""
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 1]], nx=3, ny=2, nz=2, cellshape=shape))
    model = amaru.FEModel(mesh, [("all", amaru.LinearElastic(E=1.0, nu=0.3))], nips=27 if shape == "HEX20" else 0)
    om = fem.Model(mesh.coords, mesh.conn, oshape, [("LinearElastic", (1.0, 0.3))],
                   nips=27 if shape == "HEX20" else 0)
    model.configure_dofs([])
    om.configure_dofs([])
    dev = amaru.DeviceModel(model)
    dev.assemble()
    A = dev.K_csc()
    B = om.mount_K().tocsc()
    topo = 9 * len({(a, b) for el in mesh.conn for a in el for b in el})
    assert B.nnz < topo  # the eps rule really drops entries here
    assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
    assert np.array_equal(A.data, B.data)
    dev.close()


@pytest.mark.parametrize("shape,nips", CASES)
@pytest.mark.parametrize("mat", ["LE", "VM", "DP"])
def test_update_state_and_tangent_bit_exact(amaru, shape, nips, mat):
    mesh, model, om = pair(amaru, shape, mat, nips, seed=1)
    bcs = [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0))]
    model.configure_dofs(bcs)
    om.configure_dofs([("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0))])
    dev = amaru.DeviceModel(model)
    rng = np.random.default_rng(7)
    # large enough to yield both materials at many IPs
    scale = {"LE": 1e-3, "VM": 4e-3, "DP": 2e-2}[mat]
    dU = rng.normal(size=model.ndofs) * scale
    dFg = dev.update_state(dU)
    dFo, sto, msg = om.update_state(dU, om.state)
    assert msg is None
    stg = dev.state_get()
    assert np.array_equal(stg, sto), f"max |dstate| = {np.abs(stg - sto).max()}"
    assert np.array_equal(dFg, dFo)
    if mat != "LE":
        assert (sto[:, 13] != 0).any(), "no IP yielded; the case does not test plasticity"
    # tangent from the trial state (mount_K after update_state, as in the NR loop)
    dev.assemble()
    assert np.array_equal(dev.Ke(), om.elem_stiffness(sto))
    # commit / restore semantics
    dev.restore()
    assert np.array_equal(dev.state_get(), np.zeros_like(sto))
    dev.update_state(dU)
    dev.commit()
    assert np.array_equal(dev.state_get(committed=True), sto)
    dev.close()


@pytest.mark.parametrize("precision,precond", [("fp64", "jacobi"), ("mixed", "jacobi"), ("mixed64acc", "jacobi"),
                                               ("fp64", "block_jacobi"), ("mixed64acc", "block_jacobi")])
@pytest.mark.parametrize("shape,nips", [("HEX8", 0), ("HEX20", 27), ("TET10", 0)])
def test_solve_system_vs_direct(amaru, shape, nips, precision, precond):
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1.0, 0.2, 0.2]], nx=10, ny=2, nz=2, cellshape=shape))
    model = amaru.FEModel(mesh, [("all", amaru.LinearElastic(E=200e9, nu=0.3))], nips=nips)
    bcs = [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", amaru.SurfaceBC(tz=-1e5)),
           ("x==1 && z==0", amaru.NodeBC(ux=1e-6))]
    ndofs, nu = model.configure_dofs(bcs)
    U, F = model.bc_vals(bcs)
    dev = amaru.DeviceModel(model)
    dev.assemble()
    K = dev.K_csc().tocsr()
    Ug, Fg = U.copy(), F.copy()
    info = dev.solve_system(Ug, Fg, precision=precision, rtol=1e-12, precond=precond)
    Uo, Fo = U.copy(), F.copy()
    assert fem.solve_system(K, Uo, Fo, nu) is None
    err = np.linalg.norm(Ug - Uo) / np.linalg.norm(Uo)
    assert err <= 1e-8, (err, info)
    assert np.array_equal(Ug[nu:], U[nu:]) and np.array_equal(Fg[:nu], F[:nu])
    rerr = np.abs(Fg[nu:] - Fo[nu:]).max() / np.abs(Fo[nu:]).max()
    assert rerr <= 1e-8
    # converged means relres <= rtol or the residual reached its rounding floor
    # (64 eps || |K||x| + |b| ||); either way far below 1e-10 here
    assert info["converged"] and info["relres"] <= 1e-10
    dev.close()


def test_solve_failure_reports(amaru):
    """No essential BCs: K11 is singular; the solver must fail, not return garbage."""
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 1]], nx=2, ny=2, nz=2, cellshape="HEX8"))
    model = amaru.FEModel(mesh, [("all", amaru.LinearElastic(E=1.0, nu=0.3))])
    model.configure_dofs([])
    dev = amaru.DeviceModel(model)
    dev.assemble()
    U = np.zeros(model.ndofs)
    F = np.ones(model.ndofs)
    with pytest.raises(amaru.AmaruError):
        dev.solve_system(U, F, maxiter=2000)
    dev.close()


def test_spmv_lanes_do_not_change_results(amaru):
    """Every lanes-per-row choice of the SpMV (set explicitly, the default
    rule's, or the timed one) gives bit-identical solutions (fixed 64-slot
    reduction tree), for both precisions."""
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1.0, 1.0, 1.0]], nx=4, ny=4, nz=4, cellshape="HEX20"))
    model = amaru.FEModel(mesh, [("all", amaru.LinearElastic(E=30e9, nu=0.2))], nips=27)
    bcs = [("z==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("z==1", amaru.SurfaceBC(tz=-1e6))]
    model.configure_dofs(bcs)
    U, F = model.bc_vals(bcs)
    dev = amaru.DeviceModel(model)
    dev.assemble()
    for precision in ("fp64", "mixed", "mixed64acc"):
        outs = []
        for lanes in (8, 16, 32, 64, "default", "tuned"):
            if lanes == "default":
                dev.set_spmv_lanes(0, 0)
            elif lanes == "tuned":
                dev.tune_spmv_lanes()
                assert set(dev.spmv_lanes().values()) <= {8, 16, 32, 64}
            else:
                dev.set_spmv_lanes(lanes, lanes)
            Ug, Fg = U.copy(), F.copy()
            info = dev.solve_system(Ug, Fg, precision=precision, rtol=1e-10)
            outs.append((Ug, Fg, info["iterations"]))
        for Ug, Fg, it in outs[1:]:
            assert it == outs[0][2]
            assert np.array_equal(Ug, outs[0][0]) and np.array_equal(Fg, outs[0][1])
    dev.close()


TILES = {("HEX8", 0): ((1, 2, 4), (32, 16, 8)), ("HEX20", 27): ((7, 10, 20), (9, 4, 2)),
         ("HEX20", 8): ((7, 10, 20), (16, 8, 32)), ("TET10", 0): ((2, 5, 10), (32, 16, 64))}


@pytest.mark.parametrize("shape,nips", CASES)
def test_element_tiles_do_not_change_results(amaru, shape, nips):
    """Every stiffness tiling (row nodes per lane) and update tiling
    (elements per workgroup) gives the oracle's bits; the tuner picks one of
    them and later models of the shape use it."""
    mesh, model, om = pair(amaru, shape, "VM", nips, n=(5, 3, 2), seed=2)
    model.configure_dofs([("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0))])
    om.configure_dofs([("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0))])
    dU = np.random.default_rng(9).normal(size=model.ndofs) * 4e-3
    dFo, sto, msg = om.update_state(dU, om.state)
    assert msg is None
    Ko = om.elem_stiffness(sto)
    dev = amaru.DeviceModel(model)
    achs, epws = TILES[(shape, nips)]
    for ach, epw in zip(achs, epws):
        dev.set_elem_tiles(ach, epw)
        assert dev.elem_tiles() == (ach, epw)
        dev.restore()
        assert np.array_equal(dev.update_state(dU), dFo)
        assert np.array_equal(dev.state_get(), sto)
        dev.assemble()
        assert np.array_equal(dev.Ke(), Ko)
    ach, epw = dev.tune_elem_tiles()
    assert ach in achs and epw in epws and dev.elem_tiles() == (ach, epw)
    dev.close()
    dev2 = amaru.DeviceModel(model)
    assert dev2.elem_tiles() == (ach, epw)     # the shape's tuned choice
    dev2.close()
