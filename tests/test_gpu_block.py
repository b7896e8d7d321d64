"""GPU: the 812k-DOF block's workload at sizes SuperLU solves in seconds.

The block of the headline measurement: unit cube of HEX20 elements with 27
integration points (changequadrature!, src/element.jl:185-188), E = 30 GPa,
nu = 0.2, the z = 0 face clamped and a 1 MPa pressure on z = 1 (tz = -1e6).
At 6^3 and 8^3 elements the device K must equal the oracle's mount_K bit for
bit, and the displacements and reactions of both solve paths (the drop-in
solve_system! and the device-resident NR iteration the benchmark times)
must be within 1e-8 of SuperLU on the oracle's K11, for the plain FP64 path
and the FP32-storage / FP64-accumulation path at rtol 1e-10.
"""
import numpy as np
import pytest

from oracle import HEX20 as O20, fem

pytestmark = pytest.mark.gpu


def block(amaru, n):
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 1]], nx=n, ny=n, nz=n, cellshape="HEX20"))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid, amaru.LinearElastic(E=30e9, nu=0.2))], nips=27)
    bcs = [("z==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("z==1", amaru.SurfaceBC(tz=-1e6))]
    ndofs, nu = model.configure_dofs(bcs)
    U, F = model.bc_vals(bcs)
    om = fem.Model(mesh.coords, mesh.conn, O20, [("LinearElastic", (30e9, 0.2))], nips=27)
    obcs = [("z==0", fem.SurfaceBC(ux=0, uy=0, uz=0)), ("z==1", fem.SurfaceBC(tz=-1e6))]
    assert om.configure_dofs(obcs) == (ndofs, nu)
    Uo, Fo = om.bc_vals(obcs)
    assert np.array_equal(U, Uo) and np.allclose(F, Fo, rtol=0, atol=1e-9 * np.abs(Fo).max())
    return mesh, model, om, bcs, U, F


@pytest.mark.parametrize("n", [6, 8])
def test_block_workload_vs_superlu(amaru, n):
    mesh, model, om, bcs, U, F = block(amaru, n)
    nu = model.nu
    dev = amaru.DeviceModel(model)
    dev.assemble()
    Ko = om.mount_K().tocsc()
    Kg = dev.K_csc()
    assert np.array_equal(Kg.indptr, Ko.indptr) and np.array_equal(Kg.indices, Ko.indices)
    assert np.array_equal(Kg.data, Ko.data)
    # the FP64 direct oracle: SuperLU on K11, reactions K21 U1 + K22 U2
    Ud, Fd = U.copy(), F.copy()
    assert fem.solve_system(Ko.tocsr(), Ud, Fd, nu) is None
    for precision in ("fp64", "mixed64acc"):
        # drop-in solve_system!
        Ug, Fg = U.copy(), F.copy()
        info = dev.solve_system(Ug, Fg, precision=precision, rtol=1e-10)
        assert info["converged"]
        err = np.linalg.norm(Ug - Ud) / np.linalg.norm(Ud)
        rerr = np.linalg.norm(Fg[nu:] - Fd[nu:]) / np.linalg.norm(Fd[nu:])
        assert err <= 1e-8 and rerr <= 1e-8, (precision, err, rerr)
        # the benchmark's step: one NR iteration of a one-increment stage on
        # the device (mount_K + solve_system! + update_state!); the reactions
        # are the internal forces at the prescribed dofs (mech-solver.jl:394-397)
        dev.stage_begin(U * 0, F)
        dev.increment_begin(1.0, 1.0, False)
        r = dev.nr_iteration(1, dev.opts(precision=precision, rtol=1e-10))
        assert r["status"] == 0
        dUa, dFin = dev.stage_get(last=True)
        err = np.linalg.norm(dUa[:nu] - Ud[:nu]) / np.linalg.norm(Ud[:nu])
        rerr = np.linalg.norm(dFin[nu:] - Fd[nu:]) / np.linalg.norm(Fd[nu:])
        assert err <= 1e-8 and rerr <= 1e-8, (precision, err, rerr)
        # the residual the NR loop tests is the unbalanced force, far below tol
        assert r["res"] < 1e-4 * np.abs(F).max()
    dev.close()


def test_column_offsets_do_not_change_results(amaru):
    """The SpMV reads 16-bit column offsets when every block's fits (all
    structured meshes here) and 32-bit columns otherwise; both give the same
    bits, for the stage solve and for a pool."""
    mesh, model, om, bcs, U, F = block(amaru, 4)
    dev = amaru.DeviceModel(model)
    dev.assemble()
    assert dev.spmv_col_bytes() == 2
    outs = {}
    for nbytes in (2, 4, 2):
        dev.set_spmv_col_bytes(nbytes)
        assert dev.spmv_col_bytes() == nbytes
        for precision in ("fp64", "mixed", "mixed64acc"):
            Ug, Fg = U.copy(), F.copy()
            info = dev.solve_system(Ug, Fg, precision=precision, rtol=1e-10)
            key = (precision,)
            if key in outs:
                assert info["iterations"] == outs[key][2]
                assert np.array_equal(Ug, outs[key][0]) and np.array_equal(Fg, outs[key][1])
            else:
                outs[key] = (Ug, Fg, info["iterations"])
    with pytest.raises(amaru.AmaruError):
        dev.set_spmv_col_bytes(3)
    dev.close()



@pytest.mark.parametrize("shape,nips,mat", [("HEX20", 27, "LE"), ("HEX20", 8, "LE"), ("HEX20", 27, "VM"),
                                            ("HEX8", 8, "LE"), ("TET10", 4, "LE")])
def test_mfma_stiffness_matches_oracle(amaru, shape, nips, mat):
    """The element stiffness on the FP64 matrix cores (the NR iterations'
    path when selected): linear elastic models through the Gram form
    lambda S + mu S' + mu tr(S) I of S = sum_q coef g g' (every shape), von
    Mises HEX20 through sum_q B'(coef D)B from yielded states.  Every element
    matrix within 1e-12 relative Frobenius of the oracle's (the summation
    order differs); the exported K keeps the exact path's values and
    pattern."""
    import test_gpu_parity as T
    mesh, model, om = T.pair(amaru, shape, mat, nips, n=(4, 3, 2), seed=3)
    bcs = [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0))]
    model.configure_dofs(bcs)
    om.configure_dofs([("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0))])
    dev = amaru.DeviceModel(model)
    st = om.state
    if mat == "VM":
        dU = np.random.default_rng(11).normal(size=model.ndofs) * 4e-3
        _, st, msg = om.update_state(dU, om.state)
        assert msg is None and (st[:, 13] != 0).any()
        dev.update_state(dU)    # trial states = the oracle's st
    Ko = om.elem_stiffness(st)
    dev.set_ke_path("mfma")
    U, F = model.bc_vals(bcs)
    dev.stage_begin(U, F)
    dev.increment_begin(1.0, 1.0, False)
    dev.nr_prepare(1)           # mount_K on the matrix cores
    Km = dev.Ke()
    err = np.linalg.norm((Km - Ko).reshape(len(Ko), -1), axis=1) / np.linalg.norm(Ko.reshape(len(Ko), -1), axis=1)
    assert err.max() <= 1e-12, err.max()
    # amr_assemble (mount_K with the reference's pattern) stays exact
    dev.assemble()
    assert np.array_equal(dev.Ke(), Ko)
    dev.close()


def test_mfma_block_solve_vs_superlu(amaru):
    """The benchmark's step with the matrix-core stiffness on the 6^3 block:
    displacements and reactions within 1e-8 of SuperLU."""
    mesh, model, om, bcs, U, F = block(amaru, 6)
    nu = model.nu
    Ko = om.mount_K().tocsc()
    Ud, Fd = U.copy(), F.copy()
    assert fem.solve_system(Ko.tocsr(), Ud, Fd, nu) is None
    dev = amaru.DeviceModel(model)
    dev.set_ke_path("mfma")
    dev.stage_begin(U * 0, F)
    dev.increment_begin(1.0, 1.0, False)
    r = dev.nr_iteration(1, dev.opts(precision="mixed64acc", rtol=1e-10))
    assert r["status"] == 0
    dUa, dFin = dev.stage_get(last=True)
    assert np.linalg.norm(dUa[:nu] - Ud[:nu]) <= 1e-8 * np.linalg.norm(Ud[:nu])
    assert np.linalg.norm(dFin[nu:] - Fd[nu:]) <= 1e-8 * np.linalg.norm(Fd[nu:])
    dev.close()
