"""GPU: distributed loads integrated on the device (amr_cell_loads) against
the host restatement of mech_boundary_forces and the body forces, which the
oracle fixes to 1e-13 (tests/test_host_model.py), and the reference's
elastic-hex8 known answers solved with device-built loads.

Bar: force vectors within 1e-13 of their largest entry (the host sums with
numpy's einsum and LAPACK's determinant, the device in node order with a
cofactor determinant; sin/exp come from different math libraries).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VALUES = [2.5e5, "3*z", "x^2 + sin(y) - 0.5*t", "-1e6*(1-x)*exp(-z)",
          "max(x, 0.3)*tan(z) - min(y, 0.5) + log10(2+x) + div(7*y, 0.3)"]


def mesh(am, shape, seed=0):
    hi = np.array([1.0, 0.6, 0.5])
    m = am.Mesh(am.Block([[0, 0, 0], list(hi)], nx=3, ny=2, nz=2, cellshape=shape))
    X = m.coords
    inner = np.all((X > 1e-9) & (X < hi - 1e-9), axis=1)
    X[inner] += np.random.default_rng(seed).uniform(-0.016, 0.016, (inner.sum(), 3))
    return m


def close(a, b):
    return np.abs(a - b).max() <= 1e-13 * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("shape", ["HEX8", "HEX20", "TET10"])
def test_facet_and_edge_loads_match_host(amaru, shape):
    from amaru_amd import model as M
    m = mesh(amaru, shape)
    _, fn = m.facets()
    _, en = m.edges()
    for val in VALUES:
        for key in ("tx", "ty", "tz", "tn"):
            C = m.coords[fn]
            assert close(M.distributed_forces_device(0, C, key, val, t=0.3), M.distributed_forces(C, key, val, t=0.3))
        for key in ("qx", "qz"):
            C = m.coords[en]
            assert close(M.distributed_forces_device(0, C, key, val, t=0.3), M.distributed_forces(C, key, val, t=0.3))


@pytest.mark.parametrize("shape,nips", [("HEX8", 0), ("HEX20", 27), ("TET10", 0)])
def test_body_loads_match_host(amaru, shape, nips):
    from amaru_amd.expr import compile_expr
    m = mesh(amaru, shape)
    model = amaru.FEModel(m, [("all", amaru.LinearElastic(E=1e9, nu=0.3))], nips=nips)
    elems = np.arange(model.ne)
    for val in VALUES:
        for key in ("wx", "wz"):
            host = model._body_forces(elems, key, compile_expr(val))
            dev = model._body_forces(elems, key, compile_expr(val), val, device=0)
            assert close(dev, host)


def test_bc_vals_and_reference_pins_with_device_loads(amaru, pins):
    """The reference's elastic-hex8 load cases (test/mech/elem/elastic-hex8.jl):
    right-hand sides equal to the host's, and the golden displacements."""
    p = pins["elastic_hex8"]
    base = [("x==0 && y==0 && z==0", amaru.NodeBC(ux=0, uy=0)), ("x==1 && y==0 && z==0", amaru.NodeBC(uy=0)),
            ("x==0 && y==1 && z==0", amaru.NodeBC(ux=0)), ("z==0", amaru.NodeBC(uz=0))]
    loads = [("z==1", amaru.NodeBC(fz=1)), ("y==1 && z==1", amaru.EdgeBC(qy=2)),
             ("x==1", amaru.SurfaceBC(tx="3*z")), ("x>=0", amaru.BodyC(wz=-1))]
    for case, load in enumerate(loads):
        m = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 1]], nx=1, ny=1, nz=1, cellshape="HEX8", tag="solid"))
        model = amaru.FEModel(m, [("solid", amaru.MechSolid, amaru.LinearElastic(E=p["E"], nu=p["nu"]))])
        bcs = base + [load]
        model.configure_dofs(bcs)
        U0, F0 = model.bc_vals(bcs)
        U1, F1 = model.bc_vals(bcs, device=0)
        assert np.array_equal(U0, U1) and close(F1, F0)
        dev = amaru.DeviceModel(model)
        dev.assemble()
        U, F = U1.copy(), F1.copy()
        dev.solve_system(U, F)
        uz = np.zeros(model.nnodes)
        has = model.eq[:, 2] >= 0
        uz[has] = U[model.eq[has, 2]]
        assert np.allclose(uz, p["uz"][case], atol=p["atol"])
        dev.close()


def test_bad_programs_are_refused(amaru):
    from amaru_amd import _lib
    C = np.zeros((1, 4, 3))
    N = np.full((1, 4), 0.25)
    D = np.zeros((1, 4, 2))
    w = np.ones(1)
    F = np.zeros_like(C)
    bad = np.array([[14, 0]], dtype=np.int32)   # ADD with an empty stack
    rc = _lib.lib().amr_cell_loads(0, 1, 4, 2, 1, _lib.ptr(C), _lib.ptr(N), _lib.ptr(D), _lib.ptr(w), 2,
                                   _lib.ptr(bad), None, 1, 0, 0.0, _lib.ptr(F))
    assert rc == -1 and "underflow" in _lib.lib().amr_last_error().decode()
