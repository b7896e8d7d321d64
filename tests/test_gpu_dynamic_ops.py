"""GPU: the dynamic solver's use of K and M on the device (amr_combine_KM,
amr_matvec) against the oracle's mount_K and mount_M.

Bars: the pseudo-stiffness Kp = a K + b M equal entry for entry to the
oracle's a K + b M (the same two products and one sum per entry); K x and
M x within 1e-13 of scipy's products (different summation order); a solve
with Kp within 1e-8 of SuperLU.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import HEX8 as O8, HEX20 as O20, TET10 as O10, fem

pytestmark = pytest.mark.gpu

OSHAPE = {"HEX8": O8, "HEX20": O20, "TET10": O10}


def pair(am, shape, nips):
    hi = np.array([1.0, 0.6, 0.5])
    mesh = am.Mesh(am.Block([[0, 0, 0], list(hi)], nx=3, ny=2, nz=2, cellshape=shape))
    X = mesh.coords
    inner = np.all((X > 1e-9) & (X < hi - 1e-9), axis=1)
    X[inner] += np.random.default_rng(1).uniform(-0.016, 0.016, (inner.sum(), 3))
    model = am.FEModel(mesh, [("all", am.MechSolid(rho=7850.0), am.LinearElastic(E=200e9, nu=0.3))], nips=nips)
    om = fem.Model(mesh.coords, mesh.conn, OSHAPE[shape], [("LinearElastic", (200e9, 0.3))], nips=nips)
    model.configure_dofs([("x==0", am.SurfaceBC(ux=0, uy=0, uz=0))])
    om.configure_dofs([("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0))])
    return model, om


@pytest.mark.parametrize("shape,nips", [("HEX8", 0), ("HEX20", 27), ("TET10", 0)])
def test_products_and_pseudo_stiffness(amaru, shape, nips):
    model, om = pair(amaru, shape, nips)
    dev = amaru.DeviceModel(model)
    dev.assemble()
    dev.assemble_M()
    K = om.mount_K().tocsc()
    M = om.mount_M(7850.0).tocsc()
    x = np.random.default_rng(2).standard_normal(model.ndofs)
    for which, A in (("K", K), ("M", M)):
        y = dev.matvec(which, x)
        ref = A @ x
        scale = abs(A) @ np.abs(x)   # the size of each row's terms
        assert np.all(np.abs(y - ref) <= 1e-13 * scale)
    # Newmark pseudo-stiffness with Rayleigh damping C = alpha M + beta K
    dt, alpha, beta = 1e-3, 0.2, 1e-4
    a, b = 1 + 2 * beta / dt, 4 / dt ** 2 + 2 * alpha / dt
    dev.combine_KM(a, b)
    Kp = dev.K_csc()
    ref = a * K.toarray() + b * M.toarray()
    assert np.array_equal(Kp.toarray(), ref)
    # solve_system! on Kp: unknowns first, prescribed zero
    nu = model.nu
    F = np.zeros(model.ndofs)
    F[:nu] = np.random.default_rng(3).standard_normal(nu) * 1e3
    U = np.zeros(model.ndofs)
    dev.solve_system(U, F, rtol=1e-12)
    Kpc = Kp.tocsc()
    ref_u = spla.spsolve(Kpc[:nu, :nu].tocsc(), F[:nu])
    assert np.linalg.norm(U[:nu] - ref_u) <= 1e-8 * np.linalg.norm(ref_u)
    dev.close()


def test_combine_needs_both_matrices(amaru):
    model, _ = pair(amaru, "HEX8", 0)
    dev = amaru.DeviceModel(model)
    dev.assemble()
    with pytest.raises(amaru.AmaruError, match="amr_assemble_M"):
        dev.combine_KM(1.0, 1.0)
    with pytest.raises(amaru.AmaruError, match="M has not been assembled"):
        dev.matvec("M", np.zeros(model.ndofs))
    dev.close()
