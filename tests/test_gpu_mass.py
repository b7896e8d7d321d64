"""GPU: mount_M on the device (amr_assemble_M, amr_get_M_csc) against the
oracle's restatement of elem_mass and the mount_K triplet rule.

Bar: pattern and values bit-exact (the kernel follows the oracle's
operation order: J by fma chains, coef = (rho * detJ) * w, m_ab + coef *
(N_a * N_b) over the IPs in order; assembly sums survivors in element
order).  At a larger size, the total mass equals rho times the volume.
"""
import numpy as np
import pytest

from oracle import HEX8 as O8, HEX20 as O20, TET10 as O10, fem

pytestmark = pytest.mark.gpu

OSHAPE = {"HEX8": O8, "HEX20": O20, "TET10": O10}


def pair(am, shape, nips=0, n=(3, 2, 2), seed=0):
    hi = np.array([1.0, 0.6, 0.5])
    mesh = am.Mesh(am.Block([[0, 0, 0], list(hi)], nx=n[0], ny=n[1], nz=n[2], cellshape=shape))
    X = mesh.coords
    inner = np.all((X > 1e-9) & (X < hi - 1e-9), axis=1)
    X[inner] += np.random.default_rng(seed).uniform(-0.016, 0.016, (inner.sum(), 3))
    model = am.FEModel(mesh, [("all", am.MechSolid(rho=7850.0), am.LinearElastic(E=200e9, nu=0.3))], nips=nips)
    om = fem.Model(mesh.coords, mesh.conn, OSHAPE[shape], [("LinearElastic", (200e9, 0.3))], nips=nips)
    return model, om


@pytest.mark.parametrize("shape,nips", [("HEX8", 0), ("HEX20", 8), ("HEX20", 27), ("TET10", 0)])
def test_M_bit_exact(amaru, shape, nips):
    model, om = pair(amaru, shape, nips)
    model.configure_dofs([("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0))])
    om.configure_dofs([("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0))])
    dev = amaru.DeviceModel(model)
    rho = np.linspace(1000.0, 8000.0, model.ne)
    dev.assemble_M(rho)
    A = dev.M_csc()
    B = om.mount_M(rho).tocsc()
    assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
    assert np.array_equal(A.data, B.data)
    # default densities from the binding's MechSolid(rho=...)
    dev.assemble_M()
    C = dev.M_csc()
    D = om.mount_M(7850.0).tocsc()
    assert np.array_equal(C.indptr, D.indptr) and np.array_equal(C.data, D.data)
    # K is untouched by the mass assembly
    dev.assemble()
    K = dev.K_csc()
    assert np.array_equal(K.data, om.mount_K().tocsc().data)
    dev.close()


def test_total_mass_at_size(amaru):
    """HEX20 (27 IP) 16x16x16 block: every direction carries rho * V."""
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [2.0, 1.0, 0.5]], nx=16, ny=16, nz=16, cellshape="HEX20"))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid(rho=2400.0), amaru.LinearElastic(E=30e9, nu=0.2))],
                          nips=27)
    model.configure_dofs([("z==0", amaru.SurfaceBC(ux=0))])
    dev = amaru.DeviceModel(model)
    dev.assemble_M()
    M = dev.M_csc()
    assert abs(M.sum() - 3 * 2400.0 * 1.0) <= 1e-11 * 3 * 2400.0
    dev.close()


def test_negative_density_is_refused(amaru):
    model, _ = pair(amaru, "HEX8")
    model.configure_dofs([("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0))])
    dev = amaru.DeviceModel(model)
    with pytest.raises(amaru.AmaruError, match="rho"):
        dev.assemble_M(-1.0)
    with pytest.raises(amaru.AmaruError, match="not been assembled"):
        dev.M_csc()
    dev.close()
