"""GPU: the reference's failure and edge branches on the device, against the
oracle.

- a negative Jacobian determinant (src/mech/elem/mech-solid.jl:150) fails
  mount_K with the first bad cell's id;
- von Mises "Negative value for √J2D" (src/mech/mat/von-mises.jl:146) fails
  update_state! for a crafted state (negative accumulated plastic strain);
- NaN material parameters give "NaN values in internal forces vector"
  (src/mech/mech-solver.jl:142);
- Drucker-Prager's return to the apex (src/mech/mat/drucker-prager.jl:
  135-140), chosen on the previous increment's plastic multiplier (:130),
  with the states bit-exact against the oracle.
"""
import numpy as np
import pytest

from oracle import HEX8 as O8, HEX20 as O20, TET10 as O10, fem

pytestmark = pytest.mark.gpu

OSHAPE = {"HEX8": O8, "HEX20": O20, "TET10": O10}
# node orders mirrored through the element's mid-plane z (HEX) or with two
# corners swapped (TET10: corners 1, 2 and the mid-edge nodes that follow
# them), which turns the element inside out
INVERT = {"HEX8": [4, 5, 6, 7, 0, 1, 2, 3],
          "HEX20": [4, 5, 6, 7, 0, 1, 2, 3, 12, 13, 14, 15, 8, 9, 10, 11, 16, 17, 18, 19],
          "TET10": [0, 2, 1, 3, 6, 5, 4, 7, 9, 8]}


def models(amaru, shape, mat, omat, nips=0, n=(3, 2, 2), bad=None, clamp=True):
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1.0, 0.6, 0.5]], nx=n[0], ny=n[1], nz=n[2], cellshape=shape))
    if bad is not None:
        mesh.conn[bad] = mesh.conn[bad][:, INVERT[shape]]
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid, mat)], nips=nips)
    om = fem.Model(mesh.coords, mesh.conn, OSHAPE[shape], [omat], nips=nips)
    bcs = [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0))] if clamp else []
    model.configure_dofs(bcs)
    om.configure_dofs([("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0))] if clamp else [])
    return mesh, model, om


@pytest.mark.parametrize("shape,nips", [("HEX8", 0), ("HEX20", 27), ("TET10", 0)])
def test_negative_jacobian(amaru, shape, nips):
    bad = [4, 7]
    mesh, model, om = models(amaru, shape, amaru.LinearElastic(E=1e9, nu=0.3), ("LinearElastic", (1e9, 0.3)),
                             nips, bad=bad)
    with pytest.raises(RuntimeError, match=r"Negative Jacobian determinant in cell 5\b"):
        om.mount_K()
    dev = amaru.DeviceModel(model)
    with pytest.raises(amaru.AmaruError) as ei:
        dev.assemble()
    assert ei.value.code == -3 and "Negative Jacobian determinant in cell 5" in str(ei.value)
    dev.close()


@pytest.mark.parametrize("shape,nips", [("HEX8", 0), ("HEX20", 8), ("TET10", 0)])
def test_von_mises_sqrt_j2d_failure(amaru, shape, nips):
    """State with eps_pa = -1 in one element: sigma_y + H eps_pa < 0 makes the
    trial state plastic with a multiplier larger than sqrt(J2)/(sqrt3 G)."""
    mat = amaru.VonMises(E=200e9, nu=0.3, fy=250e6, H=2e9)
    mesh, model, om = models(amaru, shape, mat, ("VonMises", (200e9, 0.3, 250e6, 2e9)), nips)
    st = np.zeros((model.ne * model.nips, 14))
    e = 3
    st[e * model.nips:(e + 1) * model.nips, 12] = -1.0
    dU = np.random.default_rng(3).normal(size=model.ndofs) * 1e-6
    _, _, msg = om.update_state(dU, st)
    assert msg == f"material update failed in cell {e + 1}"
    dev = amaru.DeviceModel(model)
    dev.state_set(st)
    with pytest.raises(amaru.AmaruError) as ei:
        dev.update_state(dU)
    assert ei.value.code == -4 and "Negative value for √J2D" in str(ei.value) and f"cell {e + 1}" in str(ei.value)
    # the same state away from that element updates normally
    st[:] = 0.0
    dev.state_set(st)
    dFg = dev.update_state(dU)
    dFo, sto, msg = om.update_state(dU, st)
    assert msg is None and np.array_equal(dFg, dFo) and np.array_equal(dev.state_get(), sto)
    dev.close()


@pytest.mark.parametrize("kind", ["LE", "DP"])
def test_nan_parameter_gives_nan_internal_forces(amaru, kind):
    if kind == "LE":
        mat = amaru.LinearElastic(E=1e9, nu=0.3)
        mat.E = float("nan")
        omat = ("LinearElastic", (float("nan"), 0.3))
    else:
        mat = amaru.DruckerPrager(E=50e6, nu=0.3, alpha=0.23094, kappa=24000.0)
        mat.kappa = float("nan")
        omat = ("DruckerPrager", (50e6, 0.3, 0.23094, float("nan"), 0.0))
    mesh, model, om = models(amaru, "HEX8", mat, omat)
    dU = np.random.default_rng(4).normal(size=model.ndofs) * 1e-4
    dFo, _, msg = om.update_state(dU, om.state)
    assert msg == "solve_system!: NaN values in internal forces vector"
    dev = amaru.DeviceModel(model)
    with pytest.raises(amaru.AmaruError) as ei:
        dev.update_state(dU)
    assert ei.value.code == -6 and "NaN values in internal forces vector" in str(ei.value)
    dev.close()


@pytest.mark.parametrize("shape,nips", [("HEX8", 0), ("HEX20", 27), ("TET10", 0)])
@pytest.mark.parametrize("H", [0.0, 1e6])
def test_drucker_prager_apex_return(amaru, shape, nips, H):
    """Previous multiplier 1e-3 at every IP and a uniform volumetric
    expansion of 1e-3 (alpha j1 > kappa; no supports, so the field is the
    same at every IP) with a tiny random part: sqrt(J2tr) stays below
    dgamma_prev n G, so every IP returns to the apex, where the stress is
    hydrostatic.  A larger random part sends some IPs down the conventional
    branch in the same launch."""
    mat = amaru.DruckerPrager(E=50e6, nu=0.3, alpha=0.23094, kappa=24000.0, H=H)
    mesh, model, om = models(amaru, shape, mat, ("DruckerPrager", (50e6, 0.3, 0.23094, 24000.0, H)), nips,
                             clamp=False)
    st = np.zeros((model.ne * model.nips, 14))
    st[:, 12] = 0.01      # eps_pa (enters the apex multiplier through H)
    st[:, 13] = 1e-3      # the previous increment's dgamma
    X = np.zeros(model.ndofs)
    has = model.eq >= 0
    X[model.eq[has]] = mesh.coords[has]
    rng = np.random.default_rng(5)
    dev = amaru.DeviceModel(model)
    seen = []
    for amp in (1e-7, 1e-4):
        dU = 1e-3 * X + rng.normal(size=model.ndofs) * amp
        dFo, sto, msg = om.update_state(dU, st)
        assert msg is None
        dev.state_set(st)
        dFg = dev.update_state(dU)
        stg = dev.state_get()
        assert np.array_equal(stg, sto), f"max |dstate| = {np.abs(stg - sto).max()}"
        assert np.array_equal(dFg, dFo)
        apex = (stg[:, 3:6] == 0.0).all(axis=1) & (stg[:, 0] == stg[:, 1]) & (stg[:, 1] == stg[:, 2])
        seen.append(apex.mean())
    assert seen[0] == 1.0            # all IPs at the apex
    assert 0.0 < seen[1] < 1.0       # both branches in one launch
    dev.close()
