"""GPU: output post-processing (amr_ip_values, amr_nodal_recovery) against
the oracle's restatement of ip_state_vals and nodal_patch_recovery.

Bars: the copied components and the material extras bit-exact; principal
stresses and von Mises within 1e-13 of the largest stress; recovered nodal
values within 1e-9 of each field's largest value (the device fits on a
centred, scaled basis by QR, the oracle by Julia's SVD pseudo-inverse on raw
coordinates: the same least-squares polynomial, different rounding); the
patch counts per node exact.
"""
import numpy as np
import pytest

from oracle import HEX8 as O8, HEX20 as O20, TET10 as O10, recovery as R

pytestmark = pytest.mark.gpu

OSHAPE = {"HEX8": O8, "HEX20": O20, "TET10": O10}
KIND = {"LE": 0, "VM": 1, "DP": 2}


def material(am, kind):
    if kind == "LE":
        return am.LinearElastic(E=200e9, nu=0.3)
    if kind == "VM":
        return am.VonMises(E=200e9, nu=0.3, fy=250e6, H=2e9)
    return am.DruckerPrager(E=50e6, nu=0.3, alpha=0.23094, kappa=24000.0, H=0.0)


def device_model(am, shape, kind, n=(3, 2, 2), nips=0, seed=0, lo=(0.0, 0.0, 0.0), hi=(1.0, 0.6, 0.5)):
    mesh = am.Mesh(am.Block([list(lo), list(hi)], nx=n[0], ny=n[1], nz=n[2], cellshape=shape))
    X = mesh.coords
    inner = np.all((X > np.array(lo) + 1e-9) & (X < np.array(hi) - 1e-9), axis=1)
    X[inner] += np.random.default_rng(seed).uniform(-0.016, 0.016, (inner.sum(), 3))
    model = am.FEModel(mesh, [("all", am.MechSolid, material(am, kind))], nips=nips)
    model.configure_dofs([(f"x=={lo[0]}", am.SurfaceBC(ux=0, uy=0, uz=0))])
    return model, am.DeviceModel(model)


def random_states(nip, seed):
    rng = np.random.default_rng(seed)
    st = np.zeros((nip, 14))
    st[:, 0:6] = rng.standard_normal((nip, 6)) * 1e6
    st[:, 6:12] = rng.standard_normal((nip, 6)) * 1e-3
    st[:, 12] = rng.uniform(0, 0.02, nip)
    return st


@pytest.mark.parametrize("shape,nips,kind", [("HEX8", 0, "LE"), ("HEX20", 27, "VM"), ("HEX20", 8, "DP"),
                                             ("TET10", 0, "DP"), ("TET10", 0, "LE")])
def test_ip_values_and_nodal_recovery_match_oracle(amaru, shape, nips, kind):
    model, dev = device_model(amaru, shape, kind, nips=nips)
    nip = model.ne * model.nips
    st = random_states(nip, 7)
    dev.state_set(st)
    V = dev.ip_values()
    Vo = R.ip_values(st, np.full(nip, KIND[kind]))
    exact = list(range(0, 6)) + list(range(10, 17))
    assert np.array_equal(V[:, exact], Vo[:, exact])
    smax = np.abs(st[:, :6]).max()
    assert np.abs(V[:, 6:10] - Vo[:, 6:10]).max() <= 1e-13 * smax
    assert np.abs(V[:, 17:] - Vo[:, 17:]).max() <= 1e-13 * smax
    N, reps = dev.nodal_values()
    No, repso = R.nodal_patch_recovery(model.coords, model.conn, OSHAPE[shape], Vo, nips)
    assert np.array_equal(reps, repso)
    scale = np.maximum(np.abs(No).max(0), 1e-300)
    err = (np.abs(N - No) / scale).max()
    assert err <= 1e-9, f"max scaled difference {err:.3e}"
    # committed copy (state_set writes both) gives the same values
    Nc, _ = dev.nodal_values(committed=True)
    assert np.array_equal(Nc, N)
    dev.close()


def test_recovery_reproduces_polynomials_at_size(amaru):
    """Hex20 (27 IP) 12x12x12 block far from the origin: fields in the
    regression space come back exactly at every node."""
    lo, hi = (10.0, -3.0, 4.0), (11.0, -2.0, 5.0)
    model, dev = device_model(amaru, "HEX20", "LE", n=(12, 12, 12), nips=27, lo=lo, hi=hi)
    XI = R.ip_coords(model.coords, model.conn, O20, 27)
    st = np.zeros((len(XI), 14))
    x, y, z = XI[:, 0] - 10.5, XI[:, 1] + 2.5, XI[:, 2] - 4.5

    def f(x, y, z):
        return 1e6 * (1.0 + 2 * x - y + 3 * z + 4 * x * y - 2 * y * z + x * z)
    st[:, 0] = f(x, y, z)
    st[:, 7] = 1e-3 * (x - 2 * z)
    dev.state_set(st)
    N, reps = dev.nodal_values()
    X = model.coords
    xn, yn, zn = X[:, 0] - 10.5, X[:, 1] + 2.5, X[:, 2] - 4.5
    assert np.all(reps > 0)
    assert np.abs(N[:, 0] - f(xn, yn, zn)).max() <= 1e-10 * np.abs(f(xn, yn, zn)).max()
    assert np.abs(N[:, 11] - 1e-3 * (xn - 2 * zn)).max() <= 1e-10 * 1.5e-3
    dev.close()


def test_node_fields_after_a_solve(amaru, tmp_path):
    """A plastic Drucker-Prager footing: the recovered fields carry the
    reference's names, match the oracle on the solved states, and go to a
    VTU file with the displacements."""
    from amaru_amd import workloads as W
    ana, kw = W.footing(n=3, settlement=0.04)
    ok, msg = amaru.solve(ana, **kw)
    assert ok, msg
    dev, model = ana.dev, ana.model
    fields = dev.node_fields(committed=True)
    assert list(fields)[:3] == ["σxx", "σyy", "σzz"] and list(fields)[-3:] == ["epa", "j1", "srj2d"]
    st = dev.state_get(committed=True)
    assert (st[:, 13] != 0).any()
    Vo = R.ip_values(st, np.full(len(st), 2))
    No, _ = R.nodal_patch_recovery(model.coords, model.conn, O8, Vo)
    scale = np.maximum(np.abs(No).max(0), 1e-300)
    got = np.stack([fields[name] for _, name in amaru.output_fields(model)], 1)
    cols = [c for c, _ in amaru.output_fields(model)]
    assert (np.abs(got - No[:, cols]) / scale[cols]).max() <= 1e-9
    from amaru_amd.vtu import save_vtu
    import xml.etree.ElementTree as ET
    path = tmp_path / "footing.vtu"
    save_vtu(model, str(path), node_data={"U": model.uvals, **fields})
    arrays = {da.get("Name"): da for da in ET.parse(path).getroot().iter("DataArray")}
    assert set(fields) | {"U", "Points", "connectivity", "offsets", "types"} == set(arrays)
    epa = np.array(arrays["epa"].text.split(), dtype=float)
    assert np.allclose(epa, fields["epa"].astype(np.float32), rtol=1e-9, atol=0)
    dev.close()


@pytest.mark.parametrize("shape,plastic", [("HEX8", "VM"), ("TET10", "VM"), ("TET10", "DP"), ("HEX20", "DP")])
def test_mixed_material_kinds_match_oracle(amaru, shape, plastic):
    """Linear elastic and plastic elements in one model: the stress and
    strain fields are fitted on whole patches, the plastic kind's extras on
    the subpatches of the plastic elements (fe-model.jl:614-680), with pinv's
    minimum-norm fit where a subpatch's basis is rank-deficient (small TET10
    subpatches at the interface)."""
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [2.0, 1.2, 0.9]], nx=4, ny=3, nz=2, cellshape=shape))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid, material(amaru, "LE")),
                                 ("x>=1", amaru.MechSolid, material(amaru, plastic))])
    model.configure_dofs([("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0))])
    dev = amaru.DeviceModel(model)
    nip = model.ne * model.nips
    st = random_states(nip, 5)
    dev.state_set(st)
    ekind = np.array([KIND["LE"] if type(model.mats[k]).__name__ == "LinearElastic" else KIND[plastic]
                      for k in model.emat])
    assert (ekind == 0).any() and (ekind != 0).any()
    Vo = R.ip_values(st, np.repeat(ekind, model.nips))
    has = np.ones((model.ne, Vo.shape[1]), bool)
    has[:, 16:] = (ekind != 0)[:, None]
    N, reps = dev.nodal_values()
    No, repso = R.nodal_patch_recovery(model.coords, model.conn, OSHAPE[shape], Vo, 0, has=has)
    assert np.array_equal(reps, repso)
    scale = np.maximum(np.abs(No).max(0), 1e-300)
    err = (np.abs(N - No) / scale).max()
    assert err <= 1e-9, f"max scaled difference {err:.3e}"
    dev.close()


def test_von_mises_with_drucker_prager_is_refused(amaru):
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 1]], nx=2, ny=1, nz=1, cellshape="HEX8"))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid, material(amaru, "VM")),
                                 ("x>=0.5", amaru.MechSolid, material(amaru, "DP"))])
    model.configure_dofs([("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0))])
    dev = amaru.DeviceModel(model)
    dev.ip_values()   # per-IP values need no patches
    with pytest.raises(amaru.AmaruError, match="share columns"):
        dev.nodal_values()
    dev.close()
