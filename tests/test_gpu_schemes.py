"""GPU: the predictor-corrector schemes :FE, :ME, :BE and :Ralston
(src/mech/mech-solver.jl:280-349) on the device.

1. One Newton-Raphson iteration on a homogeneous uniaxial-stress cube
   against the closed form of tests/test_oracle_schemes.py (strain
   Dc^-1 (S, 0, ...), Dc = a1 De + a2 Dep with the reference's continuum
   tangent).
2. Full von Mises stages (a cantilever with linear hardening under a tip
   traction, fixed increments and auto-increment) against the oracle's
   stage solver with the same scheme: the same outcome (with fixed
   increments :ME and :Ralston stop when the residual grows, as the
   oracle's do), the same Newton-Raphson path, the same yielded integration
   points, displacements and reactions within 1e-7.
"""
import numpy as np
import pytest

from oracle import HEX8 as O8, fem

pytestmark = pytest.mark.gpu

E, NU, FY, H = 200e9, 0.3, 250e6, 2e9


@pytest.mark.parametrize("scheme", ["FE", "ME", "BE", "Ralston"])
def test_one_iteration_uniaxial_closed_form(amaru, scheme):
    import test_oracle_schemes as T
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 1]], nx=1, ny=1, nz=1, cellshape="HEX8"))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid, amaru.VonMises(E=E, nu=NU, fy=FY, H=H))])
    S = 2 * FY
    bcs = [("x==0", amaru.NodeBC(ux=0)), ("y==0", amaru.NodeBC(uy=0)), ("z==0", amaru.NodeBC(uz=0)),
           ("x==1", amaru.SurfaceBC(tx=S))]
    model.configure_dofs(bcs)
    U, F = model.bc_vals(bcs)
    dev = amaru.DeviceModel(model)
    dev.set_scheme(scheme)
    dev.stage_begin(U, F)
    dev.increment_begin(1.0, 1.0, False)
    r = dev.nr_iteration(1, dev.opts(precision="fp64", rtol=1e-14))
    assert r["status"] == 0, r["message"]
    dUa, _ = dev.stage_get(last=True)
    a1, a2 = T.COMB[scheme]
    eps = np.linalg.solve(a1 * T.De() + a2 * T.Dep(), np.array([S, 0, 0, 0, 0, 0]))
    corner = int(np.nonzero((mesh.coords == 1.0).all(axis=1))[0][0])
    got = np.array([dUa[model.eq[corner, d]] for d in range(3)])
    assert np.allclose(got, eps[:3], rtol=1e-10, atol=0), (got, eps[:3])
    assert (dev.state_get()[:, 12] > 0).all()
    dev.close()


@pytest.mark.parametrize("scheme", ["ME", "BE", "Ralston"])
@pytest.mark.parametrize("autoinc", [False, True])
def test_von_mises_stage_matches_oracle(amaru, scheme, autoinc):
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1.0, 0.2, 0.2]], nx=10, ny=2, nz=2, cellshape="HEX8"))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid, amaru.VonMises(E=E, nu=NU, fy=FY, H=H))])
    ana = amaru.MechAnalysis(model)
    tz = -4.5e7 if autoinc else -3.0e7
    amaru.addstage(ana, [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", amaru.SurfaceBC(tz=tz))], nincs=6)
    ok, msg = amaru.solve(ana, scheme=scheme, autoinc=autoinc, maxits=8, tol=1.0, precision="fp64",
                          pcg_rtol=1e-13)
    om = fem.Model(mesh.coords, mesh.conn, O8, [("VonMises", (E, NU, FY, H))])
    olog = []
    ook, omsg = fem.solve_stage(om, [("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", fem.SurfaceBC(tz=tz))],
                                nincs=6, maxits=8, tol=1.0, autoinc=autoinc, log=olog, scheme=scheme)
    assert ok == ook, (msg, omsg)
    assert ok or scheme != "BE"
    assert [(inc, it) for _, inc, it, _ in ana.log] == [(inc, it) for inc, it, _ in olog]
    m = ana.model
    assert (m.state[:, 12] > 0).sum() > 0.05 * len(m.state)   # it yielded
    assert np.array_equal(m.state[:, 12] > 0, om.state[:, 12] > 0)
    for a, b in ((m.uvals, om.uvals), (m.fvals, om.fvals)):
        assert np.linalg.norm(a - b) <= 1e-7 * np.linalg.norm(b)
    assert np.abs(m.state - om.state).max() <= 1e-6 * np.abs(om.state).max()
    ana.dev.close()
