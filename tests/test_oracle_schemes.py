"""The oracle's predictor-corrector schemes (src/mech/mech-solver.jl:280-349)
pinned by closed forms.

1. Homogeneous uniaxial stress: a HEX8 unit cube on symmetry rollers with a
   traction S = 2 sigma_y on x = 1, von Mises with linear hardening, one
   Newton-Raphson iteration from the virgin state.  The elastic predictor
   reaches a plastic state whose flow direction is the uniaxial one,
   a = (1, -1/2, -1/2, 0, 0, 0), so the reference's continuum tangent there
   is Dep = De - De a a' De / (a' De a + sqrt(1.5) H)
   (src/mech/mat/von-mises.jl:112-125).  With the combined tangent
   Dc = a1 De + a2 Dep (Dc = De for :FE) the iteration's strain is
   Dc^-1 (S, 0, 0, 0, 0, 0), exactly (a homogeneous field on one element),
   and the corner node (1, 1, 1) moves by that strain's normal components.
   Ralston's predictor is scaled by 2/3: 2/3 S = 1.33 sigma_y is plastic too.
2. A linear material: the tangent does not change, so every scheme gives
   the :FE displacements (to rounding).
"""
import math

import numpy as np
import pytest

from oracle import HEX8 as O8, fem

E, NU, FY, H = 200e9, 0.3, 250e6, 2e9


def De():
    lam = E * NU / ((1 + NU) * (1 - 2 * NU))
    G = E / (2 * (1 + NU))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[:3, :3] += np.diag([2 * G] * 3)
    D[3:, 3:] = np.diag([2 * G] * 3)
    return D


def Dep():
    D = De()
    a = np.array([1.0, -0.5, -0.5, 0, 0, 0])
    Da = D @ a
    return D - np.outer(Da, Da) / (a @ Da + math.sqrt(1.5) * H)


COMB = {"FE": (1.0, 0.0), "ME": (0.5, 0.5), "BE": (0.0, 1.0), "Ralston": (0.25, 0.75)}


@pytest.mark.parametrize("scheme", ["FE", "ME", "BE", "Ralston"])
def test_one_iteration_uniaxial_closed_form(scheme):
    X, C = fem.block_mesh((0, 0, 0), (1, 1, 1), 1, 1, 1, O8)
    m = fem.Model(X, C, O8, [("VonMises", (E, NU, FY, H))])
    S = 2 * FY
    bcs = [("x==0", fem.NodeBC(ux=0)), ("y==0", fem.NodeBC(uy=0)), ("z==0", fem.NodeBC(uz=0)),
           ("x==1", fem.SurfaceBC(tx=S))]
    ndofs, nu = m.configure_dofs(bcs)
    U, F = m.bc_vals(bcs)
    st = np.zeros_like(m.state)
    dUi, dFin, st2, msg = fem.nr_step(m, st, st, np.zeros(ndofs), U.copy(), F.copy(), nu, scheme)
    assert msg is None
    a1, a2 = COMB[scheme]
    eps = np.linalg.solve(a1 * De() + a2 * Dep(), np.array([S, 0, 0, 0, 0, 0]))
    corner = int(np.nonzero((X == 1.0).all(axis=1))[0][0])
    got = np.array([dUi[m.eq[corner, d]] for d in range(3)])
    assert np.allclose(got, eps[:3], rtol=1e-12, atol=0), (got, eps[:3])
    # the update after the iteration is plastic everywhere
    assert (st2[:, 12] > 0).all()


@pytest.mark.parametrize("scheme", ["ME", "BE", "Ralston"])
def test_linear_material_schemes_equal_fe(scheme):
    X, C = fem.block_mesh((0, 0, 0), (1.0, 0.2, 0.2), 8, 2, 2, O8)
    bcs = [("x==0", fem.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", fem.SurfaceBC(tz=-1e6))]
    out = {}
    for sc in ("FE", scheme):
        m = fem.Model(X, C, O8, [("LinearElastic", (E, NU))])
        ok, msg = fem.solve_stage(m, bcs, nincs=2, scheme=sc)
        assert ok, msg
        out[sc] = m.uvals.copy()
    assert np.allclose(out[scheme], out["FE"], rtol=0, atol=1e-12 * np.abs(out["FE"]).max())
