"""Pins for the oracle's constitutive updates (CPU).

Closed forms that a plausible slip (a dropped term, a wrong sqrt factor, the
wrong hardening denominator) would break:
  * elastic: dsig = lam tr(de) I + 2 mu de (Lamé form, Mandel shear),
  * von Mises radial return with H = 0 lands on the yield surface; with H > 0
    the reference's denominator 3G + sqrt(1.5) H leaves
    f = ftr (sqrt(1.5) - 1) H / (3G + sqrt(1.5) H) (src/mech/mat/von-mises.jl:145),
  * Drucker-Prager conventional and apex returns end on f = 0,
  * the continuum tangents annihilate the flow direction when H = 0.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import lib, ptr

SQ2 = math.sqrt(2.0)


def calcDe(E, nu):
    D = np.zeros(36)
    lib().ora_calcDe(E, nu, ptr(D))
    return D.reshape(6, 6)


def update(kind, prm, st, de):
    p = np.zeros(6)
    p[:len(prm)] = prm
    st = np.array(st, dtype=float)
    ds = np.zeros(6)
    rc = lib().ora_update_state(kind, ptr(p), ptr(st), ptr(np.asarray(de, dtype=float)), ptr(ds))
    return rc, st, ds


def calcD(kind, prm, st):
    p = np.zeros(6)
    p[:len(prm)] = prm
    D = np.zeros(36)
    lib().ora_calcD(kind, ptr(p), ptr(np.asarray(st, dtype=float)), ptr(D))
    return D.reshape(6, 6)


def J2(s):
    t = [s[0], s[1], s[2], s[3] / SQ2, s[4] / SQ2, s[5] / SQ2]
    return ((t[0] - t[1]) ** 2 + (t[1] - t[2]) ** 2 + (t[2] - t[0]) ** 2) / 6 + t[3] ** 2 + t[4] ** 2 + t[5] ** 2


def dev(s):
    p = (s[0] + s[1] + s[2]) / 3
    return np.array([s[0] - p, s[1] - p, s[2] - p, s[3], s[4], s[5]])


def test_elastic_lame():
    E, nu = 200e9, 0.3
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    de = np.array([1e-4, -2e-4, 3e-5, 4e-5 * SQ2, -1e-5 * SQ2, 2e-5 * SQ2])
    rc, st, ds = update(0, (E, nu), np.zeros(14), de)
    expect = 2 * mu * de + lam * de[:3].sum() * np.array([1, 1, 1, 0, 0, 0])
    assert rc == 0
    assert np.allclose(ds, expect, rtol=1e-14, atol=1e-14 * abs(expect).max())
    assert np.array_equal(st[6:12], de)
    assert np.allclose(calcDe(E, nu) @ de, expect, rtol=1e-14, atol=1e-14 * abs(expect).max())


@pytest.mark.parametrize("H", [0.0, 2e9])
def test_von_mises_return(H):
    E, nu, fy = 200e9, 0.3, 250e6
    G = E / (2 * (1 + nu))
    rng = np.random.default_rng(3)
    for _ in range(20):
        de = rng.normal(size=6) * 3e-3
        st0 = np.zeros(14)
        st0[12] = 0.01
        rc, st, ds = update(1, (E, nu, fy, H), st0, de)
        assert rc == 0
        sig = st[:6]
        strial = calcDe(E, nu) @ de
        ftr = math.sqrt(3 * J2(strial)) - fy - H * 0.01
        assert ftr > 0
        dl = st[13]
        assert dl == pytest.approx(ftr / (3 * G + math.sqrt(1.5) * H), rel=1e-13)
        f_after = math.sqrt(3 * J2(sig)) - fy - H * st[12]
        expect = ftr * (math.sqrt(1.5) - 1) * H / (3 * G + math.sqrt(1.5) * H)
        assert f_after == pytest.approx(expect, abs=1e-9 * fy)
        # radial: deviator direction is kept, pressure unchanged
        assert np.allclose(dev(sig) / np.linalg.norm(dev(sig)), dev(strial) / np.linalg.norm(dev(strial)),
                           atol=1e-12)
        assert sig[:3].sum() == pytest.approx(strial[:3].sum(), rel=1e-12, abs=1e-3)
        assert st[12] == pytest.approx(0.01 + dl, rel=1e-14)
        assert np.allclose(ds, sig, rtol=0, atol=1e-12 * abs(sig).max())


def test_von_mises_elastic_branch_and_failure_flag():
    rc, st, ds = update(1, (200e9, 0.3, 250e6, 0.0), np.zeros(14), np.full(6, 1e-9))
    assert rc == 0 and st[13] == 0.0


def test_von_mises_tangent_annihilates_flow_direction():
    E, nu, fy = 200e9, 0.3, 250e6
    rc, st, _ = update(1, (E, nu, fy, 0.0), np.zeros(14), np.array([3e-3, -1e-3, 0, 1e-3, 0, 0]))
    D = calcD(1, (E, nu, fy, 0.0), st)
    n = dev(st[:6])
    assert np.abs(D @ n).max() < 1e-9 * np.abs(calcDe(E, nu) @ n).max()
    assert np.abs(D - D.T).max() < 1e-6 * np.abs(D).max()


def test_drucker_prager_returns():
    E, nu, alpha, kappa = 50e6, 0.3, 0.23094, 24000.0
    K = E / (3 * (1 - 2 * nu))
    # conventional (deviatoric-dominated) return
    rc, st, _ = update(2, (E, nu, alpha, kappa, 0.0), np.zeros(14), np.array([1e-3, -1e-3, 0, 2e-3, 0, 0]))
    s = st[:6]
    f = alpha * s[:3].sum() + math.sqrt(J2(s)) - kappa
    assert rc == 0 and st[13] > 0 and abs(f) < 1e-9 * kappa
    # apex return: large volumetric extension, taken when the PREVIOUS
    # dgamma makes sqrt(J2tr) - dgamma n G <= 0 (src/mech/mat/drucker-prager.jl:130)
    de = np.array([5e-2, 5e-2, 5e-2, 1e-4, 0, 0])
    prev = np.zeros(14)
    prev[13] = 1.0
    rc, st, _ = update(2, (E, nu, alpha, kappa, 0.0), prev, de)
    s = st[:6]
    assert rc == 0 and np.all(s[3:] == 0.0) and s[0] == s[1] == s[2]
    assert abs(alpha * s[:3].sum() - kappa) < 1e-9 * kappa
    dg_apex = st[13]
    assert dg_apex == pytest.approx((alpha * 3 * K * 0.15 - kappa) / (3 * math.sqrt(3) * alpha * K), rel=1e-12)
    # with a zero previous dgamma the same trial state takes the conventional
    # branch instead (the reference's "todo fix this")
    rc, st2, _ = update(2, (E, nu, alpha, kappa, 0.0), np.zeros(14), de)
    assert rc == 0 and st2[13] != dg_apex and st2[3] != 0.0
    D = calcD(2, (E, nu, alpha, kappa, 0.0), st)
    assert np.all(np.isfinite(D))


def test_drucker_prager_tangent_annihilates_flow_direction():
    E, nu, alpha, kappa = 50e6, 0.3, 0.23094, 24000.0
    rc, st, _ = update(2, (E, nu, alpha, kappa, 0.0), np.zeros(14), np.array([1e-3, -1e-3, 0, 2e-3, 0, 0]))
    D = calcD(2, (E, nu, alpha, kappa, 0.0), st)
    s = dev(st[:6])
    V = alpha * np.array([1, 1, 1, 0, 0, 0]) + (s / np.linalg.norm(s)) / SQ2
    Nu = V / np.linalg.norm(V)
    assert np.abs(D @ Nu).max() < 1e-9 * np.abs(calcDe(E, nu) @ Nu).max()
