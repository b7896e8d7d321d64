"""Pins for the oracle's shape functions and quadrature (CPU).

Mirrors the reference's own shape tests (test/mesh/shape/shape_deriv.jl:6-59:
Kronecker delta at nodes, partition of unity, finite-difference derivatives)
and adds closed-form quadrature moments.
"""
import math

import numpy as np
import pytest

from oracle import HEX8, HEX20, TET10, rule, shape_deriv, shape_func

HEX_NAT = np.array([
    [-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1],
    [0, -1, -1], [1, 0, -1], [0, 1, -1], [-1, 0, -1], [0, -1, 1], [1, 0, 1], [0, 1, 1], [-1, 0, 1],
    [-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], dtype=float)
TET_NAT = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [.5, 0, 0], [.5, .5, 0], [0, .5, 0],
                    [0, 0, .5], [.5, 0, .5], [0, .5, .5]])
NAT = {HEX8: HEX_NAT[:8], HEX20: HEX_NAT, TET10: TET_NAT}


@pytest.mark.parametrize("et", [HEX8, HEX20, TET10])
def test_kronecker_delta(et):
    P = NAT[et]
    M = np.array([shape_func(et, p) for p in P])
    assert np.allclose(M, np.eye(len(P)), atol=1e-15)


@pytest.mark.parametrize("et,nips", [(HEX8, 8), (HEX20, 8), (HEX20, 27), (TET10, 4)])
def test_partition_of_unity_and_fd_derivative(et, nips):
    R, w = rule(et, nips)
    h = 1e-6
    for r in R:
        N = shape_func(et, r)
        D = shape_deriv(et, r)
        assert abs(N.sum() - 1.0) < 1e-14
        assert np.allclose(D.sum(axis=0), 0.0, atol=1e-14)
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fd = (shape_func(et, r + e) - shape_func(et, r - e)) / (2 * h)
            assert np.allclose(fd, D[:, k], atol=1e-6)


def test_quadrature_weights_and_moments():
    # 2x2x2 Gauss: exact for degree <= 3 per axis; the literals carry 15 digits
    R, w = rule(HEX8, 8)
    assert abs(w.sum() - 8.0) < 1e-14
    assert abs((w * R[:, 0] ** 2 * R[:, 1] ** 2 * R[:, 2] ** 2).sum() - (2 / 3) ** 3) < 1e-13
    R, w = rule(HEX20, 27)
    assert abs(w.sum() - 8.0) < 1e-13
    assert abs((w * R[:, 0] ** 4 * R[:, 1] ** 2).sum() - (2 / 5) * (2 / 3) * 2) < 1e-13
    # the HEX20 default rule is 8 points (src/shape/solids3d.jl:706)
    assert len(rule(HEX20, 0)[1]) == 8
    R, w = rule(TET10, 4)
    assert abs(w.sum() - 1 / 6) < 1e-16
    # unit tetrahedron moments: int x^a y^b z^c = a! b! c! / (a+b+c+3)!
    assert abs((w * R[:, 0] ** 2).sum() - 2 / math.factorial(5)) < 1e-15
    assert abs((w * R[:, 0] * R[:, 1]).sum() - 1 / math.factorial(5)) < 1e-15


def test_ip_order_r_fastest():
    R, _ = rule(HEX20, 27)
    assert R[1, 0] == 0.0 and R[1, 1] < 0 and R[3, 1] == 0.0 and R[9, 2] == 0.0
