"""Host tables of solid shape functions at IP rules (used for body loads).

HEX8 / HEX20 / TET10 as in src/shape/solids3d.jl (node numbering of the
reference), rules HEX_IP8, HEX_IP27, TET_IP4 (src/shape/quadrature.jl).
The device kernels receive dN/dR tables built by the C library itself.
"""
import numpy as np

from .mesh import HEX8, HEX20, TET10

_HEX = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1],
                 [0, -1, -1], [1, 0, -1], [0, 1, -1], [-1, 0, -1], [0, -1, 1], [1, 0, 1], [0, 1, 1], [-1, 0, 1],
                 [-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], dtype=float)
_TET_EDGES = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]


def rule(shape, nips=0):
    g2, g3 = 0.577350269189626, 0.774596669241483
    if shape in (HEX8, HEX20) and nips in (0, 8):
        R = np.array([[i, j, k] for k in (-g2, g2) for j in (-g2, g2) for i in (-g2, g2)])
        return R, np.ones(8)
    if shape in (HEX8, HEX20) and nips == 27:
        w3 = [0.171467764060357, 0.274348422496571, 0.438957475994513, 0.702331961591221]
        pts = (-g3, 0.0, g3)
        R, w = [], []
        for k in range(3):
            for j in range(3):
                for i in range(3):
                    R.append([pts[i], pts[j], pts[k]])
                    w.append(w3[(i == 1) + (j == 1) + (k == 1)])
        return np.array(R), np.array(w)
    if shape == TET10 and nips in (0, 4):
        a, b = 0.5854101966249685, 0.1381966011250105
        return np.array([[a, b, b], [b, a, b], [b, b, a], [b, b, b]]), np.full(4, 0.04166666666666667)
    raise ValueError(f"no {nips}-point rule for {shape}")


def func_deriv(shape, R):
    r, s, t = R
    if shape == HEX8:
        H = _HEX[:8]
        fr, fs, ft = 1 + H[:, 0] * r, 1 + H[:, 1] * s, 1 + H[:, 2] * t
        N = 0.125 * fr * fs * ft
        D = np.stack([0.125 * H[:, 0] * fs * ft, 0.125 * H[:, 1] * fr * ft, 0.125 * H[:, 2] * fr * fs], axis=1)
        return N, D
    if shape == HEX20:
        N, D = np.zeros(20), np.zeros((20, 3))
        for a in range(20):
            ra, sa, ta = _HEX[a]
            fr, fs, ft = 1 + ra * r, 1 + sa * s, 1 + ta * t
            if a < 8:
                g = ra * r + sa * s + ta * t - 2
                N[a] = 0.125 * fr * fs * ft * g
                D[a] = [0.125 * ra * fs * ft * (g + fr), 0.125 * sa * fr * ft * (g + fs), 0.125 * ta * fr * fs * (g + ft)]
            elif ra == 0:
                N[a] = 0.25 * (1 - r * r) * fs * ft
                D[a] = [-0.5 * r * fs * ft, 0.25 * (1 - r * r) * sa * ft, 0.25 * (1 - r * r) * fs * ta]
            elif sa == 0:
                N[a] = 0.25 * fr * (1 - s * s) * ft
                D[a] = [0.25 * (1 - s * s) * ra * ft, -0.5 * s * fr * ft, 0.25 * (1 - s * s) * fr * ta]
            else:
                N[a] = 0.25 * fr * fs * (1 - t * t)
                D[a] = [0.25 * (1 - t * t) * ra * fs, 0.25 * (1 - t * t) * fr * sa, -0.5 * t * fr * fs]
        return N, D
    L = np.array([1 - r - s - t, r, s, t])
    dL = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
    N = np.zeros(10)
    D = np.zeros((10, 3))
    N[:4] = L * (2 * L - 1)
    D[:4] = (4 * L - 1)[:, None] * dL
    for e, (i, j) in enumerate(_TET_EDGES):
        N[4 + e] = 4 * L[i] * L[j]
        D[4 + e] = 4 * (dL[i] * L[j] + L[i] * dL[j])
    return N, D


def rule_tables(shape, nips=0):
    R, w = rule(shape, nips)
    ND = [func_deriv(shape, r) for r in R]
    return np.array([a for a, _ in ND]), np.array([b for _, b in ND]), w
