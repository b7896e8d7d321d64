"""Seeded synthetic workloads of the batched and served forms.

* ``cantilever_batch(seed, n)``: independent Tet10 cantilevers 1 x 0.1 x 0.1 m,
  element counts log-uniform in [500, 20000] (m = round((N/60)^(1/3)),
  mesh 10m x m x m cells of 6 tets), E ~ U[10, 200] GPa, nu ~ U[0.2, 0.35],
  x = 0 clamped, tip load resultant ~ U[-1, 1] kN in z on x = 1 (4 IPs).
* ``footing(n)``: elastoplastic soil block 10 x 10 x 5 m, (2n x 2n x n) Hex8,
  Drucker-Prager (E = 50 MPa, nu = 0.3, c = 20 kPa, phi = 30 deg as the
  compression cone: alpha = 0.23094, kappa = 24 kPa, H = 0), base clamped,
  lateral faces on rollers, a 2 x 2 m centred rigid footing pushed down
  0.04 m in 20 increments (n = 30 is the 60 x 60 x 30 block, 346,053 DOFs).
* ``served_stream(seed, n, rate)``: Poisson arrivals; each request a box
  4a x a x a (a = 0.25 m) meshed with Hex8 or Tet10 (50/50) at a DOF count
  log-uniform in [1e3, 3e5]; 70 % linear elastic under a tip traction, 30 %
  von Mises (fy = 250 MPa, H = 2 GPa) under a prescribed tip deflection of
  twice the elastic-limit deflection, applied in 5 increments.
"""
import math

import numpy as np

from .analysis import MechAnalysis, addstage
from .mesh import HEX8, TET10, Block, Mesh
from .model import DruckerPrager, FEModel, LinearElastic, MechSolid, NodeBC, SurfaceBC, VonMises


def cantilever_params(seed, n):
    rng = np.random.default_rng(seed)
    N = np.exp(rng.uniform(math.log(500.0), math.log(20000.0), n))
    E = rng.uniform(10e9, 200e9, n)
    nu = rng.uniform(0.2, 0.35, n)
    P = rng.uniform(-1e3, 1e3, n)
    return [dict(nelem=float(N[i]), E=float(E[i]), nu=float(nu[i]), P=float(P[i])) for i in range(n)]


def cantilever(p):
    """(MechAnalysis, kwargs) of one Tet10 cantilever."""
    m = max(1, int(round((p["nelem"] / 60.0) ** (1.0 / 3.0))))
    mesh = Mesh(Block([[0, 0, 0], [1.0, 0.1, 0.1]], nx=10 * m, ny=m, nz=m, cellshape=TET10))
    model = FEModel(mesh, [("all", MechSolid, LinearElastic(E=p["E"], nu=p["nu"]))])
    ana = MechAnalysis(model)
    addstage(ana, [("x==0", SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", SurfaceBC(tz=p["P"] / 0.01))])
    return ana


def cantilever_batch(seed=0, n=256):
    return [cantilever(p) for p in cantilever_params(seed, n)]


def footing(n=30, settlement=0.04, nincs=20):
    """(MechAnalysis, solver kwargs) of the footing problem."""
    mesh = Mesh(Block([[0, 0, 0], [10.0, 10.0, 5.0]], nx=2 * n, ny=2 * n, nz=n, cellshape=HEX8))
    model = FEModel(mesh, [("all", MechSolid, DruckerPrager(E=50e6, nu=0.3, alpha=0.23094, kappa=24000.0, H=0.0))])
    ana = MechAnalysis(model)
    bcs = [("z==0", SurfaceBC(ux=0, uy=0, uz=0)), ("x==0 || x==10", SurfaceBC(ux=0)),
           ("y==0 || y==10", SurfaceBC(uy=0)), ("z==5 && x>=4 && x<=6 && y>=4 && y<=6", NodeBC(uz=-settlement))]
    addstage(ana, bcs, nincs=nincs)
    return ana, {"maxits": 5}


def request_params(seed, n):
    rng = np.random.default_rng(seed)
    dofs = np.exp(rng.uniform(math.log(1e3), math.log(3e5), n))
    shape = rng.integers(0, 2, n)
    plastic = rng.uniform(0.0, 1.0, n) < 0.3
    E = rng.uniform(70e9, 210e9, n)
    out = []
    for i in range(n):
        out.append(dict(dofs=float(dofs[i]), shape="HEX8" if shape[i] == 0 else "TET10", vm=bool(plastic[i]),
                        E=float(E[i]), nu=0.3))
    return out


def box_request(p, a=0.25):
    """(MechAnalysis, solver kwargs) of one served request."""
    if p["shape"] == "HEX8":
        m = max(1, int(round((p["dofs"] / 12.0) ** (1.0 / 3.0))))
        cell = HEX8
    else:
        m = max(1, int(round((p["dofs"] / 96.0) ** (1.0 / 3.0))))
        cell = TET10
    L = 4 * a
    mesh = Mesh(Block([[0, 0, 0], [L, a, a]], nx=4 * m, ny=m, nz=m, cellshape=cell))
    fx = f"x=={L:g}"
    if p["vm"]:
        fy, H = 250e6, 2e9
        mat = VonMises(E=p["E"], nu=p["nu"], fy=fy, H=H)
        dy = 2.0 * fy * L * L / (3.0 * p["E"] * a)   # tip deflection at first yield (beam theory)
        bcs = [("x==0", SurfaceBC(ux=0, uy=0, uz=0)), (fx, SurfaceBC(uz=-2.0 * dy))]
        nincs = 5
    else:
        mat = LinearElastic(E=p["E"], nu=p["nu"])
        bcs = [("x==0", SurfaceBC(ux=0, uy=0, uz=0)), (fx, SurfaceBC(tz=-1e6))]
        nincs = 1
    model = FEModel(mesh, [("all", MechSolid, mat)])
    ana = MechAnalysis(model)
    addstage(ana, bcs, nincs=nincs)
    return ana, {}


def arrival_times(seed, n, rate):
    rng = np.random.default_rng(seed + 7919)
    return np.cumsum(rng.exponential(1.0 / rate, n))
