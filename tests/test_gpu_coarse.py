"""GPU: two-level preconditioner (Jacobi + rigid-body-mode coarse space).

Pinned independently of the device code:
* the aggregates equal a numpy restatement of the recursive coordinate
  bisection rule (split the longest extent at the median of (coordinate,
  node id));
* Kc^-1 equals the inverse of Phi^T K11 Phi built in numpy from the exported
  K (rotation modes from cross products), to 1e-9 relative;
* solutions match SuperLU to 1e-8 in every precision, with several times
  fewer iterations than Jacobi on a slender cantilever.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu


def rcb(X, target):
    n = len(X)
    idx = np.arange(n)
    agg = np.zeros(n, dtype=np.int32)
    st = [(0, n)]
    k = 0
    while st:
        lo, hi = st.pop()
        if hi - lo <= target:
            agg[idx[lo:hi]] = k
            k += 1
            continue
        sub = idx[lo:hi]
        ax = int(np.argmax(X[sub].max(0) - X[sub].min(0)))
        # lower half: the smallest by (coordinate, node id)
        idx[lo:hi] = sub[np.lexsort((sub, X[sub, ax]))]
        mid = lo + (hi - lo) // 2
        st.append((mid, hi))
        st.append((lo, mid))
    return agg, k


def cantilever(am, shape, n, L=1.0, h=0.1):
    mesh = am.Mesh(am.Block([[0, 0, 0], [L, h, h]], nx=n[0], ny=n[1], nz=n[2], cellshape=shape))
    model = am.FEModel(mesh, [("all", am.MechSolid, am.LinearElastic(E=100e9, nu=0.3))])
    bcs = [("x==0", am.SurfaceBC(ux=0, uy=0, uz=0)), (f"x=={L:g}", am.SurfaceBC(tz=-1e5))]
    model.configure_dofs(bcs)
    U, F = model.bc_vals(bcs)
    return model, bcs, U, F


@pytest.mark.parametrize("shape,n,target", [("HEX8", (12, 3, 3), 32), ("TET10", (8, 2, 2), 64)])
def test_coarse_space_matches_host_galerkin(amaru, shape, n, target):
    model, bcs, U, F = cantilever(amaru, shape, n)
    dev = amaru.DeviceModel(model)
    dev.set_coarse(target_nodes=target, max_aggregates=170)
    dev.assemble()
    info = dev.solve_system(U.copy(), F.copy(), precond="two_level", rtol=1e-10)
    assert info["converged"]
    agg, kinv = dev.coarse_space()
    X = model.coords
    ragg, nagg = rcb(X, max(target, -(-model.nnodes // 170)))
    assert np.array_equal(agg, ragg)
    assert kinv.shape == (6 * nagg, 6 * nagg)
    # Phi on the unknowns (eq order), rotation modes as e_k x (x - c)
    nu = model.nu
    cent = np.array([X[agg == a].mean(0) for a in range(nagg)])
    Phi = np.zeros((nu, 6 * nagg))
    for node in range(model.nnodes):
        a = agg[node]
        for d in range(3):
            q = model.eq[node, d]
            if q < 0 or q >= nu:
                continue
            Phi[q, 6 * a + d] = 1.0
            for k in range(3):
                e = np.zeros(3)
                e[k] = 1.0
                Phi[q, 6 * a + 3 + k] = np.cross(e, X[node] - cent[a])[d]
    K = dev.K_csc().tocsr()[:nu, :nu]
    Kc = Phi.T @ (K @ Phi)
    zero = np.abs(np.diag(Kc)) == 0
    Kc[zero, zero] = 1.0
    ref = np.linalg.inv(0.5 * (Kc + Kc.T))
    assert np.abs(kinv - kinv.T).max() == 0.0
    assert np.abs(kinv - ref).max() <= 1e-9 * np.abs(ref).max()
    dev.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed64acc", "mixed"])
@pytest.mark.parametrize("shape,n", [("HEX8", (20, 2, 2)), ("TET10", (12, 2, 2)), ("HEX20", (10, 2, 2))])
def test_two_level_vs_direct(amaru, shape, n, precision):
    model, bcs, U, F = cantilever(amaru, shape, n)
    dev = amaru.DeviceModel(model)
    dev.set_coarse(target_nodes=32)
    dev.assemble()
    Ug, Fg = U.copy(), F.copy()
    info = dev.solve_system(Ug, Fg, precision=precision, precond="two_level", rtol=1e-12)
    K = dev.K_csc().tocsr()
    nu = model.nu
    x = spla.splu(K[:nu, :nu].tocsc()).solve(F[:nu] - K[:nu, nu:] @ U[nu:])
    err = np.linalg.norm(Ug[:nu] - x) / np.linalg.norm(x)
    assert info["converged"] and err <= 1e-8, (err, info)
    reac = K[nu:, :nu] @ x + K[nu:, nu:] @ U[nu:]
    assert np.abs(Fg[nu:] - reac).max() <= 1e-8 * np.abs(reac).max()
    dev.close()


def test_two_level_cuts_iterations(amaru):
    model, bcs, U, F = cantilever(amaru, "TET10", (30, 3, 3))
    dev = amaru.DeviceModel(model)
    dev.assemble()
    its = {}
    sols = {}
    for pc in ("jacobi", "two_level"):
        Ug, Fg = U.copy(), F.copy()
        info = dev.solve_system(Ug, Fg, precision="fp64", precond=pc, rtol=1e-10)
        its[pc] = info["iterations"]
        sols[pc] = Ug
    assert its["two_level"] * 4 <= its["jacobi"], its
    nu = model.nu
    assert np.linalg.norm(sols["two_level"][:nu] - sols["jacobi"][:nu]) <= 1e-8 * np.linalg.norm(sols["jacobi"][:nu])
    # re-assembly refreshes Kc^-1 (same K: same inverse, same iterates)
    _, k1 = dev.coarse_space()
    dev.assemble()
    Ug = U.copy()
    info = dev.solve_system(Ug, F.copy(), precision="fp64", precond="two_level", rtol=1e-10)
    _, k2 = dev.coarse_space()
    assert np.array_equal(k1, k2) and info["iterations"] == its["two_level"]
    assert np.array_equal(Ug, sols["two_level"])
    dev.close()


def test_coarse_refresh_policies(amaru):
    """A von Mises cantilever over several increments: with the coarse
    operator formed once per DOF map its inverse stays the one of the first
    (elastic) assembly; per increment and per assembly it is formed again from
    the tangent. All three converge to the same displacements (the coarse
    operator only preconditions)."""
    vm = amaru.VonMises(E=200e9, nu=0.3, fy=250e6, H=2e9)
    out = {}
    for policy in ("dof_map", "increment", "assembly"):
        mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 0.2, 0.2]], nx=12, ny=2, nz=2, cellshape="HEX8"))
        model = amaru.FEModel(mesh, [("all", amaru.MechSolid, vm)])
        ana = amaru.MechAnalysis(model)
        amaru.addstage(ana, [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", amaru.SurfaceBC(uz=-0.02))], nincs=4)
        model.configure_dofs(ana.stages[0].bcs)
        ana.dev = amaru.DeviceModel(model)
        ana.dev.set_coarse(target_nodes=16)
        ana.dev.set_coarse_refresh(policy)
        ok, msg = amaru.solve(ana, precision="fp64", precond="two_level", pcg_rtol=1e-12)
        assert ok, msg
        assert (model.state[:, 12] > 0).any()   # it yielded (accumulated plastic strain)
        out[policy] = (model.uvals.copy(), ana.dev.coarse_space()[1].copy())
        ana.dev.close()
    u0 = out["dof_map"][0]
    for policy in ("increment", "assembly"):
        assert np.linalg.norm(out[policy][0] - u0) <= 1e-8 * np.linalg.norm(u0)
    # the plastic tangent changed the coarse operator for the refreshing policies
    assert not np.array_equal(out["dof_map"][1], out["increment"][1])
    assert not np.array_equal(out["dof_map"][1], out["assembly"][1])
    # per DOF map: the inverse is the one of the elastic K of the first assembly
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 0.2, 0.2]], nx=12, ny=2, nz=2, cellshape="HEX8"))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid, vm)])
    bcs = [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", amaru.SurfaceBC(uz=-0.02))]
    model.configure_dofs(bcs)
    U, F = model.bc_vals(bcs)
    dev = amaru.DeviceModel(model)
    dev.set_coarse(target_nodes=16)
    dev.assemble()
    dev.solve_system(U * 0.0, F, precision="fp64", precond="two_level", rtol=1e-10)
    assert np.array_equal(dev.coarse_space()[1], out["dof_map"][1])
    dev.close()


def test_batched_formation_matches_alone(amaru):
    """Six cantilevers with coarse spaces of 1 to 16 Gauss-Jordan block rows
    (30 to 1020 coarse dofs) start solving in one pool at the same step:
    their coarse operators are formed in one batch, and each Kc^-1 and
    displacement field equals the one of solving that model alone."""
    from amaru_amd import pool as P, workloads as W
    params = W.cantilever_params(11, 6)
    targets = [4000, 400, 120, 60, 24, 16]

    def build(i):
        ana = W.cantilever(params[i])
        ana.model.configure_dofs(ana.stages[0].bcs)
        ana.dev = amaru.DeviceModel(ana.model)
        ana.dev.set_coarse(targets[i], 170)
        return ana

    kw = dict(precision="fp64", precond="two_level", pcg_rtol=1e-10)
    alone = []
    for i in range(6):
        ana = build(i)
        ok, msg = amaru.solve(ana, pcg_check_every=16, **kw)
        assert ok, msg
        alone.append((ana.dev.coarse_space()[1], ana.model.uvals.copy()))
        ana.dev.close()
    ncs = [k.shape[0] for k, _ in alone]
    assert min(ncs) <= 64 and max(ncs) > 15 * 64, ncs
    anas = [build(i) for i in range(6)]
    ps = P.PoolSet(1, 0, max_members=6, check_every=16)
    try:
        out = P.run_analyses(anas, pool=ps, **kw)
        assert all(ok for ok, _ in out), out
        assert ps.pools[0].coarse_batches() == (1, 6)
        for ana, (kinv, U) in zip(anas, alone):
            assert np.array_equal(ana.dev.coarse_space()[1], kinv)
            assert np.array_equal(ana.model.uvals, U)
    finally:
        for ana in anas:
            ps.pools[0].forget(ana.dev)
            ana.dev.close()
        ps.close()
